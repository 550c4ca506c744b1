"""fdtd_solver_farFieldNoAlloc, the allocation-free variant of
fdtd_solver_farField, through the C ABI: the same far field (FP64 sums in
another order: <= 1e-12 relative L2, well inside the 1e-10 of the FP64 path)
for FP64 and FP32 solvers, TM and TE, on two calls in a row, the second on a
smaller grid (the shared scratch keeps the first call's contents); a contour
beyond the scratch's limit is refused, and queued FP32 steps are refused."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

CASES = [  # solver, precision, width_nm, height_nm, h, steps
    (4, "fp64", 4000, 4000, 25, 500),
    (5, "fp64", 6000, 4000, 25, 500),
    (2, "fp32", 4000, 5000, 25, 400),
    (5, "fp32", 5000, 5000, 25, 400),
]


def _run(m, solver, precision, w, h, hu, steps):
    sim = m.Simulator(solver, m.MIE_CYLINDER, m.field_info(w, h, hu, 10, 500, 30, steps), ntff_time=False,
                      precision=precision, T=4)
    sim.calc(steps)
    return sim


@pytest.mark.parametrize("case", range(len(CASES)))
def test_noalloc_far_field_equals_far_field(case):
    import mpifdtd_amd as m
    solver, prec, w, h, hu, steps = CASES[case]
    print("no-alloc far field:", CASES[case], flush=True)
    # first call on the full grid, second on a grid of half the size
    for scale in (1, 2):
        sim = _run(m, solver, prec, w // scale, h // scale, hu, steps)
        try:
            ref = sim.far_field()
            got = sim.far_field_noalloc()
            again = sim.far_field_noalloc()
        finally:
            sim.finish()
        assert np.linalg.norm(ref) > 0
        assert rel_l2(got, ref) <= 1e-12, (scale, rel_l2(got, ref))
        assert np.array_equal(got, again)


def test_noalloc_far_field_limits():
    import mpifdtd_amd as m
    # a contour of 2 * 66540 + 2 * 10 cells: beyond the 65536 the scratch holds
    sim = m.Simulator(m.MPI_TM_UPML_2D, m.NO_MODEL, m.field_info(666000, 400, 10, 10, 500, 0, 10),
                      ntff_time=False)
    try:
        with pytest.raises(RuntimeError, match="exceeds the limit"):
            sim.far_field_noalloc()
    finally:
        sim.finish()
    sim = _run(m, 4, "fp32", 4000, 4000, 25, 400)  # T = 4: whole blocks, nothing queued
    try:
        sim.calc(2)  # two steps queued
        with pytest.raises(RuntimeError, match="queued"):
            sim.far_field_noalloc()
        ref = sim.far_field()  # launches them
        assert rel_l2(sim.far_field_noalloc(), ref) <= 1e-12
    finally:
        sim.finish()
