"""The FP32 fused path at the full sizes of the benchmark configurations, checked
through properties that do not depend on a CPU run of that size:

- the 1024 x 1024 Mie cylinder grid (1056 x 1056 with its 16-cell UPML), TM and
  TE, in the launch shapes the autotuner picks for it: fields and far field
  against the FP64 reference-order GPU path (itself pinned to the CPU oracle at
  1e-10 on smaller grids) after 1200 steps, by when the scattered wave has
  crossed the ~410 cells from the cylinder to the NTFF contour (the far field
  of the reference run is checked to be non-zero);
- the 8192 x 4096 Morpho-scale grid (8256 x 4160 with its 32-cell UPML): the
  launch shapes agree bit for bit at full size (different tile launches, the
  same per-cell arithmetic), and the fields follow the FP64 path.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

FIELD_TOL = 2e-4   # relative L2 of FP32 fields after up to 1200 steps
FF_TOL = 1e-4      # relative L2 of the far field

MIE = dict(params={"circle_radius_nm": 500, "circle_n": 1.6})
C2 = (1, 5280, 5280, 5, 16, 550, 0)          # MIE_CYLINDER, 1056 x 1056 cells
C4 = (3, 41280, 20800, 5, 32, 550, 90)       # MORPHO_SCALE, 8256 x 4160 cells


def _sim(m, solver, cfg, steps, precision, T=8, R=16, params=None):
    model, w, h, hu, pml, lam, ang = cfg
    return m.Simulator(solver, model, m.field_info(w, h, hu, pml, lam, ang, steps),
                       params=m.model_params(**(params or {})), ntff_time=False, precision=precision,
                       **({"T": T, "R": R} if precision == "fp32" else {}))


def _run(m, solver, cfg, steps, precision, shape=None, T=8, R=16, params=None, far=True):
    sim = _sim(m, solver, cfg, steps, precision, T, R, params)
    try:
        if shape is not None:
            sim.set_launch_mode(shape)
        sim.calc(steps)
        return [np.array(f) for f in sim.fields()], (sim.far_field() if far else None)
    finally:
        sim.finish()


@pytest.mark.parametrize("solver", [4, 5])
def test_mie1056_fp32_against_fp64(solver):
    import mpifdtd_amd as m
    print("1056 x 1056 Mie cylinder, solver", solver, flush=True)
    steps = 1200
    ref, ffr = _run(m, solver, C2, steps, "fp64", **MIE)
    assert np.linalg.norm(ffr) > 0, "the scattered wave has not reached the contour"
    for shape, R in ((6, 8), (4, 16)):
        got, ffg = _run(m, solver, C2, steps, "fp32", shape=shape, R=R, **MIE)
        for k in range(3):
            assert rel_l2(got[k], ref[k]) <= FIELD_TOL, (shape, R, k, rel_l2(got[k], ref[k]))
        assert rel_l2(ffg, ffr) <= FF_TOL, (shape, R, rel_l2(ffg, ffr))


def test_morpho8256_shapes_bit_identical_and_fp64():
    import mpifdtd_amd as m
    print("8256 x 4160 Morpho scale, MPI TM", flush=True)
    steps = 48
    out = {}
    for shape in (2, 4, 5, 6):
        out[shape], _ = _run(m, 4, C4, steps, "fp32", shape=shape, far=False)
    for shape in (4, 5, 6):
        for k in range(3):
            assert np.array_equal(out[shape][k], out[2][k]), (shape, k)
    ref, _ = _run(m, 4, C4, steps, "fp64", far=False)
    for k in range(3):
        assert np.abs(ref[k]).max() > 0
        assert rel_l2(out[4][k], ref[k]) <= FIELD_TOL, (k, rel_l2(out[4][k], ref[k]))
