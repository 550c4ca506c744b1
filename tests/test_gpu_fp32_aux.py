"""UPML auxiliaries of the FP32 path.

The fused FP32 kernels run the UPML recurrence only in UPML cells, so their six
auxiliaries (getState slots 1 2 4 5 7 8: TM Jz Dz Mx Bx My By, TE Jx Dx Jy Dy
Mz Bz) are stored only outside the box that holds no UPML cell, in a compact
layout (whole rows below and above the box, the columns left and right of it
for the rows between).  The reference runs the recurrence in every cell
(mpiTM_UPML.c:424-455); there sigma = 0 and the auxiliaries feed nothing back
into the fields, which both forms update alike.  Checked here: the auxiliaries
of the UPML frame against the FP64 oracle, the export / import of the compact
layout (a value set outside the box reads back, the box reads as zero), and a
state read and written back through every slot continuing bit-identically.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

AUX_SLOTS = (1, 2, 4, 5, 7, 8)
TOL = 2e-4  # relative L2 of FP32 against FP64 after a few hundred steps


def _sim(m, solver, steps, T=4):
    fi = m.field_info(4000, 4000, 50, 10, 500, 0, steps)
    return m.Simulator(solver, m.MIE_CYLINDER, fi, precision="fp32", T=T)


def _physical(sim, a):
    return a[sim.off:sim.off + sim.N_PX, sim.off:sim.off + sim.N_PY]


def _frame(sim, width):
    i = np.arange(sim.N_PX)[:, None]
    j = np.arange(sim.N_PY)[None, :]
    return (i < width) | (i >= sim.N_PX - width) | (j < width) | (j >= sim.N_PY - width)


@pytest.mark.parametrize("solver", [4, 5])
def test_fp32_upml_auxiliaries_against_fp64_oracle(solver):
    import mpifdtd_amd as m
    from oracle import pyoracle as po
    steps, pml = 400, 10
    print(f"fp32 auxiliaries, solver {solver}", flush=True)
    sim = _sim(m, solver, steps)
    try:
        sim.calc(steps)
        got = {s: _physical(sim, sim.state(s)) for s in AUX_SLOTS}
    finally:
        sim.finish()
    o = po.Sim(width_nm=4000, height_nm=4000, h_u_nm=50, pml=pml, lambda_nm=500, angle_deg=0, stepNum=steps,
               solver=solver, model=1, ntff_time=0)
    o.step(steps)
    # the cells of the UPML frame (two cells short of its inner edge, where the
    # profile's sigma is certainly nonzero)
    fr = _frame(sim, pml - 2)
    for s in AUX_SLOTS:
        ref = _physical(sim, o.field(s))
        assert rel_l2(got[s][fr], ref[fr]) <= TOL, s


def test_fp32_auxiliary_export_import_compact_layout():
    import mpifdtd_amd as m
    rng = np.random.default_rng(7)
    sim = _sim(m, 4, 40)
    try:
        sim.calc(8)
        for s in AUX_SLOTS:
            x = rng.standard_normal((sim.NPXa, sim.NPYa)) + 1j * rng.standard_normal((sim.NPXa, sim.NPYa))
            sim.set_state(s, x)
            out = _physical(sim, sim.state(s))
            want = _physical(sim, x).astype(np.complex64).astype(np.complex128)
            zero = out == 0
            # outside the box the value set reads back (rounded to FP32) ...
            assert np.array_equal(out[~zero], want[~zero]), s
            # ... the box reads as zero: one rectangle around the centre, clear of the frame
            rows = np.where(zero.any(axis=1))[0]
            cols = np.where(zero.any(axis=0))[0]
            assert rows.size and cols.size, s
            assert zero[rows[0]:rows[-1] + 1, cols[0]:cols[-1] + 1].all(), s
            assert zero.sum() == rows.size * cols.size, s
            assert rows[0] > 0 and cols[0] > 0 and rows[-1] < sim.N_PX - 1 and cols[-1] < sim.N_PY - 1, s
            assert zero[sim.N_PX // 2, sim.N_PY // 2], s
            assert not zero[_frame(sim, 1)].any(), s
    finally:
        sim.finish()


def test_fp32_state_round_trip_continues_bit_identically():
    import mpifdtd_amd as m
    n0, n1 = 24, 40
    print("fp32 state round trip", flush=True)
    sim = _sim(m, 4, n0 + n1)
    try:
        sim.calc(n0 + n1)
        ref = [f.copy() for f in sim.fields()] + [sim.state(s) for s in AUX_SLOTS]
    finally:
        sim.finish()
    sim = _sim(m, 4, n0 + n1)
    try:
        sim.calc(n0)
        for s in range(9):
            sim.set_state(s, sim.state(s))
        sim.calc(n1)
        got = [f.copy() for f in sim.fields()] + [sim.state(s) for s in AUX_SLOTS]
    finally:
        sim.finish()
    for k, (g, r) in enumerate(zip(got, ref)):
        assert np.array_equal(g, r), k
