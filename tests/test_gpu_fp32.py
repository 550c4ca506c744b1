"""FP32 fused, temporally blocked path against the FP64 oracle.

The FP32 path stores fields and coefficients in single precision; its error
against the FP64 oracle is bounded here (fields and far field), and its
temporal blocking is checked to be exact: the per-cell arithmetic does not
depend on how many steps one launch advances, so any block depth T gives
bit-identical results for the same kernel variant.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

FIELD_TOL = 2e-4   # relative L2 of FP32 fields after a few hundred steps
FF_TOL = 1e-4      # relative L2 of the far field


def _mods():
    import mpifdtd_amd as m
    from mpifdtd_amd import _native as N
    from oracle import pyoracle as po
    return m, N, po


def run_fp32(m, N, solver, model, w, h, hu, pml, lam, ang, steps, T=4, R=16, ntff_time=False, params=None):
    sim = m.Simulator(solver, model, m.field_info(w, h, hu, pml, lam, ang, steps),
                      params=m.model_params(**(params or {})), ntff_time=ntff_time, precision="fp32", T=T, R=R)
    assert N.lib().fdtd_solver_getPrecision() == 0  # the setting for the next init is restored
    assert sim.blocking() == (T, R)  # the solver runs the fused FP32 kernels
    sim.calc(steps)
    return sim


CASES = [
    (4, 1, 4000, 4000, 50, 10, 500, 0, 400),     # MPI TM Mie
    (5, 1, 4000, 3600, 50, 10, 550, 30, 400),    # MPI TE Mie, oblique, non-square
    (2, 1, 4000, 4000, 50, 10, 500, 0, 700),     # plain TM pulse
    (3, 1, 4000, 4000, 50, 10, 500, 0, 700),     # plain TE pulse
    (4, 1, 6000, 6000, 20, 16, 500, 0, 500),     # MPI TM, 300x300 with a real interior region
]


@pytest.mark.parametrize("case", CASES)
def test_fp32_against_fp64_oracle(case):
    m, N, po = _mods()
    solver, model, w, h, hu, pml, lam, ang, steps = case
    print(f"fp32 case {case}", flush=True)
    sim = run_fp32(m, N, solver, model, w, h, hu, pml, lam, ang, steps)
    try:
        f = sim.fields()
        ff = sim.far_field()
    finally:
        sim.finish()
    o = po.Sim(width_nm=w, height_nm=h, h_u_nm=hu, pml=pml, lambda_nm=lam, angle_deg=ang, stepNum=steps,
               solver=solver, model=model, ntff_time=0)
    o.step(steps)
    for k, which in enumerate((0, 3, 6)):
        err = rel_l2(f[k], o.field(which))
        assert err <= FIELD_TOL, (case, k, err)
    assert rel_l2(ff, o.far_field()) <= FF_TOL


@pytest.mark.parametrize("solver", [4, 5, 2, 3])
def test_temporal_blocking_is_exact(solver):
    m, N, _ = _mods()
    args = (solver, 1, 6000, 5000, 20, 16, 500, 0, 120)
    ref = None
    for T in (1, 3, 4, 7):
        sim = run_fp32(m, N, *args, T=T, R=16)
        try:
            f = sim.fields()
        finally:
            sim.finish()
        if ref is None:
            ref = f
        else:
            for a, b in zip(f, ref):
                assert np.array_equal(a, b), (solver, T)


def test_rows_per_wave_variants_agree():
    m, N, _ = _mods()
    args = (4, 1, 6000, 6000, 20, 16, 500, 0, 150)
    out = []
    for R in (4, 8, 16):
        sim = run_fp32(m, N, *args, T=4, R=R)
        try:
            out.append(sim.fields())
        finally:
            sim.finish()
    for f in out[1:]:
        for a, b in zip(f, out[0]):
            assert rel_l2(a, b) <= 1e-6


@pytest.mark.parametrize("solver,RR", [(2, 20), (3, 20), (4, 20), (5, 20), (2, 18), (3, 18), (4, 18), (5, 18)])
def test_rows_per_wave_18_20_in_the_eight_wavefront_shapes(solver, RR):
    """R = 18 / 20: interior tiles of 72 / 80 rows as 8 wavefronts of 9 / 10 rows,
    in the four shapes that have them (split / two streams, with 8- or
    16-wavefront UPML tiles): bit-identical among themselves, within 1e-6 of
    R = 16 and within the FP32 tolerance of the FP64 oracle; another shape
    refuses them."""
    m, N, po = _mods()
    args = (solver, 1, 6000, 6000, 20, 16, 500, 0, 120)
    print(f"R={RR}, solver {solver}", flush=True)
    out = []
    for R, shape in ((RR, 6), (RR, 7), (RR, 8), (RR, 9), (16, 8)):
        sim = m.Simulator(solver, 1, m.field_info(*args[2:]), precision="fp32", T=4, R=R)
        try:
            sim.set_launch_mode(shape)
            sim.calc(args[-1])
            out.append([np.array(f) for f in sim.fields()])
        finally:
            sim.finish()
    for o in out[1:4]:
        for x, y in zip(o, out[0]):
            assert np.array_equal(x, y)
    for x, y in zip(out[4], out[0]):
        assert rel_l2(x, y) <= 1e-6
    o = po.Sim(width_nm=6000, height_nm=6000, h_u_nm=20, pml=16, lambda_nm=500, angle_deg=0, stepNum=args[-1],
               solver=solver, model=1, ntff_time=0)
    o.step(args[-1])
    for k, which in enumerate((0, 3, 6)):
        assert rel_l2(out[0][k], o.field(which)) <= FIELD_TOL, which
    # split with 4-wavefront interior tiles has no R = 20 body: the launch fails
    # when the queued steps go out, here at the state read (4 steps queued of a
    # block of 8, so the step calls themselves launch nothing)
    sim = m.Simulator(solver, 1, m.field_info(*args[2:]), precision="fp32", T=8, R=RR)
    try:
        sim.set_launch_mode(2)
        sim.calc(4)
        with pytest.raises(RuntimeError):
            sim.state(0)
    finally:
        sim.finish()


@pytest.mark.parametrize("solver,RR", [(2, 18), (2, 20), (4, 18), (4, 20), (3, 20), (5, 18)])
def test_rows_per_wave_18_20_on_an_odd_grid(solver, RR):
    """72- and 80-row interior tiles on a 573 x 335 grid (no side a multiple of a
    tile; oblique incidence): within 1e-6 of R = 16 and within the FP32
    tolerance of the FP64 oracle."""
    m, N, po = _mods()
    w, h, hu, pml, lam, ang, steps = 5730, 3350, 10, 12, 500, 30, 160
    print(f"odd grid R={RR}, solver {solver}", flush=True)
    out = []
    for R in (RR, 16):
        sim = m.Simulator(solver, 1, m.field_info(w, h, hu, pml, lam, ang, steps), precision="fp32", T=8, R=R)
        try:
            sim.set_launch_mode(8)
            sim.calc(steps)
            out.append([np.array(f) for f in sim.fields()])
        finally:
            sim.finish()
    for x, y in zip(out[0], out[1]):
        assert rel_l2(x, y) <= 1e-6
    o = po.Sim(width_nm=w, height_nm=h, h_u_nm=hu, pml=pml, lambda_nm=lam, angle_deg=ang, stepNum=steps,
               solver=solver, model=1, ntff_time=0)
    o.step(steps)
    for k, which in enumerate((0, 3, 6)):
        assert rel_l2(out[0][k], o.field(which)) <= FIELD_TOL, which


@pytest.mark.parametrize("solver", [2, 3, 4, 5])
def test_fp32_time_domain_ntff_against_oracle(solver):
    """FP32 fields with the per-step time-domain NTFF binned in FP64: the fused
    kernel samples the contour after every inner step, so temporal blocking
    keeps the per-step NTFF, and its accumulators do not depend on T."""
    m, N, po = _mods()
    args = (solver, 1, 4000, 4000, 50, 10, 500, 0, 700)
    out = {}
    for T in (1, 4, 7):
        sim = run_fp32(m, N, *args, T=T, ntff_time=True)
        try:
            out[T] = [sim.ntff_time(a) for a in range(3)]
            if solver in (2, 3):
                out[T].append(sim.time_spectra())
        finally:
            sim.finish()
    # the fields are bit-identical for every T; the bins are summed with atomics,
    # whose order varies, so the accumulators agree to rounding
    for T in (4, 7):
        for x, y in zip(out[T], out[1]):
            assert rel_l2(x, y) <= 1e-12, (T, rel_l2(x, y))
    o = po.Sim(width_nm=4000, height_nm=4000, h_u_nm=50, pml=10, lambda_nm=500, stepNum=700, solver=solver, model=1)
    o.step(700)
    for a in range(3):
        assert rel_l2(out[4][a], o.ntff_time(a)) <= 2e-4, a
    if solver in (2, 3):
        assert rel_l2(out[4][3], o.time_spectra()) <= 1e-4


def test_blocking_change_mid_run():
    """setBlocking on a live FP32 solver takes effect and keeps the trajectory
    exact: the depth only changes how steps are grouped into launches."""
    m, N, _ = _mods()
    args = (4, 1, 6000, 5000, 20, 16, 500, 0, 130)
    sim = run_fp32(m, N, *args[:-1], 0, T=4, R=16)
    try:
        assert sim.blocking() == (4, 16)
        sim.calc(37)
        sim.set_blocking(7, 16)
        assert sim.blocking() == (7, 16)
        sim.calc(93)
        f = sim.fields()
    finally:
        sim.finish()
    ref = run_fp32(m, N, *args, T=4, R=16)
    try:
        g = ref.fields()
    finally:
        ref.finish()
    for a, b in zip(f, g):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("solver", [2, 3])
def test_fp32_thin_upml_plain(solver):
    """Plain solvers with a one-cell UPML: the interior tiles reach the fixed
    edge cells, which the tiling keeps outside their dependency cones."""
    m, N, po = _mods()
    args = (solver, 1, 4000, 4000, 50, 1, 500, 0, 300)
    sim = run_fp32(m, N, *args, T=4, R=8)
    try:
        f = sim.fields()
    finally:
        sim.finish()
    o = po.Sim(width_nm=4000, height_nm=4000, h_u_nm=50, pml=1, lambda_nm=500, angle_deg=0, stepNum=300,
               solver=solver, model=1, ntff_time=0)
    o.step(300)
    for k, which in enumerate((0, 3, 6)):
        assert rel_l2(f[k], o.field(which)) <= FIELD_TOL, (solver, k)


def test_fp32_large_grid_split_launch_against_fp64():
    """A grid with thousands of interior tiles takes the two-launch path (interior
    tiles, then UPML tiles); its fields agree with the FP64 reference-order path."""
    m, N, _ = _mods()
    args = (4, 1, 8000, 8000, 5, 16, 550, 0, 200)   # 1600 x 1600 cells, MPI TM Mie
    sim = run_fp32(m, N, *args, T=4, R=4, params={"circle_radius_nm": 500, "circle_n": 1.6})
    try:
        f = sim.fields()
    finally:
        sim.finish()
    ref = m.Simulator(4, 1, m.field_info(8000, 8000, 5, 16, 550, 0, 200),
                      params=m.model_params(circle_radius_nm=500, circle_n=1.6), ntff_time=False)
    try:
        ref.calc(200)
        g = ref.fields()
    finally:
        ref.finish()
    for k in range(3):
        assert rel_l2(f[k], g[k]) <= FIELD_TOL, k


@pytest.mark.parametrize("solver,R", [(4, 16), (3, 16), (5, 8)])
def test_launch_shapes_are_bit_identical(solver, R):
    """Merged, split and two-stream launches of the UPML and interior tiles, with
    the UPML tiles as 4 wavefronts of 8 rows, 8 of 4 or 16 of 2, and the interior tiles as
    4 wavefronts of R rows or 8 of R/2, run the same per-cell code: the fields
    agree bit for bit."""
    m, N, po = _mods()
    out = []
    for shape in (0, 1, 2, 3, 4, 5, 6, 7, 8, 9):
        sim = m.Simulator(solver, 1, m.field_info(6000, 6000, 20, 16, 500, 0, 120), precision="fp32", T=4, R=R)
        try:
            sim.set_launch_mode(shape)
            sim.calc(120)
            out.append([np.array(f) for f in sim.fields()])
        finally:
            sim.finish()
    for o in out[1:]:
        for x, y in zip(o, out[0]):
            assert np.array_equal(x, y)
