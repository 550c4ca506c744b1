"""GPU parity of the FP64 reference-order path, through the C ABI.

Every case runs the reference's driver flow (models_setModel, simulator_setSolver,
simulator_init, simulator_calc x steps, the solver's getters) on the GPU and
compares with (a) the golden outputs of the compiled reference and (b) the CPU
oracle on the same inputs.  Tolerance: 1e-10 relative L2 (FP64); eps maps are
bit-exact.
"""
import numpy as np
import pytest

from conftest import load_golden, rel_l2, swept_params

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _mods():
    import mpifdtd_amd as m
    from oracle import pyoracle as po
    return m, po


def run_gpu(m, solver, model, w, h, hu, pml, lam, ang, steps, params=None, **kw):
    fi = m.field_info(w, h, hu, pml, lam, ang, steps)
    sim = m.Simulator(solver, model, fi, params=m.model_params(**(params or {})), **kw)
    sim.calc(steps)
    return sim


def run_oracle(po, solver, model, w, h, hu, pml, lam, ang, steps, params=None, **kw):
    o = po.Sim(width_nm=w, height_nm=h, h_u_nm=hu, pml=pml, lambda_nm=lam, angle_deg=ang, stepNum=steps,
               solver=solver, model=model, **(params or {}), **kw)
    o.step(steps)
    return o


RUN_CASES = ["mpitm_mie_a0", "mpitm_mie_a30_rect", "mpite_mie_a0", "mpite_mie_a60_rect",
             "tm_mie_pulse", "te_mie_pulse", "mpitm_morpho", "mpite_zigzag"]


@pytest.mark.parametrize("name", RUN_CASES)
def test_reference_golden_parity(name):
    m, po = _mods()
    g = load_golden(name)
    solver, model, sweep, w, h, hu, pml, lam, ang, steps = [int(v) for v in g["cfg"]]
    params = swept_params(model, sweep)
    print(f"case {name}: solver {solver} model {model} {w}x{h} h={hu} steps={steps}", flush=True)
    sim = run_gpu(m, solver, model, w, h, hu, pml, lam, ang, steps, params)
    try:
        f = sim.fields()
        for k in range(3):
            assert f[k].shape == g[f"f{k}"].shape
            err = rel_l2(f[k], g[f"f{k}"])
            assert err <= TOL, (name, k, err)
        eps = sim.eps()
        if sim.off:
            eps[0, :] = eps[-1, :] = 0
            eps[:, 0] = eps[:, -1] = 0
        assert np.array_equal(eps, g["eps"]), name
        if "ff" in g:
            assert rel_l2(sim.far_field(), g["ff"]) <= TOL
        if "spectra" in g:
            assert rel_l2(sim.time_spectra()[::4], g["spectra"]) <= TOL
        # per-step time-domain NTFF accumulators against the oracle
        o = run_oracle(po, solver, model, w, h, hu, pml, lam, ang, steps, params)
        for a in range(3):
            ref = o.ntff_time(a)
            assert rel_l2(sim.ntff_time(a), ref) <= TOL, (name, a)
        for which in range(9):  # full state including the UPML auxiliaries
            assert rel_l2(sim.state(which), o.field(which)) <= TOL, (name, which)
    finally:
        sim.finish()


def test_point_source_nomodel_tm():
    """Config-1 shape (NoModel, soft point source at the centre), small grid, FP64."""
    m, po = _mods()
    args = (po.MPI_TM_UPML, po.NO_MODEL, 1380, 1380, 5, 10, 500, 0, 400)
    sim = run_gpu(m, *args, point_source=True)
    try:
        o = run_oracle(po, *args, point_source=1)
        f = sim.fields()
        for k, w in enumerate((0, 3, 6)):
            assert rel_l2(f[k], o.field(w)) <= TOL
        assert np.abs(f[0]).max() > 1e-3
        assert rel_l2(sim.far_field(), o.far_field()) <= TOL
    finally:
        sim.finish()


def test_nomodel_without_source_stays_zero():
    """With eps = 1 everywhere the reference's sources inject exactly zero (Q2)."""
    m, _ = _mods()
    sim = run_gpu(m, m.MPI_TM_UPML_2D, m.NO_MODEL, 1000, 1000, 10, 10, 500, 0, 50)
    try:
        for f in sim.fields():
            assert not np.any(f)
    finally:
        sim.finish()


@pytest.mark.parametrize("solver", [4, 5, 2, 3])
def test_far_field_against_oracle(solver):
    m, po = _mods()
    args = (solver, po.MIE_CYLINDER, 4000, 4000, 50, 10, 500, 0, 700)
    sim = run_gpu(m, *args)
    try:
        o = run_oracle(po, *args)
        assert rel_l2(sim.far_field(), o.far_field()) <= TOL
    finally:
        sim.finish()


@pytest.mark.parametrize("model,sweep", [(1, 0), (2, 7), (5, 4), (3, 9)])
@pytest.mark.parametrize("solver", [4, 5, 2, 3])
def test_device_eps_fill_bit_exact(model, sweep, solver):
    """The GPU eps fill (shared host/device model code) reproduces the oracle's maps."""
    m, po = _mods()
    params = swept_params(model, sweep)
    w, h = {1: (2920, 2920), 2: (900, 1500), 5: (800, 2200), 3: (900, 1900)}[model]
    hu = 20 if model == 1 else 10
    sim = run_gpu(m, solver, model, w, h, hu, 10, 500, 0, 1, params, ntff_time=False)
    try:
        o = po.Sim(width_nm=w, height_nm=h, h_u_nm=hu, pml=10, lambda_nm=500, stepNum=1, solver=solver,
                   model=model, **params)
        which = 1 if solver == po.MPI_TE_UPML else 0
        got, want = sim.eps(), o.eps(which)
        if sim.off:
            want = want.copy()
            want[0, :] = want[-1, :] = 0
            want[:, 0] = want[:, -1] = 0
        assert np.array_equal(got, want)
    finally:
        sim.finish()


def test_reset_reruns_identically():
    m, _ = _mods()
    sim = run_gpu(m, m.MPI_TE_UPML_2D, m.MIE_CYLINDER, 4000, 4000, 50, 10, 500, 0, 150)
    try:
        a = sim.fields()
        sim.reset()
        sim.calc(150)
        b = sim.fields()
        for x, y in zip(a, b):
            assert rel_l2(x, y) <= 1e-14
    finally:
        sim.finish()


def test_angle_change_after_reset_matches_oracle():
    """simulator_reset + field_setWaveAngle, the reference's sweep step (main.c:207-210)."""
    m, po = _mods()
    sim = run_gpu(m, m.MPI_TM_UPML_2D, m.MIE_CYLINDER, 4000, 4000, 50, 10, 500, 0, 100)
    try:
        sim.reset(angle_deg=45)
        sim.calc(120)
        o = run_oracle(po, po.MPI_TM_UPML, po.MIE_CYLINDER, 4000, 4000, 50, 10, 500, 45, 120)
        assert rel_l2(sim.fields()[0], o.field(0)) <= TOL
    finally:
        sim.finish()


@pytest.mark.parametrize("solver,pol", [(4, "tm"), (5, "te")])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_far_field_follows_mie_series_at_100_cells_per_wavelength(solver, pol, precision):
    """The far fields against the Mie series of the dielectric cylinder at h = 5 nm
    (100 cells per wavelength): the pattern error falls to a few per cent, as the
    staircasing of the cylinder shrinks (about 10 % at 25 cells per wavelength)."""
    import math
    from mie import compare
    m, _ = _mods()
    print(f"mie {pol} {precision}", flush=True)
    kw = {"ntff_time": False}
    if precision == "fp32":
        kw.update(precision="fp32", T=8, R=16)
    sim = run_gpu(m, solver, 1, 4000, 4000, 5, 10, 500, 0, 5000, {"circle_radius_nm": 500, "circle_n": 1.6}, **kw)
    try:
        ff = sim.far_field()
    finally:
        sim.finish()
    err, other, asym = compare(ff, 2 * math.pi, 1.6, pol)
    assert err < 0.05, err
    assert other > 0.3, other
    assert asym < 0.01, asym


@pytest.mark.parametrize("name", ["tm2d_mie_a0", "tm2d_zigzag_a30_rect", "te2d_mie_a0", "te2d_layer_a45_rect",
                                  "nstm_mie_a0", "nstm_zigzag_a30_rect", "nste_mie_a0", "nste_layer_a45_rect"])
def test_split_field_parity(name, tmp_path, monkeypatch):
    """The split-field PML solvers (fdtdTM.c / fdtdTE.c, solvers TM_2D / TE_2D) and
    the non-standard FDTD solvers (nsFdtdTM.c / nsFdtdTE.c, NS_TM_2D / NS_TE_2D): the three field getters,
    eps and the far field against the compiled reference's golden run, every field
    and split component against the oracle, and the tm_ / te_ / ns_tm_ / ns_te_<h>nm.txt file
    that finish() writes through reset() against the reference's (%lf resolution)."""
    m, po = _mods()
    g = load_golden(name)
    solver, model, sweep, w, h, hu, pml, lam, ang, steps = [int(v) for v in g["cfg"]]
    params = swept_params(model, sweep)
    tm = solver in (po.TM_2D, po.NS_TM_2D)
    prefix = {po.TM_2D: "tm", po.TE_2D: "te", po.NS_TM_2D: "ns_tm", po.NS_TE_2D: "ns_te"}[solver]
    print(f"case {name}: {prefix} model {model} {w}x{h} h={hu} steps={steps}", flush=True)
    monkeypatch.chdir(tmp_path)
    sim = run_gpu(m, solver, model, w, h, hu, pml, lam, ang, steps, params, file_output=True)
    o = run_oracle(po, solver, model, w, h, hu, pml, lam, ang, steps, params)
    try:
        f = sim.fields()
        for k in range(3):
            assert f[k].shape == g[f"f{k}"].shape
            assert rel_l2(f[k], g[f"f{k}"]) <= TOL, (name, k, rel_l2(f[k], g[f"f{k}"]))
        assert np.array_equal(sim.eps(), g["eps"]), name
        if solver in (po.NS_TM_2D, po.NS_TE_2D):
            # getEpsX / Y / Z: TM EPS_HX, EPS_HY, EPS_EZ; TE EPS_EX, EPS_EY, EPS_HZ
            import ctypes
            L = m.lib()
            pre, order = ("nsFdtdTM", (1, 2, 0)) if tm else ("nsFdtdTE", (0, 1, 2))
            for ax, which in zip("XYZ", order):
                f = getattr(L, f"{pre}_getEps{ax}")
                f.restype = ctypes.c_void_p
                got = m._native.as_array(f(), (sim.NPXa, sim.NPYa), np.float64)
                assert np.array_equal(got, o.eps(which)), (name, ax)
        if "ff" in g:
            assert rel_l2(sim.far_field(), g["ff"]) <= TOL
        # TM: Ez, Ezx, Ezy, Hx, Hy; TE: Ex, Ey, Hz, Hzx, Hzy
        for which in ((0, 1, 2, 3, 6) if tm else (0, 3, 6, 7, 8)):
            assert np.abs(o.field(which)).max() > 0
            assert rel_l2(sim.state(which), o.field(which)) <= TOL, (name, which)
        ell_oracle = o.elliptic()
    finally:
        sim.finish()
    rows = np.loadtxt(tmp_path / f"{prefix}_{hu}nm.txt")
    assert np.array_equal(rows[:, 0], np.arange(181))
    assert np.max(np.abs(rows[:, 1] - g["elliptic"])) <= 1.0000001e-6
    assert np.max(np.abs(rows[:, 1] - ell_oracle)) <= 5.0000001e-7
