"""The benchmark configurations as stated, on the GPU.

- Config 1: NoModel, 256 x 256 cells + a 10-cell UPML (276 x 276) at h = 5 nm,
  the soft point source at the centre, lambda = 500 nm, Courant number 0.5
  (field_setLightSpeed), FP64, 2000 steps: Ez, Hx, Hy against the CPU oracle run
  with the same Courant number, <= 1e-10.
- Config 2 with the NTFF contour 10 cells inside the UPML (field_setNTFFInset):
  the FP64 far field against the oracle with the same inset on a 264 x 264 twin
  of the cylinder, <= 1e-10; and on the full 1056 x 1056 grid the FP32 far field
  against FP64 after 1200 steps (the scattered wave has reached the contour by
  then: the check asserts a non-zero reference), <= 1e-4.
- Config 3: the seeded 64-instance sweep (sides 512 ... 2048 + 16-cell UPML,
  6000 ... 10000 steps) advanced as one batch; two instances again through the
  FP64 single-instance path: far field <= 1e-4.
- Config 4 at its full size (8256 x 4160, 32-cell UPML): TE in every launch
  shape bit for bit and against FP64; the 81-wavelength running DFT of the
  plain pulse solver, TM and TE, FP32 against FP64 after 4000 steps, <= 1e-4.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-10
FF_TOL = 1e-4
FIELD_TOL = 2e-4
C2 = (1, 5280, 5280, 5, 16, 550, 0)     # MIE_CYLINDER, 1056 x 1056 cells
C4 = (3, 41280, 20800, 5, 32, 550, 90)  # MORPHO_SCALE, 8256 x 4160 cells
MIE = {"circle_radius_nm": 500, "circle_n": 1.6}


def _sim(m, solver, cfg, steps, precision="fp64", T=8, R=16, params=None, **kw):
    model, w, h, hu, pml, lam, ang = cfg
    return m.Simulator(solver, model, m.field_info(w, h, hu, pml, lam, ang, steps),
                       params=m.model_params(**(params or {})), precision=precision,
                       **({"T": T, "R": R} if precision == "fp32" else {}), **kw)


def test_config1_point_source_courant_half():
    import mpifdtd_amd as m
    from oracle import pyoracle as po
    print("config 1: 276 x 276 NoModel point source, c0 = 0.5, 2000 steps", flush=True)
    steps = 2000
    sim = _sim(m, m.MPI_TM_UPML_2D, (m.NO_MODEL, 1380, 1380, 5, 10, 500, 0), steps, c0=0.5, point_source=True,
               ntff_time=False)
    try:
        assert sim.N_PX == 276 and sim.N_PY == 276
        sim.calc(steps)
        f = sim.fields()
    finally:
        sim.finish()
    o = po.Sim(width_nm=1380, height_nm=1380, h_u_nm=5, pml=10, lambda_nm=500, angle_deg=0, stepNum=steps,
               solver=po.MPI_TM_UPML, model=po.NO_MODEL, point_source=1, ntff_time=0, c0=0.5)
    o.step(steps)
    for k, which in enumerate((0, 3, 6)):
        assert rel_l2(f[k], o.field(which)) <= TOL, (k, rel_l2(f[k], o.field(which)))
    assert np.abs(f[0]).max() > 1e-3


@pytest.mark.parametrize("solver", [4, 5])
def test_config2_ntff_inset_10_against_oracle(solver):
    """The contour 10 cells inside the UPML on a 264 x 264 twin of the cylinder
    (h = 20 nm, r = 500 nm, n = 1.6, 16-cell UPML), 900 steps."""
    import mpifdtd_amd as m
    from oracle import pyoracle as po
    print("config 2 twin, inset 10, solver", solver, flush=True)
    steps = 900
    sim = _sim(m, solver, (1, 5280, 5280, 20, 16, 550, 0), steps, params=MIE, ntff_inset=10, ntff_time=False)
    try:
        sim.calc(steps)
        ff = sim.far_field()
        f0 = sim.fields()[0]
    finally:
        sim.finish()
    o = po.Sim(width_nm=5280, height_nm=5280, h_u_nm=20, pml=16, lambda_nm=550, angle_deg=0, stepNum=steps,
               solver=solver, model=po.MIE_CYLINDER, ntff_time=0, ntff_inset=10, **MIE)
    o.step(steps)
    ni = o.ntff_info()
    assert ni["bottom"] == 16 + 10 and ni["left"] == 16 + 10
    assert rel_l2(f0, o.field(0)) <= TOL
    assert rel_l2(ff, o.far_field()) <= TOL
    # the inset moves the contour: the far field differs from the default contour's
    o5 = po.Sim(width_nm=5280, height_nm=5280, h_u_nm=20, pml=16, lambda_nm=550, angle_deg=0, stepNum=steps,
                solver=solver, model=po.MIE_CYLINDER, ntff_time=0, **MIE)
    o5.step(steps)
    assert rel_l2(o5.far_field(), ff) > 1e-6


@pytest.mark.parametrize("solver", [4, 5])
def test_config2_full_size_fp32_far_field_inset_10(solver):
    import mpifdtd_amd as m
    print("config 2 full size, inset 10, 1200 steps, solver", solver, flush=True)
    steps = 1200
    out = {}
    for prec in ("fp64", "fp32"):
        sim = _sim(m, solver, C2, steps, prec, T=8, R=8, params=MIE, ntff_inset=10, ntff_time=False)
        try:
            if prec == "fp32":
                sim.set_launch_mode(6)
            sim.calc(steps)
            out[prec] = sim.far_field()
        finally:
            sim.finish()
    assert np.linalg.norm(out["fp64"]) > 0  # the scattered wave has reached the contour
    assert rel_l2(out["fp32"], out["fp64"]) <= FF_TOL, rel_l2(out["fp32"], out["fp64"])


def test_config3_batch_of_64_against_fp64():
    import mpifdtd_amd as m
    from mpifdtd_amd import batch as B
    spec = B.sweep_spec(1, 64)
    print("config 3: batch of", len(spec), "seeded instances", flush=True)
    with B.Batch(B.sweep_instances(spec), T=8, R=16) as b:
        b.run()
        done = b.sync()
        assert done == [s["steps"] for s in spec]
        # the first two and the largest instance
        big = max(range(len(spec)), key=lambda i: spec[i]["side"])
        pick = [0, 1, big]
        got = {i: b.far_field(i) for i in pick}
    for i in pick:
        s = spec[i]
        solver = m.MPI_TM_UPML_2D if s["mode"] == "tm" else m.MPI_TE_UPML_2D
        model = m.LAYER if s["kind"] == "multilayer" else m.ZIGZAG
        ref = m.Simulator(solver, model, m.field_info(s["width_nm"], s["height_nm"], s["h_nm"], s["pml"],
                                                      s["lambda_nm"], 0, s["steps"]),
                          params=m.model_params(**s["params"]), ntff_time=False)
        try:
            ref.calc(s["steps"])
            ff64 = ref.far_field()
        finally:
            ref.finish()
        assert np.linalg.norm(ff64) > 0
        assert rel_l2(got[i], ff64) <= FF_TOL, (i, s, rel_l2(got[i], ff64))


def test_config4_te_full_size_shapes_and_fp64():
    import mpifdtd_amd as m
    print("config 4 full size, MPI TE, 48 steps", flush=True)
    steps = 48
    out = {}
    for shape in (2, 4, 6, 7):
        sim = _sim(m, m.MPI_TE_UPML_2D, C4, steps, "fp32", T=8, R=16, ntff_time=False)
        try:
            sim.set_launch_mode(shape)
            sim.calc(steps)
            out[shape] = [np.array(f) for f in sim.fields()]
        finally:
            sim.finish()
    for shape in (4, 6, 7):
        for k in range(3):
            assert np.array_equal(out[shape][k], out[2][k]), (shape, k)
    sim = _sim(m, m.MPI_TE_UPML_2D, C4, steps, "fp64", ntff_time=False)
    try:
        sim.calc(steps)
        ref = sim.fields()
    finally:
        sim.finish()
    for k in range(3):
        assert rel_l2(out[2][k], ref[k]) <= FIELD_TOL, (k, rel_l2(out[2][k], ref[k]))


@pytest.mark.parametrize("solver", [2, 3])
def test_config4_sweep81_dft_full_size(solver):
    """The plain pulse solver on the 8256 x 4160 Morpho grid with the running DFT at
    the 81 wavelengths 380-780 nm, 4000 steps: FP32 against FP64 (accumulators and
    far-field spectra; the scattered pulse has reached the contour's near sides)."""
    import mpifdtd_amd as m
    print("config 4 sweep DFT, 4000 steps, solver", solver, flush=True)
    lams = np.arange(380.0, 781.0, 5.0)
    steps = 4000
    out = {}
    for prec in ("fp64", "fp32"):
        sim = _sim(m, solver, C4, steps, prec, T=10, R=16, ntff_time=False)
        try:
            if prec == "fp32":
                sim.set_launch_mode(7 if solver == 2 else 6)
            sim.set_dft(lams)
            sim.calc(steps)
            out[prec] = (sim.dft(), sim.dft_spectra())
        finally:
            sim.finish()
    d64, s64 = out["fp64"]
    d32, s32 = out["fp32"]
    assert np.linalg.norm(d64) > 0 and np.all(np.isfinite(d32))
    assert rel_l2(d32, d64) <= FF_TOL, rel_l2(d32, d64)
    assert rel_l2(s32, s64) <= FF_TOL, rel_l2(s32, s64)
