"""Pins the CPU oracle against the compiled reference's own outputs.

The fixtures under tests/golden/ were written by tests/golden/make_golden.py,
which runs the reference's C sources (mpiTM_UPML.c, mpiTE_UPML.c, fdtdTM_upml.c,
fdtdTE_upml.c, the split-field fdtdTM.c and fdtdTE.c, nsFdtdTM.c, nsFdtdTE.c, field.c, ntffTM.c, ntffTE.c, cfft.c, ntff.c, models.c and the
model files) as one MPI rank.  The oracle restates the same arithmetic in the
same order, so most comparisons are bit-exact.
"""
import numpy as np
import pytest

from conftest import load_golden, rel_l2, swept_params
from oracle import pyoracle as po

RUN_CASES = ["mpitm_mie_a0", "mpitm_mie_a30_rect", "mpite_mie_a0", "mpite_mie_a60_rect",
             "tm_mie_pulse", "te_mie_pulse", "mpitm_morpho", "mpite_zigzag",
             "tm2d_mie_a0", "tm2d_zigzag_a30_rect", "te2d_mie_a0", "te2d_layer_a45_rect",
             "nstm_mie_a0", "nstm_zigzag_a30_rect", "nste_mie_a0", "nste_layer_a45_rect"]


def oracle_for(g):
    solver, model, sweep, w, h_nm, hu, pml, lam, ang, steps = [int(v) for v in g["cfg"]]
    kw = swept_params(model, sweep)
    return po.Sim(width_nm=w, height_nm=h_nm, h_u_nm=hu, pml=pml, lambda_nm=lam, angle_deg=ang,
                  stepNum=steps, solver=solver, model=model, **kw), steps


@pytest.mark.parametrize("name", RUN_CASES)
def test_fields_eps_farfield_match_reference(name):
    g = load_golden(name)
    sim, steps = oracle_for(g)
    sim.step(steps)
    tm = sim.cfg.solver in (po.MPI_TM_UPML, po.TM_UPML, po.TM_2D, po.NS_TM_2D)
    which = (0, 3, 6)  # Ez,Hx,Hy / Ex,Ey,Hz
    for k, w in enumerate(which):
        ref = g[f"f{k}"]
        mine = sim.field(w)
        assert mine.shape == ref.shape
        assert np.array_equal(mine, ref), (name, k, rel_l2(mine, ref))
    eps = sim.eps(0 if tm else 1)
    if sim.off:
        eps[0, :] = eps[-1, :] = 0
        eps[:, 0] = eps[:, -1] = 0
    if sim.cfg.solver == po.TE_UPML:
        eps = sim.eps(0)   # fdtdTE_upml_getEps returns EPS_EX
    assert np.array_equal(eps, g["eps"])
    if "ff" in g:
        ff = sim.far_field()
        assert np.array_equal(ff, g["ff"]), rel_l2(ff, g["ff"])
    if "spectra" in g:
        sp = sim.time_spectra()[::4]
        assert rel_l2(sp, g["spectra"]) < 1e-12
    if "Eth" in g:
        # the reference prints these with "%.20lf": absolute resolution 1e-20
        eth, eph = sim.time_translate()
        rows = list(range(0, 360, 15))
        for nm, arr in (("Eth", eth), ("Eph", eph)):
            ref = g[nm]
            mine = arr[rows, :steps]
            assert np.max(np.abs(mine - ref)) < 1e-19 + 1e-12 * np.max(np.abs(ref)), nm
    if "elliptic" in g:
        # field_outputElliptic prints "%lf": absolute resolution 5e-7
        ell = sim.elliptic()
        assert np.abs(g["elliptic"]).max() > 1e-3
        assert np.max(np.abs(ell - g["elliptic"])) <= 5.0000001e-7, np.max(np.abs(ell - g["elliptic"]))


def test_plain_solver_layout_is_unpadded():
    g = load_golden("tm_mie_pulse")
    sim, _ = oracle_for(g)
    assert (sim.NPXa, sim.NPYa) == (sim.N_PX, sim.N_PY) == g["f0"].shape


EPS_CASES = ["eps_circle_h20", "eps_circle_h7", "eps_layer_s0", "eps_layer_s7",
             "eps_zigzag_s0", "eps_zigzag_s4", "eps_morpho_s0", "eps_morpho_s9"]


@pytest.mark.parametrize("name", EPS_CASES)
def test_model_eps_maps_match_reference(name):
    g = load_golden(name)
    model, sweep, w, h_nm, hu = [int(v) for v in g["cfg"]]
    sim = po.Sim(width_nm=w, height_nm=h_nm, h_u_nm=hu, pml=10, lambda_nm=500, stepNum=10,
                 solver=po.TM_UPML, model=model, **swept_params(model, sweep))
    assert np.array_equal(sim.model_eps_map(po.D_XY), g["eps_xy"])
    assert np.array_equal(sim.model_eps_map(po.D_Y, 0.5, 0.0), g["eps_dy"])
    assert np.array_equal(sim.model_eps_map(po.D_X, 0.0, 0.5), g["eps_dx"])
    # the map is not trivially vacuum
    assert g["eps_xy"].max() > 1.0
