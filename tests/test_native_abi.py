"""CPU checks of the C-ABI boundary: the library loads and exports every
function that include/*.h declares (no GPU call is made)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADERS = [os.path.join(ROOT, "include", f) for f in sorted(os.listdir(os.path.join(ROOT, "include"))) if f.endswith(".h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        src = re.sub(r"//[^\n]*", "", src)
        src = re.sub(r"#[^\n]*", "", src)
        # drop the extern "C" wrapper (its closing brace sits on a line of its own)
        src = src.replace('extern "C" {', "")
        src = re.sub(r"^\s*\}\s*$", "", src, flags=re.M)
        # drop struct / enum bodies
        src = re.sub(r"\{[^{}]*\}", ";", src)
        for stmt in src.split(";"):
            s = " ".join(stmt.split())
            if not s or s.startswith("typedef") or s.startswith("extern \"C\"") or "(" not in s:
                continue
            m = re.search(r"([A-Za-z_][A-Za-z0-9_]*)\s*\([^()]*\)\s*$", s)
            if m and m.group(1) not in ("if", "while", "sizeof"):
                names.add(m.group(1))
    return sorted(names)


def test_headers_declare_the_reference_entry_points():
    names = declared_functions()
    for must in ("mpi_fdtdTM_upml_getUpdate", "mpi_fdtdTE_upml_getEy", "fdtdTM_upml_getEz", "fdtdTE_upml_getHz",
                 "simulator_init", "simulator_calc", "field_init", "field_nextStep", "models_setModel",
                 "models_eps", "ntff_outputEnormBin", "fdtd_hip_grid_step"):
        assert must in names, must
    assert len(names) > 100


def test_library_exports_every_declared_function():
    import mpifdtd_amd
    L = ctypes.CDLL(mpifdtd_amd.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    from mpifdtd_amd import _native
    monkeypatch.setattr(_native, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(_native, "_lib", None)
    with pytest.raises(_native.NativeLibraryMissing):
        _native.lib()


def test_host_field_state_matches_oracle():
    """field_init / field_nextStep / sigma of the C host against the oracle (no GPU needed)."""
    import mpifdtd_amd as m
    from mpifdtd_amd import _native as N
    from oracle import pyoracle as po
    L = N.lib()
    L.field_setLightSpeed(0.7071)
    L.field_setNTFFInset(5)
    fi = m.field_info(5280, 3000, 5, 16, 550, 30, 800)
    L.field_init(fi)
    ref = po.Sim(width_nm=5280, height_nm=3000, h_u_nm=5, pml=16, lambda_nm=550, angle_deg=30, stepNum=800,
                 solver=po.MPI_TM_UPML, model=po.NO_MODEL)
    ni = L.field_getNTFFInfo()
    oi = ref.ntff_info()
    assert (ni.top, ni.bottom, ni.left, ni.right, ni.cx, ni.cy) == tuple(oi[k] for k in ("top", "bottom", "left", "right", "cx", "cy"))
    assert ni.RFperC == oi["RFperC"] and ni.arraySize == ref.arraySize
    assert L.field_getOmega() == ref.omega and L.field_getK() == ref.k
    for x in (0, 3, 15.5, 16, 500, 1040, 1055.5):
        assert L.field_sigmaX(x, 0) == ref.sigma_x(x)
    for y in (0, 7.5, 16, 300, 583, 599.5):
        assert L.field_sigmaY(0, y) == ref.sigma_y(y)
    for _ in range(37):
        L.field_nextStep()
    ref.step(37)
    assert L.field_getTime() == ref.time and L.field_getRayCoef() == ref.ray_coef


@pytest.mark.parametrize("model,sweep,w,h,hu", [(1, 0, 2920, 2920, 20), (2, 7, 900, 1500, 10), (5, 4, 800, 2200, 10),
                                                (3, 9, 900, 1900, 10)])
def test_host_models_eps_matches_golden(model, sweep, w, h, hu):
    """models_eps of the C host (shared host/device code) against the reference's maps."""
    import numpy as np

    import mpifdtd_amd as m
    from conftest import load_golden, swept_params
    from mpifdtd_amd import _native as N
    name = {(1, 0): "eps_circle_h20", (2, 7): "eps_layer_s7", (5, 4): "eps_zigzag_s4", (3, 9): "eps_morpho_s9"}[(model, sweep)]
    g = load_golden(name)
    L = N.lib()
    L.models_setModel(model)
    L.models_setParams(ctypes.byref(m.model_params(**swept_params(model, sweep))))
    L.field_init(m.field_info(w, h, hu, 10, 500, 0, 10))
    L.models_initModel()
    nx, ny = g["eps_xy"].shape
    got = np.array([[L.models_eps(x, y, N.D_XY) for y in range(ny)] for x in range(nx)])
    assert np.array_equal(got, g["eps_xy"])
    got = np.array([[L.models_eps(x + 0.5, y, N.D_Y) for y in range(ny)] for x in range(nx)])
    assert np.array_equal(got, g["eps_dy"])


def test_model_sweeps_match_reference_rules():
    """models_isFinish steps the parameters like the reference's nextStructure functions."""
    import mpifdtd_amd as m
    from conftest import swept_params
    from mpifdtd_amd import _native as N
    L = N.lib()
    for model, keys in ((2, ("ml_thick_nm0", "ml_thick_nm1", "ml_edge_rate", "ml_branch_nm", "ml_layer_num")),
                        (5, ("zz_degree", "zz_thick_nm")),
                        (3, ("mo_thick_nm0", "mo_thick_nm1", "mo_edge_rate", "mo_branch_nm", "mo_layer_num"))):
        L.models_setModel(model)
        L.models_setParams(ctypes.byref(m.model_params()))
        for k in range(1, 40):
            L.models_isFinish()
            p = m.ModelParams()
            L.models_getParams(ctypes.byref(p))
            want = swept_params(model, k)
            assert all(getattr(p, key) == want[key] for key in keys), (model, k)


def test_reference_sweep_enumerates_structures_then_angles():
    """fdtd_batch_sweep follows nextSimulation (main.c:114-138): every angle of a
    structure, then models_isFinish() to the next structure, until the sweep ends;
    fitted field sizes follow calcFieldSize (main.c:71-88)."""
    import mpifdtd_amd as m
    from conftest import swept_params
    from mpifdtd_amd import _native as N
    from mpifdtd_amd.batch import reference_sweep
    L = N.lib()
    # the number of structures of the multilayer sweep from the reference's start values
    L.models_setModel(2)
    L.models_setParams(ctypes.byref(m.model_params()))
    structures = 1
    while not L.models_isFinish():
        structures += 1
    assert structures > 1
    items = reference_sweep(4, 2, h_u_nm=10, pml=10, lambda_nm=500, steps=100, start_deg=0, end_deg=30,
                            delta_deg=10)
    assert len(items) == 4 * structures
    keys = ("ml_thick_nm0", "ml_thick_nm1", "ml_edge_rate", "ml_branch_nm", "ml_layer_num")
    for k, it in enumerate(items):
        s, a = divmod(k, 4)
        assert it.field.angle_deg == 10 * a and it.solver == 4 and it.model == 2
        want = swept_params(2, s)
        assert all(getattr(it.params, key) == want[key] for key in keys), k
        L.models_setParams(ctypes.byref(it.params))
        x, y = ctypes.c_int(), ctypes.c_int()
        L.models_needSize(ctypes.byref(x), ctypes.byref(y))
        assert (it.field.width_nm, it.field.height_nm) == (x.value + 10 * 15 * 2 + 200, y.value + 10 * 15 * 2 + 200)
        assert (it.field.h_u_nm, it.field.pml, it.field.lambda_nm, it.field.stepNum) == (10, 10, 500, 100)


def test_reference_sweep_limits_and_errors():
    import mpifdtd_amd as m
    from mpifdtd_amd import _native as N
    L = N.lib()
    fi = N.FieldInfo(2000, 2000, 10, 10, 500, 0, 100)
    L.models_setModel(5)
    L.models_setParams(ctypes.byref(m.model_params()))
    n = L.fdtd_batch_sweep(fi, 5, 5, 0, 0, 1, 0, None, 0)  # zigzag, one angle, fixed size
    assert n > 1
    L.models_setParams(ctypes.byref(m.model_params()))
    arr = (N.BatchItem * 3)()
    assert L.fdtd_batch_sweep(fi, 5, 5, 0, 0, 1, 0, arr, 3) == n  # writes 3, reports all
    assert [(a.field.width_nm, a.field.height_nm, a.model) for a in arr] == [(2000, 2000, 5)] * 3
    assert arr[0].params.zz_degree != arr[1].params.zz_degree
    for bad in ((fi, 2, 5, 0, 0, 1, 0, None, 0),      # not an MPI solver
                (fi, 4, 4, 0, 0, 1, 0, None, 0),      # concentric circles: out of scope, refused
                (fi, 4, 5, 0, 30, 0, 0, None, 0),     # no angle step
                (fi, 4, 5, 30, 0, 10, 0, None, 0),    # end before start
                (fi, 4, 5, 0, 0, 1, 0, None, 2)):     # no output array
        assert L.fdtd_batch_sweep(*bad) < 0, bad


def test_ns_get_solver_returns_an_unset_solver():
    """nsFdtdTM_getSolver / nsFdtdTE_getSolver (nsFdtdTM.c:45-61, nsFdtdTE.c:42-58)
    return a static solver.h Solver (ten function pointers) whose fields are all
    left unset: null, the same object on every call."""
    import ctypes
    from mpifdtd_amd import _native as N
    L = N.lib()
    for name in ("nsFdtdTM_getSolver", "nsFdtdTE_getSolver"):
        f = getattr(L, name)
        f.restype = ctypes.POINTER(ctypes.c_void_p * 10)
        p = f()
        assert all(v is None for v in p.contents), name
        assert ctypes.addressof(p.contents) == ctypes.addressof(f().contents)


def test_finite_check_host_logic():
    """The non-finite check's host side without a GPU call: no active solver
    (FDTD_ESTATE), argument checks, and the empty status."""
    import ctypes
    from mpifdtd_amd import _native as N
    L = N.lib()
    out = (ctypes.c_longlong * 3)()
    assert L.fdtd_solver_checkFinite(out) == -4  # FDTD_ESTATE: no solver
    assert L.fdtd_solver_checkFinite(None) == -1  # FDTD_EINVAL
    assert L.fdtd_solver_setFiniteCheck(-1, 0) == -1
    assert L.fdtd_solver_setFiniteCheck(0, 0) == 0
    st, cnt = ctypes.c_int(7), ctypes.c_longlong(7)
    assert L.fdtd_solver_finiteStatus(ctypes.byref(st), ctypes.byref(cnt)) == 0
    assert (st.value, cnt.value) == (-1, 0)
