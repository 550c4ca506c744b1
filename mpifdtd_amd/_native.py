"""ctypes bindings of mpifdtd_amd/lib/libmpifdtd_amd.so (the C-ABI boundary).

The library holds the C host (field / models / simulator / solver entry
points, the reference's API) and the gfx950 HIP kernels.  There is no Python
or CPU fallback: if the library is missing, every entry point raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmpifdtd_amd.so")

# enum SOLVER (simulator.h:8-18)
TM_2D, TE_2D, TM_UPML_2D, TE_UPML_2D, MPI_TM_UPML_2D, MPI_TE_UPML_2D, NS_TM_2D, NS_TE_2D = range(8)
# enum MODEL (models.h:5-14)
NO_MODEL, MIE_CYLINDER, LAYER, MORPHO_SCALE, CONCENTRIC_CIRCLE, ZIGZAG, TRACE_IMAGE = range(7)
# enum MODE (models.h:16-21)
D_X, D_Y, D_XY = range(3)


class NativeLibraryMissing(RuntimeError):
    pass


class FieldInfo(ctypes.Structure):
    """field.h:20-28"""
    _fields_ = [(n, ctypes.c_int) for n in
                ("width_nm", "height_nm", "h_u_nm", "pml", "lambda_nm", "angle_deg", "stepNum")]


class FieldInfo_S(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("N_X", "N_Y", "N_PX", "N_PY", "N_CELL", "N_PML", "DX", "DY")]


class NTFFInfo(ctypes.Structure):
    _fields_ = [("top", ctypes.c_int), ("bottom", ctypes.c_int), ("left", ctypes.c_int), ("right", ctypes.c_int),
                ("cx", ctypes.c_int), ("cy", ctypes.c_int), ("RFperC", ctypes.c_double), ("arraySize", ctypes.c_int)]


class ModelParams(ctypes.Structure):
    """fdtd_model_params (include/mpifdtd_models.h)"""
    _fields_ = [
        ("circle_radius_nm", ctypes.c_double), ("circle_n", ctypes.c_double),
        ("ml_n0", ctypes.c_double), ("ml_n1", ctypes.c_double), ("ml_n0x", ctypes.c_double), ("ml_n1x", ctypes.c_double),
        ("ml_uniaxial", ctypes.c_int), ("ml_asymmetry", ctypes.c_int), ("ml_use_gap", ctypes.c_int),
        ("ml_width_nm0", ctypes.c_int), ("ml_width_nm1", ctypes.c_int),
        ("ml_thick_nm0", ctypes.c_int), ("ml_thick_nm1", ctypes.c_int),
        ("ml_layer_num", ctypes.c_int), ("ml_branch_nm", ctypes.c_int), ("ml_left_gap_nm", ctypes.c_int),
        ("ml_edge_rate", ctypes.c_double), ("ml_curve", ctypes.c_double),
        ("zz_width_nm", ctypes.c_int), ("zz_thick_nm", ctypes.c_int), ("zz_layer_num", ctypes.c_int),
        ("zz_degree", ctypes.c_int), ("zz_n", ctypes.c_double),
        ("mo_width_nm", ctypes.c_int), ("mo_thick_nm0", ctypes.c_int), ("mo_thick_nm1", ctypes.c_int),
        ("mo_layer_num", ctypes.c_int), ("mo_branch_nm", ctypes.c_int), ("mo_left_gap_nm", ctypes.c_int),
        ("mo_edge_rate", ctypes.c_double), ("mo_curve", ctypes.c_double), ("mo_n", ctypes.c_double),
        ("mo_nx", ctypes.c_double), ("mo_uniaxial", ctypes.c_int), ("mo_asymmetry", ctypes.c_int),
        ("mo_randomness", ctypes.c_int), ("mo_seed", ctypes.c_int),
    ]


class RunConfig(ctypes.Structure):
    """fdtd_run_config of mpifdtd_simulator.h: the values of a config.txt (main.c:319-366)."""
    _fields_ = [("field_info", FieldInfo), ("startAngle", ctypes.c_int), ("endAngle", ctypes.c_int),
                ("deltaAngle", ctypes.c_int), ("model", ctypes.c_int), ("solver", ctypes.c_int)]


class BatchItem(ctypes.Structure):
    """fdtd_batch_item of mpifdtd_batch.h: one instance of a batch."""
    _fields_ = [("field", FieldInfo), ("solver", ctypes.c_int), ("model", ctypes.c_int), ("params", ModelParams)]


_lib = None


def lib():
    """Load the native library (raises NativeLibraryMissing if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeLibraryMissing(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(or make -C mpifdtd_amd/csrc); there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    I, D, V, P = ctypes.c_int, ctypes.c_double, None, ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)

    def sig(name, res, *args):
        f = getattr(L, name)
        f.restype = res
        f.argtypes = list(args)

    # field.h
    sig("field_init", V, FieldInfo)
    sig("field_reset", V)
    sig("field_nextStep", V)
    sig("field_isFinish", I)
    sig("field_setWaveAngle", V, I)
    for n in ("field_getT", "field_getK", "field_getRayCoef", "field_getOmega", "field_getLambda",
              "field_getWaveAngle", "field_getTime", "field_getMaxTime", "field_getLightSpeed",
              "field_getMu0", "field_getZ0"):
        sig(n, D)
    sig("field_sigmaX", D, D, D)
    sig("field_sigmaY", D, D, D)
    sig("field_toCellUnit", D, D)
    sig("field_toPhisycalUnit", D, D)
    sig("field_getNTFFInfo", NTFFInfo)
    sig("field_getFieldInfo_S", FieldInfo_S)
    sig("field_getFieldInfo", FieldInfo)
    sig("field_index", I, I, I)
    sig("field_subIndex", I, I, I)
    sig("field_setLightSpeed", V, D)
    sig("field_setNTFFInset", V, I)
    sig("field_setPointSource", V, I)
    sig("field_getPointSource", I)
    # models.h
    sig("models_setModel", V, I)
    sig("models_eps", D, D, D, I)
    sig("models_isFinish", I)
    sig("models_needSize", V, ip, ip)
    sig("models_initModel", V)
    sig("models_defaultParams", V, ctypes.POINTER(ModelParams))
    sig("models_setParams", V, ctypes.POINTER(ModelParams))
    sig("models_getParams", V, ctypes.POINTER(ModelParams))
    sig("models_getModel", I)
    # simulator.h
    sig("simulator_setSolver", V, I)
    sig("simulator_init", V, FieldInfo)
    sig("simulator_calc", V)
    sig("simulator_isFinish", I)
    sig("simulator_finish", V)
    sig("simulator_reset", V)
    sig("simulator_getDrawingData", P)
    sig("simulator_getEps", P)
    sig("simulator_elapsed", D)
    sig("fdtd_simulator_run", V, I)
    sig("fdtd_readConfig", I, ctypes.c_char_p, ctypes.POINTER(RunConfig))
    # solver extensions
    sig("fdtd_solver_getState", I, I, dp)
    sig("fdtd_solver_setState", I, I, dp)
    sig("fdtd_solver_getNTFFTime", I, I, dp)
    sig("fdtd_solver_farField", I, dp)
    sig("fdtd_solver_farFieldNoAlloc", I, dp)
    sig("fdtd_solver_timeSpectra", I, dp)
    sig("fdtd_solver_setDFT", I, dp, I)
    sig("fdtd_solver_getDFT", I, dp)
    sig("fdtd_solver_dftSpectra", I, dp)
    sig("fdtd_solver_setNTFFTime", V, I)
    sig("fdtd_solver_setFileOutput", V, I)
    sig("fdtd_solver_dims", I, ip)
    sig("fdtd_solver_setPrecision", V, I)
    sig("fdtd_solver_getPrecision", I)
    sig("fdtd_solver_setBlocking", I, I, I)
    sig("fdtd_solver_getBlocking", I, ip, ip)
    sig("fdtd_solver_setLaunchMode", I, I)
    sig("fdtd_solver_setNTFFOverlap", I, I)
    sig("fdtd_solver_getLaunchMode", I, ip)
    llp = ctypes.POINTER(ctypes.c_longlong)
    sig("fdtd_solver_checkFinite", I, llp)
    sig("fdtd_solver_setFiniteCheck", I, I, I)
    sig("fdtd_solver_finiteStatus", I, ip, llp)
    # mpifdtd_batch.h
    sig("fdtd_batch_create", I, ctypes.POINTER(BatchItem), I, I, I, ctypes.POINTER(P))
    sig("fdtd_batch_run", I, P)
    sig("fdtd_batch_run_blocks", I, P, I)
    sig("fdtd_batch_sync", I, P, ip)
    sig("fdtd_batch_set_launch_mode", I, P, I)
    sig("fdtd_batch_size", I, P)
    sig("fdtd_batch_dims", I, P, I, ip)
    sig("fdtd_batch_field", I, P, I, I, dp)
    sig("fdtd_batch_far_field", I, P, I, dp)
    sig("fdtd_batch_destroy", V, P)
    sig("fdtd_batch_sweep", I, FieldInfo, I, I, I, I, I, I, ctypes.POINTER(BatchItem), I)
    # mpifdtd_surrogate.h
    vp = ctypes.c_void_p
    fp = ctypes.POINTER(ctypes.c_float)
    sig("mlp_create", I, I, ip, I, ctypes.c_ulonglong, ctypes.POINTER(P))
    sig("mlp_destroy", I, P)
    sig("mlp_forward", I, P, vp, I, vp)
    sig("mlp_grads", I, P, vp, vp, I, fp)
    sig("mlp_train_step", I, P, vp, vp, I, ctypes.c_float, fp)
    sig("mlp_get_params", I, P, I, fp, fp)
    sig("mlp_set_params", I, P, I, fp, fp)
    sig("mlp_get_grads", I, P, I, fp, fp)
    sig("mlp_gemm_bf16", I, vp, vp, vp, I, I, I, I, I)
    sig("mlp_train_epoch", I, P, vp, vp, I, I, ctypes.c_ulonglong, ctypes.c_float, fp)
    sig("fdtd_hip_alloc", I, ctypes.c_size_t, ctypes.POINTER(vp))
    sig("fdtd_hip_free", I, vp)
    sig("fdtd_hip_upload", I, vp, vp, ctypes.c_size_t)
    sig("fdtd_hip_download", I, vp, vp, ctypes.c_size_t)
    sig("fdtd_hip_synchronize", I)
    sig("fdtd_hip_trim", I)
    sig("fdtd_hip_profile_enable", V, I)
    sig("fdtd_hip_profile_read", V, dp)
    sig("fdtd_hip_copy_rate", I, ctypes.c_size_t, I, dp)
    sig("fdtd_hip_last_error", ctypes.c_char_p)
    sig("fdtd_hip_device_count", I)
    for n in ("mpi_fdtdTM_upml_getEz", "mpi_fdtdTM_upml_getHx", "mpi_fdtdTM_upml_getHy", "mpi_fdtdTM_upml_getEps",
              "mpi_fdtdTE_upml_getEx", "mpi_fdtdTE_upml_getEy", "mpi_fdtdTE_upml_getHz", "mpi_fdtdTE_upml_getEps",
              "fdtdTM_upml_getEz", "fdtdTM_upml_getHx", "fdtdTM_upml_getHy", "fdtdTM_upml_getEps",
              "fdtdTE_upml_getEx", "fdtdTE_upml_getEy", "fdtdTE_upml_getHz", "fdtdTE_upml_getEps",
              "fdtdTM_getEz", "fdtdTM_getHx", "fdtdTM_getHy", "fdtdTM_getEzx", "fdtdTM_getEzy", "fdtdTM_getEps",
              "fdtdTE_getEx", "fdtdTE_getEy", "fdtdTE_getHz", "fdtdTE_getHzx", "fdtdTE_getHzy", "fdtdTE_getEps",
              "nsFdtdTM_getEz", "nsFdtdTM_getHx", "nsFdtdTM_getHy", "nsFdtdTM_getEzx", "nsFdtdTM_getEzy",
              "nsFdtdTM_getEps", "nsFdtdTE_getEx", "nsFdtdTE_getEy", "nsFdtdTE_getHz", "nsFdtdTE_getHzx",
              "nsFdtdTE_getHzy", "nsFdtdTE_getEps"):
        sig(n, P)
    _lib = L
    return L


def check(rc, what=""):
    if rc != 0:
        msg = lib().fdtd_hip_last_error()
        raise RuntimeError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
    return rc


def as_array(ptr, shape, dtype):
    """Copy `prod(shape)` elements of `dtype` from a borrowed C pointer."""
    if not ptr:
        raise RuntimeError("null pointer from the native library")
    n = int(np.prod(shape)) * np.dtype(dtype).itemsize
    buf = (ctypes.c_char * n).from_address(ptr)
    return np.frombuffer(buf, dtype=dtype).reshape(shape).copy()


def dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
