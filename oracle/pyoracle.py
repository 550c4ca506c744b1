"""ctypes wrapper of oracle/liboracle.so -- the CPU parity oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

TM_2D, TE_2D = 0, 1  # split-field PML TM / TE (fdtdTM.c, fdtdTE.c)
NS_TM_2D, NS_TE_2D = 6, 7  # non-standard FDTD TM / TE (nsFdtdTM.c, nsFdtdTE.c)
TM_UPML, TE_UPML, MPI_TM_UPML, MPI_TE_UPML = 2, 3, 4, 5
NO_MODEL, MIE_CYLINDER, LAYER, MORPHO_SCALE, ZIGZAG = 0, 1, 2, 3, 5
D_X, D_Y, D_XY = 0, 1, 2


class ModelParams(ctypes.Structure):
    _fields_ = [
        ("circle_radius_nm", ctypes.c_double), ("circle_n", ctypes.c_double),
        ("ml_n0", ctypes.c_double), ("ml_n1", ctypes.c_double),
        ("ml_n0x", ctypes.c_double), ("ml_n1x", ctypes.c_double),
        ("ml_uniaxial", ctypes.c_int), ("ml_asymmetry", ctypes.c_int), ("ml_use_gap", ctypes.c_int),
        ("ml_width_nm0", ctypes.c_int), ("ml_width_nm1", ctypes.c_int),
        ("ml_thick_nm0", ctypes.c_int), ("ml_thick_nm1", ctypes.c_int),
        ("ml_layer_num", ctypes.c_int), ("ml_branch_nm", ctypes.c_int), ("ml_left_gap_nm", ctypes.c_int),
        ("ml_edge_rate", ctypes.c_double), ("ml_curve", ctypes.c_double),
        ("zz_width_nm", ctypes.c_int), ("zz_thick_nm", ctypes.c_int),
        ("zz_layer_num", ctypes.c_int), ("zz_degree", ctypes.c_int),
        ("zz_n", ctypes.c_double),
        ("mo_width_nm", ctypes.c_int), ("mo_thick_nm0", ctypes.c_int), ("mo_thick_nm1", ctypes.c_int),
        ("mo_layer_num", ctypes.c_int), ("mo_branch_nm", ctypes.c_int), ("mo_left_gap_nm", ctypes.c_int),
        ("mo_edge_rate", ctypes.c_double), ("mo_curve", ctypes.c_double),
        ("mo_n", ctypes.c_double), ("mo_nx", ctypes.c_double),
        ("mo_uniaxial", ctypes.c_int), ("mo_asymmetry", ctypes.c_int),
        ("mo_randomness", ctypes.c_int), ("mo_seed", ctypes.c_int),
    ]


class Config(ctypes.Structure):
    _fields_ = [
        ("width_nm", ctypes.c_int), ("height_nm", ctypes.c_int), ("h_u_nm", ctypes.c_int),
        ("pml", ctypes.c_int), ("lambda_nm", ctypes.c_int), ("angle_deg", ctypes.c_int),
        ("stepNum", ctypes.c_int), ("solver", ctypes.c_int), ("model", ctypes.c_int),
        ("c0", ctypes.c_double), ("ntff_inset", ctypes.c_int), ("point_source", ctypes.c_int),
        ("ntff_time", ctypes.c_int), ("mp", ModelParams),
    ]


def build():
    """Compile liboracle.so (gcc -O2, no FMA contraction)."""
    subprocess.run(["make", "-C", _HERE, "liboracle.so"], check=True, capture_output=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_default_config.argtypes = [ctypes.POINTER(Config)]
        L.orc_default_model_params.argtypes = [ctypes.POINTER(ModelParams)]
        L.orc_create.argtypes = [ctypes.POINTER(Config)]
        L.orc_create.restype = P
        L.orc_destroy.argtypes = [P]
        L.orc_step.argtypes = [P]
        L.orc_run.argtypes = [P, I]
        L.orc_reset.argtypes = [P]
        L.orc_set_angle.argtypes = [P, I]
        L.orc_dims.argtypes = [P, ip]
        L.orc_ntff_info.argtypes = [P, ip, dp]
        for f in ("orc_time", "orc_ray_coef", "orc_omega", "orc_k"):
            getattr(L, f).argtypes = [P]
            getattr(L, f).restype = D
        L.orc_get_field.argtypes = [P, I, dp]
        L.orc_set_field.argtypes = [P, I, dp]
        L.orc_get_eps.argtypes = [P, I, dp]
        L.orc_get_ntff_time.argtypes = [P, I, dp]
        L.orc_ntff_frequency.argtypes = [P, dp]
        L.orc_time_spectra.argtypes = [P, dp]
        L.orc_time_translate.argtypes = [P, dp, dp]
        L.orc_model_eps.argtypes = [P, D, D, I]
        L.orc_model_eps.restype = D
        L.orc_model_eps_map.argtypes = [P, I, D, D, dp]
        L.orc_sigma_x.argtypes = [P, D, D]
        L.orc_sigma_x.restype = D
        L.orc_sigma_y.argtypes = [P, D, D]
        L.orc_sigma_y.restype = D
        L.orc_cfft.argtypes = [dp, I]
        L.orc_elliptic.argtypes = [P, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def default_config(**kw):
    c = Config()
    lib().orc_default_config(ctypes.byref(c))
    for k, v in kw.items():
        if hasattr(c.mp, k) and not hasattr(c, k):
            setattr(c.mp, k, v)
        else:
            setattr(c, k, v)
    return c


class Sim:
    """One oracle simulation instance (the reference's globals, made explicit)."""

    def __init__(self, cfg=None, **kw):
        self.cfg = cfg if cfg is not None else default_config(**kw)
        self.h = lib().orc_create(ctypes.byref(self.cfg))
        d = (ctypes.c_int * 8)()
        lib().orc_dims(self.h, d)
        (self.N_PX, self.N_PY, self.NPXa, self.NPYa, self.off, self.N_PML,
         self.arraySize, self.maxTime) = list(d)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def step(self, n=1):
        lib().orc_run(self.h, n)

    def reset(self):
        lib().orc_reset(self.h)

    def set_angle(self, deg):
        lib().orc_set_angle(self.h, deg)

    @property
    def time(self):
        return lib().orc_time(self.h)

    @property
    def ray_coef(self):
        return lib().orc_ray_coef(self.h)

    @property
    def omega(self):
        return lib().orc_omega(self.h)

    @property
    def k(self):
        return lib().orc_k(self.h)

    def ntff_info(self):
        o = (ctypes.c_int * 6)()
        rf = ctypes.c_double()
        lib().orc_ntff_info(self.h, o, ctypes.byref(rf))
        return dict(zip(("top", "bottom", "left", "right", "cx", "cy"), list(o)), RFperC=rf.value)

    def field(self, which):
        out = np.empty((self.NPXa, self.NPYa), np.complex128)
        lib().orc_get_field(self.h, which, _dp(out))
        return out

    def set_field(self, which, arr):
        a = np.ascontiguousarray(arr, np.complex128).reshape(self.NPXa, self.NPYa)
        lib().orc_set_field(self.h, which, _dp(a))

    def eps(self, which=0):
        out = np.empty((self.NPXa, self.NPYa), np.float64)
        lib().orc_get_eps(self.h, which, _dp(out))
        return out

    def ntff_time(self, which):
        out = np.empty((360, self.arraySize), np.complex128)
        lib().orc_get_ntff_time(self.h, which, _dp(out))
        return out

    def far_field(self):
        out = np.empty(360, np.complex128)
        lib().orc_ntff_frequency(self.h, _dp(out))
        return out

    def elliptic(self):
        """TM_2D / TE_2D: the 181 values field_outputElliptic writes (field.c:322-341), for 0..180."""
        out = np.empty(181, np.float64)
        lib().orc_elliptic(self.h, _dp(out))
        return out

    def time_spectra(self):
        out = np.empty((321, 360), np.float64)
        lib().orc_time_spectra(self.h, _dp(out))
        return out

    def time_translate(self):
        eth = np.zeros((360, self.arraySize), np.complex128)
        eph = np.zeros((360, self.arraySize), np.complex128)
        lib().orc_time_translate(self.h, _dp(eth), _dp(eph))
        return eth, eph

    def model_eps(self, x, y, mode):
        return lib().orc_model_eps(self.h, float(x), float(y), int(mode))

    def model_eps_map(self, mode, dx=0.0, dy=0.0):
        out = np.empty((self.N_PX, self.N_PY), np.float64)
        lib().orc_model_eps_map(self.h, int(mode), float(dx), float(dy), _dp(out))
        return out

    def sigma_x(self, x):
        return lib().orc_sigma_x(self.h, float(x), 0.0)

    def sigma_y(self, y):
        return lib().orc_sigma_y(self.h, 0.0, float(y))


def cfft(data):
    a = np.ascontiguousarray(data, np.complex128).copy()
    lib().orc_cfft(_dp(a), a.size)
    return a
