"""The host-array helpers of field.h / myComplex.h against the reference's own.

Fixture ref_field_util: the reference's field.c and myComplex.c run by
oracle/ref_harness (mode=util) on fixed inputs after 480 field_nextStep calls,
on a 90 x 100 grid (h = 20 nm, lambda 450 nm, angle 30).  The library's
functions take the same inputs after the same field_init and steps, on the CPU
(no GPU call): the four sources added to a caller's array, the PML / NS-PML
coefficient formulas, cbilinear / cnorm, the subdomain getters and ind()
bit for bit, and the three text writers byte for byte.
"""
import ctypes

import numpy as np

from conftest import load_golden


class C2(ctypes.Structure):
    """A double complex by value (two SSE-class doubles, as the C ABI passes it)."""
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]


def _lib():
    from mpifdtd_amd import _native as N
    L = N.lib()
    D, I, cp, dp = ctypes.c_double, ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)
    for nm in ("field_scatteredWave", "field_scatteredWaveNotUPML"):
        getattr(L, nm).argtypes = [cp, dp, D, D]
    for nm in ("field_nsScatteredWaveNotUPML", "field_scatteredPulse"):
        getattr(L, nm).argtypes = [cp, dp, D, D, D]
    for nm in ("field_pmlCoef", "field_pmlCoef_LXY", "field_ns_beta", "field_ns_beta_aster"):
        getattr(L, nm).argtypes = [D, D]
        getattr(L, nm).restype = D
    for nm in ("field_getOffsetX", "field_getOffsetY", "field_getSubNx", "field_getSubNy", "field_getSubNpx",
               "field_getSubNpy", "field_getSubNcell"):
        getattr(L, nm).restype = I
    L.ind.argtypes = [I, I]
    L.ind.restype = I
    L.cnorm.argtypes = [C2]
    L.cnorm.restype = D
    L.cbilinear.argtypes = [cp, D, D, I, I]
    L.cbilinear.restype = C2
    L.field_outputElliptic.argtypes = [ctypes.c_char_p, cp]
    L.field_outputAllDataComplex.argtypes = [ctypes.c_char_p, cp]
    L.field_outputAllDataDouble.argtypes = [ctypes.c_char_p, dp]
    return N, L


def test_field_and_complex_helpers_match_the_reference(tmp_path, monkeypatch):
    g = load_golden("field_util")
    N, L = _lib()
    w, h, hu, pml, lam, ang, steps = [int(v) for v in g["cfg"]]
    L.field_setLightSpeed(0.7071)
    fi = N.FieldInfo(w, h, hu, pml, lam, ang, steps)
    L.field_init(fi)
    for _ in range(steps):
        L.field_nextStep()
    eps = np.ascontiguousarray(g["eps"])
    npx, npy = eps.shape
    assert (L.field_getFieldInfo_S().N_PX, L.field_getFieldInfo_S().N_PY) == (npx, npy)
    dp = eps.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    src = np.stack([np.array(g["base"]) for _ in range(4)])
    L.field_scatteredWave(src[0].ctypes.data, dp, 0.5, 0.0)
    L.field_scatteredWaveNotUPML(src[1].ctypes.data, dp, 0.0, 0.5)
    L.field_nsScatteredWaveNotUPML(src[2].ctypes.data, dp, 0.0, 0.5, 0.6)
    L.field_scatteredPulse(src[3].ctypes.data, dp, 0.5, 0.0, 0.8)
    for q in range(4):
        assert np.abs(src[q] - g["base"]).max() > 0, q  # every source added something
        assert np.array_equal(src[q], g["src"][q]), q
    av, bv = (0, 0.3, 1.7), (0.2, 1.0, 2.5)
    scal = np.array([[L.field_pmlCoef(1 + a, b), L.field_pmlCoef_LXY(1 + a, b), L.field_ns_beta(a, b),
                      L.field_ns_beta_aster(a, b)] for a in av for b in bv])
    assert np.array_equal(scal, g["scal"])
    base = np.ascontiguousarray(g["base"])
    px = (1.0, 2.5, 10.25, npx / 2.0 + 0.3, npx - 2.5)
    py = (1.0, 3.75, 20.5, npy / 2.0 - 0.7, npy - 2.25)
    bil = np.array([complex(r.re, r.im) for r in (L.cbilinear(base.ctypes.data, x, y, npx, npy) for x, y in zip(px, py))])
    assert np.array_equal(bil, g["bil"])
    nrm = np.array([L.cnorm(C2(b.real, b.imag)) for b in bil])
    assert np.array_equal(nrm, g["nrm"])
    meta = [npx, npy, L.field_getOffsetX(), L.field_getOffsetY(), L.field_getSubNx(), L.field_getSubNy(),
            L.field_getSubNpx(), L.field_getSubNpy(), L.field_getSubNcell(), L.ind(3, 4)]
    assert meta == list(g["meta"])
    monkeypatch.chdir(tmp_path)
    L.field_outputElliptic(b"ell.txt", src[0].ctypes.data)
    L.field_outputAllDataComplex(b"all_c.txt", src[1].ctypes.data)
    L.field_outputAllDataDouble(b"all_d.txt", dp)
    for nm in ("ell", "all_c", "all_d"):
        assert (tmp_path / f"{nm}.txt").read_text() == str(g[nm]), nm
