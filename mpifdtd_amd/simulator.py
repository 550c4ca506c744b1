"""Python view of the reference driver flow (main.c:150-213) over the C ABI.

    sim = Simulator(MPI_TM_UPML_2D, MIE_CYLINDER, FieldInfo(...))
    sim.calc(steps)             # simulator_calc() x steps (one native call)
    ez, hx, hy = sim.fields()   # the solver's getters (borrowed pointers, copied)

Everything runs through libmpifdtd_amd.so: models_setModel, simulator_setSolver,
simulator_init (field_init -> models_initModel -> solver init), the solver's
update on the GPU, its getters, and simulator_finish.
"""
import ctypes

import numpy as np

from . import _native as N
from ._native import FieldInfo, ModelParams, check

_GETTERS = {
    N.MPI_TM_UPML_2D: ("mpi_fdtdTM_upml_getEz", "mpi_fdtdTM_upml_getHx", "mpi_fdtdTM_upml_getHy", "mpi_fdtdTM_upml_getEps"),
    N.MPI_TE_UPML_2D: ("mpi_fdtdTE_upml_getEx", "mpi_fdtdTE_upml_getEy", "mpi_fdtdTE_upml_getHz", "mpi_fdtdTE_upml_getEps"),
    N.TM_UPML_2D: ("fdtdTM_upml_getEz", "fdtdTM_upml_getHx", "fdtdTM_upml_getHy", "fdtdTM_upml_getEps"),
    N.TE_UPML_2D: ("fdtdTE_upml_getEx", "fdtdTE_upml_getEy", "fdtdTE_upml_getHz", "fdtdTE_upml_getEps"),
    N.TM_2D: ("fdtdTM_getEz", "fdtdTM_getHx", "fdtdTM_getHy", "fdtdTM_getEps"),  # split-field PML (FP64)
    N.TE_2D: ("fdtdTE_getEx", "fdtdTE_getEy", "fdtdTE_getHz", "fdtdTE_getEps"),
    N.NS_TM_2D: ("nsFdtdTM_getEz", "nsFdtdTM_getHx", "nsFdtdTM_getHy", "nsFdtdTM_getEps"),  # non-standard FDTD (FP64)
    N.NS_TE_2D: ("nsFdtdTE_getEx", "nsFdtdTE_getEy", "nsFdtdTE_getHz", "nsFdtdTE_getEps"),
}


def field_info(width_nm, height_nm, h_u_nm, pml, lambda_nm, angle_deg, steps):
    fi = FieldInfo()
    fi.width_nm, fi.height_nm, fi.h_u_nm, fi.pml = width_nm, height_nm, h_u_nm, pml
    fi.lambda_nm, fi.angle_deg, fi.stepNum = lambda_nm, angle_deg, steps
    return fi


def model_params(**kw):
    p = ModelParams()
    N.lib().models_defaultParams(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class Simulator:
    """One simulation through the reference's API (one active solver per process)."""

    def __init__(self, solver, model, fi, params=None, c0=0.7071, ntff_inset=5, point_source=False,
                 ntff_time=True, file_output=False, precision="fp64", T=None, R=16):
        """precision "fp64": reference-order FP64 kernels; "fp32": fused FP32 kernels
        advancing T steps per launch (with ntff_time on they sample the NTFF
        contour after every inner step)."""
        L = N.lib()
        if L.fdtd_hip_device_count() < 1:
            raise RuntimeError("no HIP device visible: the solvers run on the GPU only")
        self.solver = solver
        L.models_setModel(model)
        L.models_setParams(ctypes.byref(params if params is not None else model_params()))
        L.simulator_setSolver(solver)
        L.field_setLightSpeed(c0)
        L.field_setNTFFInset(ntff_inset)
        L.field_setPointSource(1 if point_source else 0)
        L.fdtd_solver_setNTFFTime(1 if ntff_time else 0)
        L.fdtd_solver_setFileOutput(1 if file_output else 0)
        self.precision = precision
        self.per_step_ntff = bool(ntff_time)
        L.fdtd_solver_setPrecision(1 if precision == "fp32" else 0)
        if precision == "fp32":
            L.fdtd_solver_setBlocking(int(T or 4), int(R))
        try:
            L.simulator_init(fi)
        finally:
            L.fdtd_solver_setPrecision(0)
        d = (ctypes.c_int * 4)()
        check(L.fdtd_solver_dims(d), "fdtd_solver_dims")
        self.NPXa, self.NPYa, self.off, self.arraySize = list(d)
        fs = L.field_getFieldInfo_S()
        self.N_PX, self.N_PY = fs.N_PX, fs.N_PY
        self._open = True

    def calc(self, n=1):
        N.lib().fdtd_simulator_run(int(n))

    def set_blocking(self, T, R=16):
        """FP32: steps per fused launch and rows per wavefront of the interior tiles
        (4, 8, 16, or 18 and 20 with the 8-wavefront interior launch shapes 6 to 9)."""
        check(N.lib().fdtd_solver_setBlocking(int(T), int(R)), "fdtd_solver_setBlocking")

    def set_launch_mode(self, mode):
        """FP32: 0 auto, 1 merged, 2 split, 3 split on two streams, 4 / 5 as 2 / 3 with the 48-row
        UPML tiles (TM 12 wavefronts of 4 rows, TE 8 of 6), 6 / 7 as 4 / 5 with the interior tiles
        as 8 wavefronts of R/2 rows, 8 / 9 as 6 / 7 with the UPML tiles as 16 wavefronts of 3 rows
        (FAST_LAUNCH_* of fdtd_hip.h)."""
        check(N.lib().fdtd_solver_setLaunchMode(int(mode)), "fdtd_solver_setLaunchMode")

    def set_ntff_overlap(self, on):
        """FP32 with contour samples: 1 the packing, binning and running DFT of a block beside
        the next block on a stream of their own, 0 in order after it, -1 by grid size."""
        check(N.lib().fdtd_solver_setNTFFOverlap(int(on)), "fdtd_solver_setNTFFOverlap")

    def launch_mode(self):
        """FP32: the launch shape setting, or with auto the shape of the last launch."""
        v = ctypes.c_int()
        check(N.lib().fdtd_solver_getLaunchMode(ctypes.byref(v)), "fdtd_solver_getLaunchMode")
        return v.value

    def blocking(self):
        """(T, R) of this FP32 solver."""
        t, r = ctypes.c_int(), ctypes.c_int()
        check(N.lib().fdtd_solver_getBlocking(ctypes.byref(t), ctypes.byref(r)), "fdtd_solver_getBlocking")
        return t.value, r.value

    @property
    def time(self):
        return N.lib().field_getTime()

    def fields(self):
        """The three getters' arrays (TM: Ez, Hx, Hy; TE: Ex, Ey, Hz) in the solver layout."""
        L = N.lib()
        out = []
        for name in _GETTERS[self.solver][:3]:
            ptr = getattr(L, name)()
            out.append(N.as_array(ptr, (self.NPXa, self.NPYa), np.complex128))
        return out

    def eps(self):
        ptr = getattr(N.lib(), _GETTERS[self.solver][3])()
        return N.as_array(ptr, (self.NPXa, self.NPYa), np.float64)

    def state(self, which):
        out = np.empty((self.NPXa, self.NPYa), np.complex128)
        check(N.lib().fdtd_solver_getState(which, N.dptr(out)), "fdtd_solver_getState")
        return out

    def set_state(self, which, arr):
        a = np.ascontiguousarray(arr, np.complex128).reshape(self.NPXa, self.NPYa)
        check(N.lib().fdtd_solver_setState(which, N.dptr(a)), "fdtd_solver_setState")

    def ntff_time(self, which):
        out = np.empty((360, self.arraySize), np.complex128)
        check(N.lib().fdtd_solver_getNTFFTime(which, N.dptr(out)), "fdtd_solver_getNTFFTime")
        return out

    def check_finite(self):
        """Values with a NaN or Inf part in each physical field (TM Ez Hx Hy, TE Ex Ey Hz)."""
        out = (ctypes.c_longlong * 3)()
        check(N.lib().fdtd_solver_checkFinite(out), "fdtd_solver_checkFinite")
        return list(out)

    def set_finite_check(self, every, abort_on_detect=False):
        """Check the fields every `every` steps inside the step loop (0: off)."""
        check(N.lib().fdtd_solver_setFiniteCheck(int(every), 1 if abort_on_detect else 0),
              "fdtd_solver_setFiniteCheck")

    def finite_status(self):
        """(step time, count) of the first non-finite detection since init / reset, or None."""
        st, cnt = ctypes.c_int(), ctypes.c_longlong()
        hit = N.lib().fdtd_solver_finiteStatus(ctypes.byref(st), ctypes.byref(cnt))
        return (st.value, cnt.value) if hit else None

    def far_field(self):
        out = np.empty(360, np.complex128)
        check(N.lib().fdtd_solver_farField(N.dptr(out)), "fdtd_solver_farField")
        return out

    def far_field_noalloc(self):
        """far_field() through the allocation-free entry point (fdtd_solver_farFieldNoAlloc)."""
        out = np.empty(360, np.complex128)
        check(N.lib().fdtd_solver_farFieldNoAlloc(N.dptr(out)), "fdtd_solver_farFieldNoAlloc")
        return out

    def time_spectra(self):
        out = np.empty((321, 360), np.float64)
        check(N.lib().fdtd_solver_timeSpectra(N.dptr(out)), "fdtd_solver_timeSpectra")
        return out

    def set_dft(self, lambdas_nm):
        """Running DFT of the NTFF contour at these wavelengths (nm), from the next step on."""
        lam = np.ascontiguousarray(np.asarray(lambdas_nm, np.float64).reshape(-1))
        check(N.lib().fdtd_solver_setDFT(N.dptr(lam), lam.size), "fdtd_solver_setDFT")
        self._ndft = lam.size

    def dft(self):
        """[n, 3, 360] complex: the DFT of the three time-domain NTFF accumulator rows
        (TM Ux Uy Wz, TE Wx Wy Uz) of every angle at each wavelength of set_dft."""
        out = np.empty((self._ndft, 3, 360), np.complex128)
        check(N.lib().fdtd_solver_getDFT(N.dptr(out)), "fdtd_solver_getDFT")
        return out

    def dft_spectra(self):
        """[n, 360]: far-field spectra at the set_dft wavelengths, as time_spectra
        translates and normalises them, at the exact frequencies."""
        out = np.empty((self._ndft, 360), np.float64)
        check(N.lib().fdtd_solver_dftSpectra(N.dptr(out)), "fdtd_solver_dftSpectra")
        return out

    def reset(self, angle_deg=None):
        N.lib().simulator_reset()
        if angle_deg is not None:
            N.lib().field_setWaveAngle(int(angle_deg))

    def finish(self):
        if self._open:
            N.lib().simulator_finish()
            self._open = False
