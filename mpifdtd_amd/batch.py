"""Many MPI-solver simulations advanced together on one GPU (mpifdtd_batch.h).

    items = sweep_instances(seed=1, n=64)       # seeded multilayer / zigzag sweep
    with Batch(items, T=8, R=16) as b:
        b.run()                                  # every instance to its stepNum
        ff = b.far_field(0)                      # 360-angle far field of instance 0

The reference runs such a sweep as one simulation per MPI rank
(main.c:114-138); here the instances share the GPU, one launch per block of T
steps pulling the tiles of every unfinished instance from device queues.
"""
import ctypes
import math

import numpy as np

from . import _native as N
from ._native import BatchItem, check


def item(solver, model, width_nm, height_nm, h_u_nm, pml, lambda_nm, angle_deg, steps, **params):
    """One batch instance: what simulator_init and the model setters take for one run."""
    it = BatchItem()
    f = it.field
    f.width_nm, f.height_nm, f.h_u_nm, f.pml = width_nm, height_nm, h_u_nm, pml
    f.lambda_nm, f.angle_deg, f.stepNum = lambda_nm, angle_deg, steps
    it.solver, it.model = solver, model
    N.lib().models_defaultParams(ctypes.byref(it.params))
    for k, v in params.items():
        setattr(it.params, k, v)
    return it


def reference_sweep(solver, model, h_u_nm, pml, lambda_nm, steps, start_deg=0, end_deg=0, delta_deg=1,
                    fit_size=True, width_nm=0, height_nm=0, **params):
    """The reference's parameter sweep (main.c:114-138 nextSimulation) as batch items:
    every incidence angle from start_deg to end_deg by delta_deg for each structure of
    `model`, from the given parameters (the reference's start values by default) until
    models_isFinish() ends the sweep; each structure's field sized as calcFieldSize
    does (main.c:71-88) when fit_size, else width_nm x height_nm."""
    L = N.lib()
    L.models_setModel(model)
    p = N.ModelParams()
    L.models_defaultParams(ctypes.byref(p))
    for k, v in params.items():
        setattr(p, k, v)
    L.models_setParams(ctypes.byref(p))
    base = N.FieldInfo(width_nm, height_nm, h_u_nm, pml, lambda_nm, start_deg, steps)
    args = (base, solver, model, start_deg, end_deg, delta_deg, 1 if fit_size else 0)
    n = L.fdtd_batch_sweep(*args, None, 0)
    check(n if n < 0 else 0, "fdtd_batch_sweep")
    L.models_setParams(ctypes.byref(p))  # the same sweep again, written this time
    arr = (BatchItem * n)()
    check(min(0, L.fdtd_batch_sweep(*args, arr, n)), "fdtd_batch_sweep")
    return list(arr)


class Batch:
    def __init__(self, items, T=8, R=16, c0=0.7071, ntff_inset=5):
        """c0 and ntff_inset: the field settings every instance is created with
        (field_setLightSpeed, field_setNTFFInset; the reference's values by
        default), set here because the C host keeps them as process globals."""
        L = N.lib()
        if L.fdtd_hip_device_count() < 1:
            raise RuntimeError("no HIP device visible: the solvers run on the GPU only")
        L.field_setLightSpeed(c0)
        L.field_setNTFFInset(ntff_inset)
        L.field_setPointSource(0)
        arr = (BatchItem * len(items))(*items)
        h = ctypes.c_void_p()
        check(L.fdtd_batch_create(arr, len(items), int(T), int(R), ctypes.byref(h)), "fdtd_batch_create")
        self._h = h
        self.n = len(items)
        self.T, self.R = T, R

    def set_launch_mode(self, mode):
        """2: interior tiles then UPML tiles; 3: the two side by side on two streams; 4 / 5: 2 / 3
        with the UPML tiles as 8 wavefronts (5 is the default); 6 / 7: 4 / 5 with the interior
        tiles as 8 wavefronts as well (FAST_LAUNCH_* of fdtd_hip.h)."""
        check(N.lib().fdtd_batch_set_launch_mode(self._h, int(mode)), "fdtd_batch_set_launch_mode")

    def run(self):
        """Every instance to its step count (synchronised)."""
        check(N.lib().fdtd_batch_run(self._h), "fdtd_batch_run")

    def run_blocks(self, blocks):
        """Queue `blocks` blocks of T steps (asynchronous; call sync() before reading)."""
        check(N.lib().fdtd_batch_run_blocks(self._h, int(blocks)), "fdtd_batch_run_blocks")

    def sync(self):
        done = (ctypes.c_int * self.n)()
        check(N.lib().fdtd_batch_sync(self._h, done), "fdtd_batch_sync")
        return list(done)

    def dims(self, i):
        d = (ctypes.c_int * 4)()
        check(N.lib().fdtd_batch_dims(self._h, i, d), "fdtd_batch_dims")
        return list(d)

    def field(self, i, which):
        """State slot `which` of instance i (TM: 0 Ez, 3 Hx, 6 Hy; TE: 0 Ex, 3 Ey, 6 Hz) in the solver layout."""
        NPXa, NPYa, _, _ = self.dims(i)
        out = np.empty((NPXa, NPYa), np.complex128)
        check(N.lib().fdtd_batch_field(self._h, i, which, N.dptr(out)), "fdtd_batch_field")
        return out

    def far_field(self, i):
        out = np.empty(360, np.complex128)
        check(N.lib().fdtd_batch_far_field(self._h, i, N.dptr(out)), "fdtd_batch_far_field")
        return out

    def close(self):
        if self._h:
            N.lib().fdtd_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014): a small seeded generator, so that a
    sweep is reproduced exactly from its seed on any machine."""

    def __init__(self, seed):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self, a, b):
        return a + (b - a) * (self.next() >> 11) / float(1 << 53)

    def integer(self, a, b):
        """uniform integer in [a, b]"""
        return a + self.next() % (b - a + 1)


SIDES = (512, 768, 1024, 1536, 2048)


def sweep_spec(seed=1, n=64, h_nm=5, pml=16, steps=(6000, 10000)):
    """Seeded sweep of uneven instances: multilayer stacks (2-20 layers, eps in
    [1, 4], thicknesses 50-300 nm) or zigzag gratings (period 200-800 nm,
    amplitude 50-300 nm, eps in [1.5, 3]) on square grids of side 512 ... 2048
    cells plus the UPML, TM or TE, lambda in [380, 780] nm (integer nm), step
    counts in `steps`.  Returns plain dicts (the instance list to commit or
    print); `sweep_instances` turns them into batch items."""
    g = SplitMix64(seed)
    out = []
    for _ in range(n):
        kind = "multilayer" if g.next() & 1 else "zigzag"
        side = SIDES[g.integer(0, len(SIDES) - 1)]
        tm = bool(g.next() & 1)
        lam = int(round(g.uniform(380, 780)))
        nsteps = g.integer(steps[0], steps[1])
        span_nm = 0.8 * side * h_nm  # the structure stays inside the inner 80 % of the grid
        if kind == "multilayer":
            n0 = math.sqrt(g.uniform(1.0, 4.0))
            n1 = math.sqrt(g.uniform(1.0, 4.0))
            t0, t1 = g.integer(50, 300), g.integer(50, 300)
            layers = min(g.integer(2, 20), max(1, int(span_nm // (t0 + t1))))
            width = int(0.5 * side * h_nm)
            params = dict(ml_n0=n0, ml_n1=n1, ml_thick_nm0=t0, ml_thick_nm1=t1, ml_layer_num=layers,
                          ml_width_nm0=width, ml_width_nm1=width, ml_branch_nm=0, ml_use_gap=0)
        else:
            period, amp = g.integer(200, 800), g.integer(50, 300)
            zn = math.sqrt(g.uniform(1.5, 3.0))
            deg = g.integer(10, 80)
            rad = math.radians(deg)
            per_layer = math.sin(rad) * period + amp * math.cos(rad)
            layers = min(g.integer(3, 11), max(1, int(span_nm // max(per_layer, 1.0)) - 1))
            params = dict(zz_width_nm=period, zz_thick_nm=amp, zz_n=zn, zz_degree=deg, zz_layer_num=layers)
        w = (side + 2 * pml) * h_nm
        out.append(dict(kind=kind, side=side, mode="tm" if tm else "te", lambda_nm=lam, steps=nsteps,
                        width_nm=w, height_nm=w, h_nm=h_nm, pml=pml, params=params))
    return out


def sweep_instances(spec):
    items = []
    for s in spec:
        solver = N.MPI_TM_UPML_2D if s["mode"] == "tm" else N.MPI_TE_UPML_2D
        model = N.LAYER if s["kind"] == "multilayer" else N.ZIGZAG
        items.append(item(solver, model, s["width_nm"], s["height_nm"], s["h_nm"], s["pml"], s["lambda_nm"], 0,
                          s["steps"], **s["params"]))
    return items
