"""Reflectance surrogate: an ensemble of MLPs trained on FDTD spectra
(mpifdtd_surrogate.h).

    geoms = morpho_geometries(seed=1, n=247)            # seeded morpho-scale geometries
    X, Y = generate_sweep(geoms)                        # 81 wavelengths per geometry: one pulse
                                                        # run on the 8256 x 4160 grid each
    spec = morpho_spec(seed=1, n=2000)                  # or: seeded (geometry, wavelength) pairs
    X, Y = generate(spec)                               # batched CW FDTD on a reduced grid
    ens = Ensemble([13, 512, 512, 512, 180], k=4)
    ens.fit(X, Y, epochs=200)                           # BF16 matrix-core training
    R, used_fdtd = predict_with_fallback(ens, spec_q)   # FDTD where the members disagree

The reference sweeps geometries one FDTD run at a time (main.c:114-138,
models_isFinish); here a network maps 12 morpho-scale parameters plus the
wavelength to the far-field reflectance at the 180 back-hemisphere angles, and
the FDTD stays both its teacher and its fallback.  The forward and backward
passes, the loss and Adam run in the library's own kernels (BF16 products on
the matrix cores, FP32 accumulation and master weights), on device buffers of
the library's own HIP runtime (DeviceArray).
"""
import ctypes
import math

import numpy as np

from . import _native as N
from ._native import check
from .batch import SplitMix64, item

# the 12 geometry parameters (morphoScaleModel.c:10-99) and their sampling ranges
GEOMETRY = (
    ("mo_width_nm", 200, 1500, int),
    ("mo_thick_nm0", 50, 100, int),
    ("mo_thick_nm1", 50, 100, int),
    ("mo_layer_num", 4, 10, int),
    ("mo_branch_nm", 0, 50, int),
    ("mo_left_gap_nm", 0, 50, int),    # lateral shelf offset (the left side's gap)
    ("mo_edge_rate", 0.0, 1.0, float),
    ("mo_curve", 0.0, 0.5, float),
    ("mo_n", 1.3, 1.8, float),
    ("mo_nx", 1.0, 1.3, float),
    ("mo_uniaxial", 0, 1, int),
    ("mo_asymmetry", 0, 1, int),
)
LAMBDA_RANGE = (380, 780)
N_FEATURES = len(GEOMETRY) + 1
N_ANGLES = 180
# generate(): the reduced grid every batched CW instance runs on: 4.4 um square
# at h = 10 nm with a 16-cell UPML, normal incidence along +y (angle 90);
# far-field angles 180..359 face back
GRID = dict(side_nm=4400, h_nm=10, pml=16, angle_deg=90, steps=2500)
# generate_sweep(): the large Morpho grid, 8192 x 4096 cells + a 32-cell UPML
# (8256 x 4160) at h = 5 nm, the plain pulse solver at normal incidence (+y),
# 20000 steps, the running DFT of the NTFF contour at 380..780 nm every 5 nm
SWEEP_GRID = dict(width_nm=41280, height_nm=20800, h_nm=5, pml=32, angle_deg=90, steps=20000)
SWEEP_LAMBDAS = tuple(range(380, 781, 5))


def _lib():
    L = N.lib()
    if L.fdtd_hip_device_count() < 1:
        raise RuntimeError("the surrogate runs on the GPU only (no HIP device visible)")
    return L


class DeviceArray:
    """A dense array in device memory of the library's own HIP runtime
    (fdtd_hip_alloc), uploaded from and downloaded to numpy."""

    def __init__(self, host=None, shape=None, dtype=np.float32, _base=None, _ptr=None):
        if host is not None:
            host = np.ascontiguousarray(host, dtype)
            shape = host.shape
        self.shape, self.dtype = tuple(int(v) for v in shape), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self._base = _base
        if _ptr is not None:
            self.ptr = _ptr
            return
        p = ctypes.c_void_p()
        check(_lib().fdtd_hip_alloc(self.nbytes, ctypes.byref(p)), "fdtd_hip_alloc")
        self.ptr = p
        if host is not None:
            check(N.lib().fdtd_hip_upload(self.ptr, host.ctypes.data_as(ctypes.c_void_p), self.nbytes), "upload")

    def rows(self, i0, i1):
        """rows [i0, i1) as a view (no copy)"""
        row = self.nbytes // self.shape[0]
        return DeviceArray(shape=(i1 - i0,) + self.shape[1:], dtype=self.dtype, _base=self,
                           _ptr=ctypes.c_void_p(self.ptr.value + i0 * row))

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        check(N.lib().fdtd_hip_download(out.ctypes.data_as(ctypes.c_void_p), self.ptr, self.nbytes), "download")
        return out

    def free(self):
        if self._base is None and getattr(self, "ptr", None) is not None:
            N.lib().fdtd_hip_free(self.ptr)
        self.ptr = None

    def __del__(self):
        self.free()


def _dev(x):
    return x if isinstance(x, DeviceArray) else DeviceArray(np.asarray(x, np.float32))


class MLP:
    """One network of the library (mlp_create).  Inputs and targets are FP32
    rows, [B, dims[0]] and [B, dims[-1]], as numpy arrays or DeviceArrays."""

    def __init__(self, dims, batch_max=1024, seed=0):
        L = _lib()
        self.dims = list(dims)
        self.batch_max = int(batch_max)
        arr = (ctypes.c_int * len(dims))(*dims)
        h = ctypes.c_void_p()
        check(L.mlp_create(len(dims) - 1, arr, self.batch_max, ctypes.c_ulonglong(seed), ctypes.byref(h)),
              "mlp_create")
        self._h = h

    def forward_into(self, x, y):
        """device to device: y[B, dims[-1]] = net(x[B, dims[0]])"""
        for i in range(0, x.shape[0], self.batch_max):
            i1 = min(i + self.batch_max, x.shape[0])
            check(N.lib().mlp_forward(self._h, x.rows(i, i1).ptr, i1 - i, y.rows(i, i1).ptr), "mlp_forward")

    def forward(self, x):
        x = _dev(x)
        y = DeviceArray(shape=(x.shape[0], self.dims[-1]))
        self.forward_into(x, y)
        return y.numpy()

    def grads(self, x, t):
        loss = ctypes.c_float()
        x, t = _dev(x), _dev(t)
        check(N.lib().mlp_grads(self._h, x.ptr, t.ptr, x.shape[0], ctypes.byref(loss)), "mlp_grads")
        return loss.value

    def train_step(self, x, t, lr):
        loss = ctypes.c_float()
        x, t = _dev(x), _dev(t)
        check(N.lib().mlp_train_step(self._h, x.ptr, t.ptr, x.shape[0], ctypes.c_float(lr), ctypes.byref(loss)),
              "mlp_train_step")
        return loss.value

    def train_epoch(self, x, t, batch, seed, lr):
        """one shuffled pass over device rows x, t; returns the mean batch loss"""
        loss = ctypes.c_float()
        check(N.lib().mlp_train_epoch(self._h, x.ptr, t.ptr, x.shape[0], int(batch), ctypes.c_ulonglong(seed),
                                      ctypes.c_float(lr), ctypes.byref(loss)), "mlp_train_epoch")
        return loss.value

    def _layer(self, fn, l):
        w = np.empty((self.dims[l + 1], self.dims[l]), np.float32)
        b = np.empty(self.dims[l + 1], np.float32)
        fp = ctypes.POINTER(ctypes.c_float)
        check(fn(self._h, l, w.ctypes.data_as(fp), b.ctypes.data_as(fp)), "mlp layer copy")
        return w, b

    def params(self, l):
        return self._layer(N.lib().mlp_get_params, l)

    def layer_grads(self, l):
        return self._layer(N.lib().mlp_get_grads, l)

    def set_params(self, l, w, b):
        w = np.ascontiguousarray(w, np.float32)
        b = np.ascontiguousarray(b, np.float32)
        fp = ctypes.POINTER(ctypes.c_float)
        check(N.lib().mlp_set_params(self._h, l, w.ctypes.data_as(fp), b.ctypes.data_as(fp)), "mlp_set_params")

    def close(self):
        if getattr(self, "_h", None):
            N.lib().mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class Ensemble:
    """k networks of the same shape from different seeds and data orders; the
    prediction is their mean and the disagreement their standard deviation."""

    def __init__(self, dims, k=4, batch=1024, seed=0):
        self.nets = [MLP(dims, batch, seed * 1000 + i) for i in range(k)]
        self.batch = batch
        self.seed = seed

    def fit(self, X, Y, epochs=200, lr=1e-3, lr_end=None, log=None):
        """Adam over `epochs` shuffled passes (one library call per pass and
        network), at a fixed rate lr, or cosine-annealed from lr to lr_end."""
        Xd, Yd = _dev(X), _dev(Y)
        hist = []
        for e in range(epochs):
            tot = 0.0
            r = lr if lr_end is None else lr_end + 0.5 * (lr - lr_end) * (1 + math.cos(math.pi * e / max(1, epochs - 1)))
            for i, net in enumerate(self.nets):
                tot += net.train_epoch(Xd, Yd, self.batch, self.seed * 1000003 + e * 131 + i, r)
            hist.append(tot / len(self.nets))
            if log and (e % max(1, epochs // 10) == 0 or e == epochs - 1):
                log(f"epoch {e}: mse {hist[-1]:.3e}")
        return hist

    def predict(self, X):
        """(mean, std) over the members, numpy float64 [B, dims[-1]]"""
        P = np.stack([net.forward(X).astype(np.float64) for net in self.nets])
        return P.mean(0), P.std(0)


def morpho_spec(seed=1, n=64):
    """Seeded morpho-scale geometries and wavelengths (integer nm): plain dicts."""
    g = SplitMix64(seed)
    out = []
    for _ in range(n):
        d = {}
        for name, lo, hi, kind in GEOMETRY:
            d[name] = g.integer(lo, hi) if kind is int else g.uniform(lo, hi)
        d["lambda_nm"] = g.integer(*LAMBDA_RANGE)
        out.append(d)
    return out


def morpho_geometries(seed=1, n=64):
    """Seeded morpho-scale geometries (the 12 parameters only): plain dicts."""
    g = SplitMix64(seed)
    out = []
    for _ in range(n):
        out.append({name: (g.integer(lo, hi) if kind is int else g.uniform(lo, hi)) for name, lo, hi, kind in GEOMETRY})
    return out


def features(spec):
    """[n, 13] in [0, 1]: the 12 parameters and the wavelength over their ranges."""
    X = np.empty((len(spec), N_FEATURES), np.float32)
    for r, d in enumerate(spec):
        for c, (name, lo, hi, _) in enumerate(GEOMETRY):
            X[r, c] = (d[name] - lo) / (hi - lo)
        X[r, -1] = (d["lambda_nm"] - LAMBDA_RANGE[0]) / (LAMBDA_RANGE[1] - LAMBDA_RANGE[0])
    return X


def reflectance(ff):
    """|E|^2 of the far field at the back-hemisphere angles 180..359 (incidence along +y)."""
    ff = np.asarray(ff)
    return (np.abs(ff[..., 180:360]) ** 2).astype(np.float64)


def generate(spec, T=8, R=16, chunk=128, log=None, grid=None):
    """Batched FDTD (MPI TM, morpho scale, CW at each wavelength) for every
    entry of spec on `grid` (default GRID): returns the features [n, 13] and
    reflectances [n, 180]."""
    from .batch import Batch
    G = grid or GRID
    w = G["side_nm"] + 2 * G["pml"] * G["h_nm"]
    Y = np.empty((len(spec), N_ANGLES))
    for s0 in range(0, len(spec), chunk):
        part = spec[s0:s0 + chunk]
        items = []
        for d in part:
            params = {k: d[k] for k, *_ in GEOMETRY}
            items.append(item(N.MPI_TM_UPML_2D, N.MORPHO_SCALE, w, w, G["h_nm"], G["pml"], d["lambda_nm"],
                              G["angle_deg"], G["steps"], mo_randomness=0, mo_seed=0, **params))
        with Batch(items, T=T, R=R) as b:
            b.run()
            for i in range(len(part)):
                Y[s0 + i] = reflectance(b.far_field(i))
        if log:
            log(f"generated {min(s0 + chunk, len(spec))}/{len(spec)}")
    return features(spec), Y


def generate_sweep(geoms, T=10, R=16, launch=8, log=None, grid=None):
    """The FDTD spectra of each geometry at the 81 wavelengths SWEEP_LAMBDAS
    from one run on the large grid (default SWEEP_GRID): the reference's plain
    TM pulse solver in FP32 with the running DFT of the NTFF contour, the
    far-field spectra |E|^2 / 8192 (ntffTM.c:197-233) at the exact wavelengths,
    back-hemisphere angles 180..359.  Returns the features [81 n, 13] and the
    spectra [81 n, 180], geometry-major."""
    from .simulator import Simulator, field_info, model_params
    G = grid or SWEEP_GRID
    lams = np.asarray(SWEEP_LAMBDAS, np.float64)
    spec, rows = [], []
    for k, d in enumerate(geoms):
        params = {name: d[name] for name, *_ in GEOMETRY}
        sim = Simulator(N.TM_UPML_2D, N.MORPHO_SCALE,
                        field_info(G["width_nm"], G["height_nm"], G["h_nm"], G["pml"], 550, G["angle_deg"], G["steps"]),
                        params=model_params(mo_randomness=0, mo_seed=0, **params), ntff_time=False, precision="fp32",
                        T=T, R=R)
        try:
            sim.set_launch_mode(launch)
            sim.set_dft(lams)
            sim.calc(G["steps"])
            S = sim.dft_spectra()
        finally:
            sim.finish()
        for f, lam in enumerate(lams):
            spec.append(dict(d, lambda_nm=int(lam)))
            rows.append(S[f, 180:360])
        if log and (k % 10 == 0 or k == len(geoms) - 1):
            log(f"swept {k + 1}/{len(geoms)} geometries")
    return features(spec), np.asarray(rows, np.float64)


def predict_with_fallback(ens, spec, scale, threshold=0.02, fdtd=None):
    """Ensemble prediction for each query; where the members' spread (std over
    the angles, relative to the mean's norm) exceeds `threshold`, the query is
    answered by FDTD instead.  `scale` is a Targets, or the float that linear
    targets were divided by.  Returns the reflectances [n, 180] and the mask of
    queries that went to FDTD."""
    tg = scale if isinstance(scale, Targets) else None
    X = features(spec)
    if tg is None:
        mean, std = ens.predict(X)
        mean, std = mean * scale, std * scale
    else:
        P = np.stack([tg.decode(net.forward(X)) for net in ens.nets])
        mean, std = P.mean(0), P.std(0)
    rel = np.linalg.norm(std, axis=1) / np.maximum(np.linalg.norm(mean, axis=1), 1e-300)
    use = rel > threshold
    out = mean.copy()
    if use.any():
        idx = np.nonzero(use)[0]
        _, Yf = (fdtd or generate)([spec[i] for i in idx])
        out[idx] = Yf
    return out, use


def rel_err(pred, ref):
    """Per-query relative L2 error over the angles."""
    return np.linalg.norm(pred - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-300)


def nearest_neighbour(Xtr, Ytr, Xq):
    """The method without a network: the training spectrum of the nearest
    geometry (Euclidean over the scaled features)."""
    d2 = (Xq ** 2).sum(1)[:, None] - 2 * Xq @ Xtr.T + (Xtr ** 2).sum(1)[None, :]
    return Ytr[np.argmin(d2, axis=1)]


class Targets:
    """The network's view of the reflectances: linear (R / scale) or log
    (log(R / scale + floor), which turns a relative error into an additive one)."""

    def __init__(self, Y, kind="linear"):
        self.kind = kind
        self.scale = float(np.mean(Y))
        self.floor = 1e-4

    def encode(self, Y):
        Z = Y / self.scale
        return (np.log(Z + self.floor) if self.kind == "log" else Z).astype(np.float32)

    def decode(self, Z):
        Z = np.asarray(Z, np.float64)
        return (np.exp(Z) - self.floor if self.kind == "log" else Z) * self.scale


def experiment(n_train=20000, n_test=2000, epochs=200, k=4, seed=1, threshold=0.02, log=print, save=None,
               load=None, lr=1e-3, lr_end=None, target="linear", source="sweep"):
    """Generate, train, evaluate: the numbers of the surrogate's acceptance
    (held-out error, fallback rate, error after fallback, throughput against
    the FDTD that made the data), as a dict.  source "sweep": the spectra of
    ceil(n / 81) seeded geometries at the 81 wavelengths each, from one pulse
    run per geometry on the large grid (generate_sweep; held-out geometries
    from another seed); "cw": n seeded (geometry, wavelength) pairs from
    batched CW runs on the reduced grid (generate).  `save` writes the spectra
    to an .npz file, `load` reads them from one instead of running the FDTD
    (its generation rate is then the one recorded in the file)."""
    import time
    L = _lib()
    if load:
        z = np.load(load)
        Xtr, Ytr, Xte, Yte = z["Xtr"], z["Ytr"], z["Xte"], z["Yte"]
        t_gen = float(z["t_gen"])
        n_train, n_test = len(Xtr), len(Xte)
    else:
        t0 = time.perf_counter()
        if source == "sweep":
            nl = len(SWEEP_LAMBDAS)
            Xtr, Ytr = generate_sweep(morpho_geometries(seed, -(-n_train // nl)), log=log)
            Xte, Yte = generate_sweep(morpho_geometries(seed + 1, -(-n_test // nl)), log=log)
            n_train, n_test = len(Xtr), len(Xte)
        else:
            Xtr, Ytr = generate(morpho_spec(seed=seed, n=n_train), log=log)
            Xte, Yte = generate(morpho_spec(seed=seed + 1, n=n_test), log=log)
        L.fdtd_hip_synchronize()
        t_gen = time.perf_counter() - t0
        if save:
            np.savez_compressed(save, Xtr=Xtr, Ytr=Ytr.astype(np.float32), Xte=Xte, Yte=Yte.astype(np.float32),
                                t_gen=t_gen)
    fdtd_rate = (n_train + n_test) / t_gen
    tg = Targets(Ytr, target)
    ens = Ensemble([N_FEATURES, 512, 512, 512, N_ANGLES], k=k, batch=1024, seed=seed)
    t0 = time.perf_counter()
    hist = ens.fit(Xtr, tg.encode(Ytr), epochs=epochs, lr=lr, lr_end=lr_end, log=log)
    t_train = time.perf_counter() - t0
    P = np.stack([tg.decode(net.forward(Xte)) for net in ens.nets])  # members in reflectance units
    pred = P.mean(0)
    spread = np.linalg.norm(P.std(0), axis=1) / np.maximum(np.linalg.norm(pred, axis=1), 1e-300)
    err = rel_err(pred, Yte)
    use = spread > threshold
    err_fb = np.where(use, 0.0, err)  # the flagged queries are answered by the FDTD itself
    nn = rel_err(nearest_neighbour(Xtr, Ytr, Xte), Yte)
    # inference throughput: 64 Ki queries through the whole ensemble, device to device
    rng = np.random.default_rng(seed)
    Xb = DeviceArray(rng.random((65536, N_FEATURES)).astype(np.float32))
    outs = [DeviceArray(shape=(65536, N_ANGLES)) for _ in ens.nets]
    for net, o in zip(ens.nets, outs):
        net.forward_into(Xb, o)
    L.fdtd_hip_synchronize()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        for net, o in zip(ens.nets, outs):
            net.forward_into(Xb, o)
    L.fdtd_hip_synchronize()
    sur_rate = reps * 65536 / (time.perf_counter() - t0)
    return {
        "n_train": n_train, "n_test": n_test, "epochs": epochs, "ensemble": k, "lr": lr, "lr_end": lr_end,
        "target": target, "source": source,
        "grid": SWEEP_GRID if source == "sweep" else GRID, "fdtd_spectra_per_s": fdtd_rate, "fdtd_seconds": t_gen,
        "train_seconds": t_train, "final_train_mse": hist[-1],
        "heldout_rel_err_mean": float(err.mean()), "heldout_rel_err_median": float(np.median(err)),
        "heldout_rel_err_p95": float(np.quantile(err, 0.95)),
        "heldout_frac_within_1pct": float((err <= 0.01).mean()),
        "fallback_threshold": threshold, "fallback_frac": float(use.mean()),
        "rel_err_after_fallback_mean": float(err_fb.mean()),
        "frac_within_1pct_after_fallback": float((err_fb <= 0.01).mean()),
        "nearest_neighbour_rel_err_mean": float(nn.mean()),
        "nearest_neighbour_rel_err_median": float(np.median(nn)),
        "surrogate_spectra_per_s": sur_rate, "speedup_vs_fdtd": sur_rate / fdtd_rate,
    }


def data_scaling(load, sizes=(2500, 5000, 10000, 20000), steps=4000, k=4, seed=1, lr=1e-3, lr_end=1e-5, log=print):
    """Held-out error against the number of training spectra: the ensemble is
    trained on nested prefixes of the training set of `load` (an .npz written by
    experiment(save=...)) for about the same number of Adam steps each (epochs
    = steps * batch / n, cosine-annealed), and evaluated on the same held-out
    set.  Returns [{"n", "epochs", "median", "mean", "p95", "train_mse"}]."""
    z = np.load(load)
    Xtr, Ytr, Xte, Yte = z["Xtr"], z["Ytr"].astype(np.float64), z["Xte"], z["Yte"].astype(np.float64)
    out = []
    for n in sizes:
        n = min(int(n), len(Xtr))
        tg = Targets(Ytr[:n])
        ens = Ensemble([N_FEATURES, 512, 512, 512, N_ANGLES], k=k, batch=1024, seed=seed)
        epochs = max(1, int(round(steps * 1024 / n)))
        hist = ens.fit(Xtr[:n], tg.encode(Ytr[:n]), epochs=epochs, lr=lr, lr_end=lr_end)
        pred = np.stack([tg.decode(net.forward(Xte)) for net in ens.nets]).mean(0)
        err = rel_err(pred, Yte)
        row = {"n": n, "epochs": epochs, "median": float(np.median(err)), "mean": float(err.mean()),
               "p95": float(np.quantile(err, 0.95)), "train_mse": hist[-1]}
        if log:
            log(row)
        out.append(row)
    return out


if __name__ == "__main__":
    import argparse
    import json
    ap = argparse.ArgumentParser(description="reflectance surrogate: generate, train, evaluate")
    ap.add_argument("--train", type=int, default=20000)
    ap.add_argument("--test", type=int, default=2000)
    ap.add_argument("--epochs", type=int, default=200)
    ap.add_argument("--k", type=int, default=4)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-end", type=float, default=None, help="cosine-anneal the rate to this value")
    ap.add_argument("--target", default="linear", choices=("linear", "log"))
    ap.add_argument("--source", default="sweep", choices=("sweep", "cw"),
                    help="sweep: 81 wavelengths per geometry on the 8256 x 4160 grid; cw: batched CW runs on the "
                         "reduced grid")
    ap.add_argument("--save", default=None, help="write the generated spectra to this .npz")
    ap.add_argument("--load", default=None, help="read the spectra from this .npz instead of running the FDTD")
    ap.add_argument("--subsets", default=None,
                    help="comma-separated training-set sizes: held-out error against the number of spectra "
                         "(needs --load)")
    ap.add_argument("--steps", type=int, default=4000, help="--subsets: Adam steps per size")
    a = ap.parse_args()
    if a.subsets:
        if not a.load:
            ap.error("--subsets needs --load")
        rows = data_scaling(a.load, [int(x) for x in a.subsets.split(",")], steps=a.steps, k=a.k, seed=a.seed,
                            lr=a.lr, lr_end=a.lr_end if a.lr_end is not None else 1e-5,
                            log=lambda r: print(json.dumps(r), flush=True))
        ns = np.log([r["n"] for r in rows])
        slope = float(np.polyfit(ns, np.log([r["median"] for r in rows]), 1)[0]) if len(rows) > 1 else None
        print(json.dumps({"data_scaling": rows, "median_exponent": slope}), flush=True)
        raise SystemExit(0)
    res = experiment(a.train, a.test, a.epochs, a.k, a.seed, log=lambda s: print(s, flush=True), save=a.save,
                     load=a.load, lr=a.lr, lr_end=a.lr_end, target=a.target, source=a.source)
    print(json.dumps(res), flush=True)
