"""Autotuner of the FP32 fused kernels: temporal block depth T, rows per
wavefront R and the launch shape of the UPML and interior tiles.

For a grid class (mode, flavour, grid size rounded to 256 cells) it times a
short run of each candidate (T, R) in each launch shape (merged, split, split
on two streams; the last two also with 8-wavefront UPML workgroups, and with
8-wavefront interior workgroups as well) on a live simulation and keeps the
fastest.  None of these changes the arithmetic of any cell, so tuning on
the live state is safe.  Results
are cached per process and, when MPIFDTD_TUNE_CACHE names a file, on disk.
"""
import json
import os
import time

CANDIDATES = [(1, 16), (2, 16), (3, 16), (4, 16), (5, 16), (6, 16), (8, 16), (10, 16), (10, 18), (8, 20), (10, 20), (4, 8),
              (6, 8), (8, 8), (10, 8), (4, 4)]

_cache = {}


def _key(sim):
    return f"{sim.solver}:{sim.N_PX // 256}x{sim.N_PY // 256}:{'ntff' if sim.per_step_ntff else 'plain'}:v11"


def _load():
    path = os.environ.get("MPIFDTD_TUNE_CACHE")
    if path and os.path.exists(path):
        try:
            _cache.update(json.load(open(path)))
        except (OSError, ValueError):
            pass


def _save():
    path = os.environ.get("MPIFDTD_TUNE_CACHE")
    if path:
        try:
            json.dump(_cache, open(path, "w"))
        except OSError:
            pass


def tune(m, make_sim, probe_steps=240, candidates=CANDIDATES):
    """make_sim() -> a fresh FP32 Simulator; returns {"T", "R", "us_per_step": {...}}."""
    _load()
    sim = make_sim()
    key = _key(sim)
    if key in _cache:
        sim.finish()
        return _cache[key]
    L = m.lib()
    timings = {}
    shapes = {}

    def probe(name, T, R, shape):
        n = T * max(1, round(probe_steps / T))  # whole blocks only: no partial block in the timing
        sim.set_blocking(T, R)
        sim.set_launch_mode(shape)
        sim.calc(2 * T)  # warm the variant
        sim.set_blocking(T, R)
        L.fdtd_hip_synchronize()
        t0 = time.perf_counter()
        sim.calc(n)
        sim.set_blocking(T, R)  # flushes nothing: n is a multiple of T
        L.fdtd_hip_synchronize()
        timings[name] = (time.perf_counter() - t0) * 1e6 / n
        shapes[name] = (T, R, shape)

    try:
        sim.calc(40)
        shapes_tried = ((1, "merged"), (2, "split"), (3, "2streams"), (4, "splitu48"), (5, "2streamsu48"),
                        (6, "splitw8"), (7, "2streamsw8"), (8, "splitp16"), (9, "2streamsp16"))
        for T, R in candidates:
            for shape, tag in shapes_tried:
                if shape in (6, 7) and R < 8:
                    continue  # R = 4 interior tiles keep 4 wavefronts: the same as shapes 4 / 5
                if R in (18, 20) and shape not in (6, 7, 8, 9):
                    continue  # 72- and 80-row interior tiles exist as 8 wavefronts only
                probe(f"T{T}R{R}{tag}", T, R, shape)
        sim.set_launch_mode(0)
    finally:
        sim.finish()
    best = min(timings, key=timings.get)
    T, R, shape = shapes[best]
    res = {"T": T, "R": R, "launch": shape, "us_per_step": {k: round(v, 2) for k, v in timings.items()}}
    _cache[key] = res
    _save()
    del L
    return res
