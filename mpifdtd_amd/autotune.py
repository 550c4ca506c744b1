"""Autotuner of the FP32 fused kernels: temporal block depth T and rows per wavefront R.

For a grid class (mode, flavour, grid size rounded to 256 cells) it times a
short run of each candidate (T, R) on a live simulation and keeps the fastest.
Block depth and rows per wavefront change only how the steps are launched,
not the arithmetic of any cell, so tuning on the live state is safe.  Results
are cached per process and, when MPIFDTD_TUNE_CACHE names a file, on disk.
"""
import json
import os
import time

CANDIDATES = [(1, 16), (2, 16), (3, 16), (4, 16), (5, 16), (6, 16), (8, 16), (4, 8), (6, 8), (8, 8), (10, 8), (4, 4)]

_cache = {}


def _key(sim):
    return f"{sim.solver}:{sim.N_PX // 256}x{sim.N_PY // 256}:{'ntff' if sim.per_step_ntff else 'plain'}"


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
    try:
        sim.calc(40)
        for T, R in candidates:
            sim.set_blocking(T, R)
            sim.calc(2 * T)  # warm the variant
            sim.set_blocking(T, R)
            L.fdtd_hip_synchronize()
            t0 = time.perf_counter()
            sim.calc(probe_steps)
            sim.set_blocking(T, R)  # flush the partial block
            L.fdtd_hip_synchronize()
            timings[f"T{T}R{R}"] = (time.perf_counter() - t0) * 1e6 / probe_steps
    finally:
        sim.finish()
    best = min(timings, key=timings.get)
    T, R = best[1:].split("R")
    res = {"T": int(T), "R": int(R), "us_per_step": {k: round(v, 2) for k, v in timings.items()}}
    _cache[key] = res
    _save()
    del L
    return res
