"""Benchmark: Mie-cylinder scattering (TM), 1024x1024 cells + 16-cell UPML at h = 5 nm.

One step = one FDTD time step of the whole 1056 x 1056 grid (E and H halves,
UPML, scattered-wave source).  The timed region is K steps of the FP32 fused,
temporally blocked kernels through the reference's own driver API
(simulator_calc via libmpifdtd_amd.so); inputs are resident in HBM.  The far
field (360 angles) of the final fields is formed in FP64 and its error is
reported against the FP64 reference-order GPU path, which the test suite pins
to the CPU oracle at 1e-10.

    python bench.py [--steps K] [--warmup W] [--precision fp32|fp64] [--mode tm|te]
prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Mie cylinder (circleModel.c): r = 500 nm, n = 1.6, lambda = 550 nm, +x incidence
W_NM, H_NM, HU, PML, LAMBDA, ANGLE = 5280, 5280, 5, 16, 550, 0
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8000)
    ap.add_argument("--warmup", type=int, default=400)
    ap.add_argument("--precision", default="fp32", choices=("fp32", "fp64"))
    ap.add_argument("--mode", default="tm", choices=("tm", "te"))
    ap.add_argument("--T", type=int, default=0, help="temporal block depth (0: autotune)")
    ap.add_argument("--R", type=int, default=0, help="rows per wavefront (0: autotune)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 far-field check")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (5 times): the reference kernel "
                         "that scales rocprofv3's FETCH_SIZE / WRITE_SIZE counters to bytes")
    return ap.parse_args()


def copy_calibration(mib):
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 5}), flush=True)


def make_sim(m, precision, mode, steps, T=4, R=16):
    solver = m.MPI_TM_UPML_2D if mode == "tm" else m.MPI_TE_UPML_2D
    fi = m.field_info(W_NM, H_NM, HU, PML, LAMBDA, ANGLE, steps)
    params = m.model_params(circle_radius_nm=500, circle_n=1.6)
    return m.Simulator(solver, m.MIE_CYLINDER, fi, params=params, ntff_time=False, precision=precision, T=T, R=R)


def prof_read(L):
    import ctypes
    out = (ctypes.c_double * 6)()
    L.fdtd_hip_profile_read(out)
    return list(out)


def cpu_baseline(mode):
    """The FP64 CPU oracle (the reference's MPI-TM/TE step restated in C, gcc -O2,
    one core) on a bounded sample of the same workload, including the per-step
    time-domain NTFF that every reference step performs."""
    from oracle import pyoracle as po
    solver = po.MPI_TM_UPML if mode == "tm" else po.MPI_TE_UPML
    o = po.Sim(width_nm=W_NM, height_nm=H_NM, h_u_nm=HU, pml=PML, lambda_nm=LAMBDA, angle_deg=ANGLE, stepNum=8000,
               solver=solver, model=po.MIE_CYLINDER, ntff_time=1)
    o.step(2)
    n = 120
    t0 = time.perf_counter()
    o.step(n)
    dt = time.perf_counter() - t0
    cells = o.N_PX * o.N_PY
    return {"value": round(cells * n / dt / 1e6, 3), "unit": "Mcell-updates/s", "cores": 1, "kind": "port",
            "sample": f"{n} steps of the same 1056x1056 Mie {mode.upper()} grid, FP64, with the per-step NTFF "
                      f"({dt:.1f} s on one host core; gcc -O2 oracle)"}


def measured_traffic(mode, T, R, cells):
    """HBM bytes per launch of this configuration from the committed rocprofv3
    counter summaries (FETCH_SIZE / WRITE_SIZE scaled by a plain copy), or None."""
    import glob
    import re
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*traffic*.md"))):
        for block in re.findall(r"```json\n(.*?)\n```", open(path).read(), re.S):
            try:
                rec = json.loads(block)
            except ValueError:
                continue
            if (rec.get("mode"), rec.get("T"), rec.get("R"), rec.get("cells")) == (mode, T, R, cells):
                return rec["traffic_bytes_per_launch"], os.path.relpath(path, ROOT)
    return None, None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    # the C host prints progress messages like the reference does (printf on stdout);
    # keep stdout for the one JSON line by sending fd 1 to stderr until the end
    import ctypes
    libc = ctypes.CDLL(None)
    real_stdout = os.dup(1)
    os.dup2(2, 1)
    if args.gpus != 1:
        sys.exit(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU sharding is "
                 "described in DESIGN.md and not built")
    import numpy as np
    import torch

    import mpifdtd_amd as m
    from mpifdtd_amd import autotune
    L = m.lib()
    assert torch.cuda.is_available(), "no GPU"
    torch.cuda.set_device(0)
    K, W = args.steps, args.warmup
    total = K + W

    T, R = args.T, args.R
    tuned = None
    if args.precision == "fp32" and (T == 0 or R == 0):
        log("autotune")
        tuned = autotune.tune(m, lambda: make_sim(m, "fp32", args.mode, 2000), probe_steps=240)
        T = T or tuned["T"]
        R = R or tuned["R"]
    log(f"timed run: {args.precision} T={T} R={R} warmup={W} steps={K}")
    sim = make_sim(m, args.precision, args.mode, total, T or 4, R or 16)
    cells = sim.N_PX * sim.N_PY
    sim.calc(W)
    ez_w = sim.fields()[0]  # flushes queued steps
    torch.cuda.synchronize()
    L.fdtd_hip_profile_enable(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.calc(K)
    if args.precision == "fp32":
        L.fdtd_solver_setBlocking(T, R)  # launches the last partial block inside the timed region
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    prof = prof_read(L)
    L.fdtd_hip_profile_enable(0)
    t1 = time.perf_counter()
    ff = sim.far_field()
    t_ff = time.perf_counter() - t1
    sim.finish()
    ms_step = dt * 1e3 / K
    mcups = cells * K / dt / 1e6

    # roofline of the dominant kernel: the fused launch (UPML + interior tiles, T steps)
    ms_k, _, nlaunch, b_int, b_pml, cupd = prof
    if args.precision == "fp32" and nlaunch > 0 and ms_k > 0:
        achieved = (b_int + b_pml) / (ms_k * 1e-3) / 1e9
        traffic, tsrc = measured_traffic(args.mode, T, R, cells)
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": tsrc,
                "kernel": f"k_fast<{args.mode.upper()},MPI,R={R}>",
                "launch_avg_us": round(ms_k * 1e3 / nlaunch, 2), "steps_per_launch": T,
                "bytes_per_launch": round((b_int + b_pml) / nlaunch),
                "bytes_per_cell_update": round((b_int + b_pml) / cupd, 2)}
    else:
        roof = {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None, "traffic": None}

    extra = {}
    # spectra/s: one spectrum = a full 8000-step run plus its far field
    spec_steps = 8000
    extra["spectra_per_s"] = round(1.0 / (spec_steps * ms_step * 1e-3 + t_ff), 3)
    if not args.no_check and args.precision == "fp32":
        # FP32 far field vs the FP64 reference-order GPU path on the same run length
        log("fp64 reference-order run for the far-field check")
        ref = make_sim(m, "fp64", args.mode, total)
        t2 = time.perf_counter()
        ref.calc(total)
        ff64 = ref.far_field()
        t64 = time.perf_counter() - t2
        ref.finish()
        extra["ff_rel_err_vs_fp64"] = float(np.linalg.norm(ff - ff64) / np.linalg.norm(ff64))
        extra["fp64_mcups"] = round(cells * total / t64 / 1e6, 1)
        extra["fp64_spectra_per_s"] = round(1.0 / (spec_steps * t64 / total), 3)
        extra["speedup_vs_fp64_spectra"] = round(extra["spectra_per_s"] / extra["fp64_spectra_per_s"], 2)
    if not args.no_cpu:
        log("cpu baseline (oracle, one core)")
    cpu = None if args.no_cpu else cpu_baseline(args.mode)
    line = {
        "metric": "Mcell-updates/s + % HBM roofline; spectra/s at <=1e-4 far-field err",
        "value": round(mcups, 1), "unit": "Mcell-updates/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": round(ms_step, 5), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f64", "data": "synthetic",
        "config": {"workload": f"Mie cylinder {args.mode.upper()}, 1024x1024 + 16-cell UPML (1056x1056), h=5 nm, "
                               f"r=500 nm, n=1.6, lambda=550 nm", "cells": cells, "precision": args.precision,
                   "T": T, "R": R, "autotune": tuned},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    line.update(extra)
    libc.fflush(None)
    sys.stdout.flush()
    os.dup2(real_stdout, 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
