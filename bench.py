"""Benchmark of the FP32 fused UPML step through the reference's driver API.

Default workload: Mie-cylinder scattering (circleModel.c, r = 500 nm, n = 1.6),
1024 x 1024 cells + 16-cell UPML (1056 x 1056) at h = 5 nm, lambda = 550 nm,
MPI TM solver, 8000 steps.  One step = one FDTD time step of the whole grid (E
and H halves, UPML, scattered-wave source).  The timed region is K steps of
simulator_calc through libmpifdtd_amd.so, whose FP32 path advances T steps per
launch; inputs are resident in HBM.  The far field (360 angles) of the final
fields is formed in FP64 and compared with the FP64 reference-order GPU path,
which the test suite pins to the CPU oracle at 1e-10.

--workload morpho8256 runs the large HBM-bound case instead: the Morpho-scale
model on 8256 x 4160 cells (h = 5 nm, 32-cell UPML), 34.3 M cells, whose FP32
state (1.6 GB) is far beyond the 256 MB Infinity Cache.

--workload sweep64 runs a batched sweep: 64 seeded instances (multilayer or
zigzag, sides 512 ... 2048 cells + 16-cell UPML, TM or TE, lambda 380-780 nm,
6000-10000 steps each) advanced together to completion; value is the aggregate
cell-update rate and spectra_per_s the completed instances per second.

    python bench.py [--steps K] [--warmup W] [--workload mie1056|morpho8256] [--mode tm|te]
prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# launch shapes of the UPML and interior tiles (FAST_LAUNCH_* of fdtd_hip.h)
SHAPES = ["auto", "merged", "split", "two_streams", "split_pml8", "two_streams_pml8", "split_w8", "two_streams_w8",
          "split_p16", "two_streams_p16"]

# name -> (model, width_nm, height_nm, h_nm, pml, lambda_nm, angle_deg, model params, description)
WORKLOADS = {
    "mie1056": ("MIE_CYLINDER", 5280, 5280, 5, 16, 550, 0, {"circle_radius_nm": 500, "circle_n": 1.6},
                "Mie cylinder {mode}, 1024x1024 + 16-cell UPML (1056x1056), h=5 nm, r=500 nm, n=1.6, lambda=550 nm"),
    "morpho8256": ("MORPHO_SCALE", 41280, 20800, 5, 32, 550, 90, {},
                   "Morpho scale {mode}, 8192x4096 + 32-cell UPML (8256x4160), h=5 nm, lambda=550 nm, "
                   "normal incidence (+y)"),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="timed steps (default 8000; 400 for morpho8256)")
    ap.add_argument("--warmup", type=int, default=-1, help="untimed steps (default 400; 40 for morpho8256)")
    ap.add_argument("--workload", default="mie1056",
                    choices=sorted(WORKLOADS) + ["sweep64", "morpho8256_sweep81", "solvers_fp64"])
    ap.add_argument("--seed", type=int, default=1, help="sweep64: generator seed")
    ap.add_argument("--precision", default="fp32", choices=("fp32", "fp64"))
    ap.add_argument("--mode", default="tm", choices=("tm", "te"))
    ap.add_argument("--T", type=int, default=0, help="temporal block depth (0: autotune)")
    ap.add_argument("--R", type=int, default=0, help="rows per wavefront (0: autotune)")
    ap.add_argument("--launch", type=int, default=-1,
                    help="launch shape of the UPML and interior tiles: 0 auto, 1 merged, 2 split, "
                         "3 split on two streams, 4 / 5: 2 / 3 with 8-wavefront UPML workgroups, "
                         "6 / 7: 4 / 5 with 8-wavefront interior workgroups "
                         "(-1: autotune's choice, or auto)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--ntff-overlap", type=int, default=-1,
                    help="contour sample work beside the next block (1), after it (0), by grid size (-1)")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 far-field check")
    ap.add_argument("--ntff", action="store_true",
                    help="run the reference's per-step time-domain NTFF inside the timed steps (the MPI solvers "
                         "do it every step; default off, the far field then comes from the final fields)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (5 times): the reference kernel "
                         "that scales rocprofv3's FETCH_SIZE / WRITE_SIZE counters to bytes")
    a = ap.parse_args()
    big = a.workload != "mie1056"
    a.steps_given = a.steps > 0
    if a.steps <= 0:
        a.steps = 400 if big else 8000
    if a.warmup < 0:
        a.warmup = 40 if big else 400
    return a


def copy_calibration(mib):
    # a plain device copy of known size, for scaling the PMC byte counters (its
    # own process: torch's runtime only, the library is not loaded)
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 5}), flush=True)


def make_sim(m, precision, mode, steps, T=4, R=16, workload="mie1056", ntff=False):
    model, w, h, hu, pml, lam, ang, params, _ = WORKLOADS[workload]
    solver = m.MPI_TM_UPML_2D if mode == "tm" else m.MPI_TE_UPML_2D
    fi = m.field_info(w, h, hu, pml, lam, ang, steps)
    return m.Simulator(solver, getattr(m, model), fi, params=m.model_params(**params), ntff_time=ntff,
                       precision=precision, T=T, R=R)


def prof_read(L):
    import ctypes
    out = (ctypes.c_double * 6)()
    L.fdtd_hip_profile_read(out)
    return list(out)


def cpu_reference(mode, workload, steps=120):
    """The reference itself on a bounded sample of the same workload: its MPI-TM /
    TE UPML solver compiled from its own sources (oracle/_ref/ref_harness, one MPI
    rank, one core), `steps` simulator_calc() calls including the per-step NTFF,
    timed inside the harness.  Runs as a child process before this process touches
    the GPU.  None when the harness is not built or does not run."""
    import subprocess
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if not os.access(exe, os.X_OK):
        return None
    model, w, h, hu, pml, lam, ang, params, _ = WORKLOADS[workload]
    if model != "MIE_CYLINDER" or params != {"circle_radius_nm": 500, "circle_n": 1.6}:
        return None  # the reference's cylinder is r = 500 nm, n = 1.6 (circleModel.c:6-10)
    cmd = [exe, "mode=time", "solver=%d" % (4 if mode == "tm" else 5), "model=1", f"width={w}", f"height={h}",
           f"h={hu}", f"pml={pml}", f"lambda={lam}", f"angle={ang}", f"steps={steps}", "warm=2"]
    try:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=240)
        res = json.loads(r.stdout.strip().splitlines()[-1])
    except (OSError, ValueError, IndexError, subprocess.SubprocessError):
        return None
    cells = res["N_PX"] * res["N_PY"]
    return {"value": round(cells * res["steps"] / res["seconds"] / 1e6, 3), "unit": "Mcell-updates/s", "cores": 1,
            "kind": "reference",
            "sample": f"{res['steps']} steps of the same {res['N_PX']}x{res['N_PY']} {workload} {mode.upper()} grid, "
                      f"the reference's own MPI {mode.upper()} UPML solver (FP64, per-step NTFF) built from its "
                      f"sources with gcc -O2, one MPI rank ({res['seconds']:.1f} s on one host core)"}


def cpu_baseline(mode, workload):
    """The FP64 CPU oracle (the reference's MPI-TM/TE step restated in C, gcc -O2,
    one core) on a bounded sample of the same workload, including the per-step
    time-domain NTFF that every reference step performs."""
    from oracle import pyoracle as po
    model, w, h, hu, pml, lam, ang, params, _ = WORKLOADS[workload]
    solver = po.MPI_TM_UPML if mode == "tm" else po.MPI_TE_UPML
    o = po.Sim(width_nm=w, height_nm=h, h_u_nm=hu, pml=pml, lambda_nm=lam, angle_deg=ang, stepNum=8000,
               solver=solver, model=getattr(po, model), ntff_time=1, **params)
    o.step(2)
    cells = o.N_PX * o.N_PY
    n = max(4, int(130e6 // cells))  # about 5 s of one core
    t0 = time.perf_counter()
    o.step(n)
    dt = time.perf_counter() - t0
    return {"value": round(cells * n / dt / 1e6, 3), "unit": "Mcell-updates/s", "cores": 1, "kind": "port",
            "sample": f"{n} steps of the same {o.N_PX}x{o.N_PY} {workload} {mode.upper()} grid, FP64, with the "
                      f"per-step NTFF ({dt:.1f} s on one host core; gcc -O2 oracle)"}


def measured_traffic(mode, T, R, cells, launch):
    """HBM bytes per launch of this configuration (launch shape included) from
    the committed rocprofv3 counter summaries (FETCH_SIZE / WRITE_SIZE scaled by
    a plain copy), or None."""
    import glob
    import re
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*traffic*.md"))):
        for block in re.findall(r"```json\n(.*?)\n```", open(path).read(), re.S):
            try:
                rec = json.loads(block)
            except ValueError:
                continue
            if (rec.get("mode"), rec.get("T"), rec.get("R"), rec.get("cells"), rec.get("launch")) == \
                    (mode, T, R, cells, launch):
                return rec["traffic_bytes_per_launch"], os.path.relpath(path, ROOT)
    return None, None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_sweep81(args, m):
    """Config 4's wavelength sweep (380-780 nm every 5 nm: 81 wavelengths) on the
    8256 x 4160 Morpho-scale grid as one run: the reference's plain pulse solver
    (TM_UPML_2D / TE_UPML_2D, a Gaussian plane-wave pulse at normal incidence
    whose band covers the sweep) with the running DFT of the NTFF contour at the
    81 wavelengths.  Each wavelength's far-field spectrum (the reference's
    |Eth|^2 / 8192 or |Eph|^2 / 8192, ntffTM.c:197-233) comes from the DFT, the
    back-hemisphere angles 180-359 being its reflectance."""
    import numpy as np
    L = m.lib()
    model, w, h, hu, pml, lam, ang, params, _ = WORKLOADS["morpho8256"]
    lams = np.arange(380.0, 781.0, 5.0)
    steps = args.steps if args.steps_given else 20000
    solver = m.TM_UPML_2D if args.mode == "tm" else m.TE_UPML_2D
    T, R = args.T or 8, args.R or 16
    launch = args.launch if args.launch >= 0 else 3

    def make(precision, n):
        return m.Simulator(solver, getattr(m, model), m.field_info(w, h, hu, pml, lam, ang, n),
                           params=m.model_params(**params), ntff_time=False, precision=precision, T=T, R=R)

    tune = {}
    if not (args.T and args.R and args.launch >= 0):
        # a short tune of depth, rows and launch shape on the same grid with the DFT on
        probe = make("fp32", 4000)
        probe.set_dft(lams)
        probe.calc(40)
        for Tc, Rc in ((8, 16), (10, 16), (8, 8), (6, 8)):
            for sh in (2, 3, 4, 5, 6, 7, 8, 9):
                n = 6 * Tc  # whole blocks only
                probe.set_blocking(Tc, Rc)
                probe.set_launch_mode(sh)
                probe.calc(2 * Tc)
                probe.set_blocking(Tc, Rc)
                L.fdtd_hip_synchronize()
                t0 = time.perf_counter()
                probe.calc(n)
                probe.set_blocking(Tc, Rc)
                L.fdtd_hip_synchronize()
                tune[(Tc, Rc, sh)] = (time.perf_counter() - t0) * 1e6 / n
        probe.finish()
        T, R, launch = min(tune, key=tune.get)
    log(f"sweep81: {len(lams)} wavelengths, {steps} steps, T={T} R={R} launch={launch}")
    sim = make("fp32", steps)
    cells = sim.N_PX * sim.N_PY
    sim.set_launch_mode(launch)
    sim.set_ntff_overlap(args.ntff_overlap)
    sim.set_dft(lams)
    L.fdtd_hip_synchronize()
    L.fdtd_hip_profile_enable(1)
    t0 = time.perf_counter()
    sim.calc(steps)
    L.fdtd_solver_setBlocking(T, R)  # launches the last partial block
    L.fdtd_hip_synchronize()
    dt = time.perf_counter() - t0
    prof = prof_read(L)
    L.fdtd_hip_profile_enable(0)
    t1 = time.perf_counter()
    spec = sim.dft_spectra()
    t_spec = time.perf_counter() - t1
    shape = SHAPES[sim.launch_mode()]
    sim.finish()
    assert np.all(np.isfinite(spec)) and spec[:, 180:360].max() > 0
    check = {}
    if not args.no_check:
        # FP32 against FP64 on a shorter run of the same grid and source: the 81
        # transforms of the three accumulator rows, and the FP64 path's step rate
        nchk = 4000  # the scattered pulse reaches the nearest contour side after about 3000 steps
        log(f"fp32 / fp64 check over {nchk} steps")
        res = {}
        for prec in ("fp32", "fp64"):
            c = make(prec, nchk)
            if prec == "fp32":
                c.set_launch_mode(launch)
            c.set_dft(lams)
            L.fdtd_hip_synchronize()
            t2 = time.perf_counter()
            c.calc(nchk)
            if prec == "fp32":
                L.fdtd_solver_setBlocking(T, R)
            L.fdtd_hip_synchronize()
            res[prec] = (c.dft(), time.perf_counter() - t2)
            c.finish()
        d32, d64 = res["fp32"][0], res["fp64"][0]
        assert np.linalg.norm(d64) > 0, "the check run ended before the pulse reached the contour"
        check["dft_rel_err_vs_fp64"] = float(np.linalg.norm(d32 - d64) / np.linalg.norm(d64))
        ms64 = res["fp64"][1] * 1e3 / nchk
        check["fp64_mcups"] = round(cells / (ms64 * 1e-3) / 1e6, 1)
        check["fp64_spectra_per_s"] = round(len(lams) / (steps * ms64 * 1e-3), 4)
    ms_k, _, nlaunch, b_int, b_pml, cupd = prof
    achieved = (b_int + b_pml) / (ms_k * 1e-3) / 1e9 if nlaunch and ms_k > 0 else None
    line = {
        "metric": "Mcell-updates/s + % HBM roofline; spectra/s at <=1e-4 far-field err",
        "value": round(cells * steps / dt / 1e6, 1), "unit": "Mcell-updates/s", "n_gpus": 1, "steps": steps,
        "warmup": 0, "ms_per_step": round(dt * 1e3 / steps, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"Morpho scale {args.mode.upper()}, 8192x4096 + 32-cell UPML (8256x4160), h=5 nm, "
                               "plain pulse solver, normal incidence (+y), running DFT at 380-780 nm every 5 nm "
                               "(81 wavelengths) in one run",
                   "cells": cells, "precision": "fp32", "T": T, "R": R, "launch": shape, "wavelengths": len(lams),
                   "tune_us_per_step": {f"T{k[0]}R{k[1]}{SHAPES[k[2]]}": round(v, 2) for k, v in tune.items()}},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1) if achieved else None, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                     "traffic": None, "kernel": f"k_fast<{args.mode.upper()},PLAIN,R={R}> with contour samples",
                     "launch_avg_us": round(ms_k * 1e3 / nlaunch, 2) if nlaunch else None,
                     "bytes_per_cell_update": round((b_int + b_pml) / cupd, 2) if cupd else None},
        "cpu_baseline": None,
        "spectra_per_s": round(len(lams) / (dt + t_spec), 4),
    }
    if check:
        line.update(check)
        line["speedup_vs_fp64_spectra"] = round(line["spectra_per_s"] / check["fp64_spectra_per_s"], 2)
    return line


def run_solvers_fp64(args, m):
    """Every solver of enum SOLVER in its FP64 reference-order path on the mie1056
    grid (the reference's cylinder; CW source for the MPI, split-field and NS
    solvers, Gaussian pulse for the plain UPML ones), without the per-step NTFF:
    the step rate of each over `steps` simulator_calc() calls after 20 untimed
    ones.  Not the headline: a measurement of the solvers beside the UPML path,
    whose kernel durations a rocprofv3 kernel trace of this command gives."""
    import numpy as np
    L = m.lib()
    model, w, h, hu, pml, lam, ang, params, _ = WORKLOADS["mie1056"]
    steps = args.steps if args.steps_given else 400
    solvers = [("TM_2D", m.TM_2D), ("TE_2D", m.TE_2D), ("NS_TM_2D", m.NS_TM_2D), ("NS_TE_2D", m.NS_TE_2D),
               ("TM_UPML_2D", m.TM_UPML_2D), ("TE_UPML_2D", m.TE_UPML_2D),
               ("MPI_TM_UPML_2D", m.MPI_TM_UPML_2D), ("MPI_TE_UPML_2D", m.MPI_TE_UPML_2D)]
    per = {}
    cells = 0
    for name, solver in solvers:
        log(f"solvers_fp64: {name}, {steps} steps")
        sim = m.Simulator(solver, getattr(m, model), m.field_info(w, h, hu, pml, lam, ang, steps + 20),
                          params=m.model_params(**params), ntff_time=False, precision="fp64")
        try:
            cells = sim.N_PX * sim.N_PY
            sim.calc(20)
            L.fdtd_hip_synchronize()
            t0 = time.perf_counter()
            sim.calc(steps)
            L.fdtd_hip_synchronize()
            dt = time.perf_counter() - t0
            f0 = sim.fields()[0]
            assert np.all(np.isfinite(f0)) and np.abs(f0).max() > 0, name
        finally:
            sim.finish()
        per[name] = {"mcups": round(cells * steps / dt / 1e6, 1), "us_per_step": round(dt * 1e6 / steps, 2)}
    return {"metric": "Mcell-updates/s, FP64 reference-order path, per solver", "value": per["NS_TE_2D"]["mcups"],
            "unit": "Mcell-updates/s", "n_gpus": 1, "steps": steps, "warmup": 20,
            "ms_per_step": round(per["NS_TE_2D"]["us_per_step"] / 1e3, 5), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "every solver of enum SOLVER on the 1056x1056 Mie cylinder grid (h=5 nm, "
                                   "r=500 nm, n=1.6, lambda=550 nm), FP64, no per-step NTFF; value: NS_TE_2D, "
                                   "the reference main.c's solver", "cells": cells, "per_solver": per},
            "roofline": None, "cpu_baseline": None}


def run_sweep(args, m):
    """The batched sweep: every instance to completion, timed as one job."""
    import numpy as np
    from mpifdtd_amd import batch as B
    T, R = args.T or 8, args.R or 16
    spec = B.sweep_spec(args.seed, 64)
    items = B.sweep_instances(spec)
    cells = [(s["side"] + 2 * s["pml"]) ** 2 for s in spec]
    upd = sum(c * s["steps"] for c, s in zip(cells, spec))
    log(f"sweep64: {len(spec)} instances, {sum(cells) / 1e6:.1f} M cells, {upd / 1e9:.1f} G cell-updates, T={T} R={R}")
    t0 = time.perf_counter()
    b = B.Batch(items, T=T, R=R)
    if args.launch >= 0:
        b.set_launch_mode(args.launch)
    m.lib().fdtd_hip_synchronize()
    t_init = time.perf_counter() - t0
    log(f"created in {t_init:.1f} s")
    t0 = time.perf_counter()
    b.run()
    m.lib().fdtd_hip_synchronize()
    dt = time.perf_counter() - t0
    done = b.sync()
    assert done == [s["steps"] for s in spec], "an instance did not finish"
    ff = [b.far_field(i) for i in (0, 1)]
    check = {}
    if not args.no_check:
        # two instances again through the FP64 reference-order single-instance path
        for i in (0, 1):
            s = spec[i]
            solver = m.MPI_TM_UPML_2D if s["mode"] == "tm" else m.MPI_TE_UPML_2D
            model = m.LAYER if s["kind"] == "multilayer" else m.ZIGZAG
            ref = m.Simulator(solver, model, m.field_info(s["width_nm"], s["height_nm"], s["h_nm"], s["pml"],
                                                         s["lambda_nm"], 0, s["steps"]),
                              params=m.model_params(**s["params"]), ntff_time=False)
            ref.calc(s["steps"])
            ff64 = ref.far_field()
            ref.finish()
            check[f"ff_rel_err_vs_fp64_inst{i}"] = float(np.linalg.norm(ff[i] - ff64) / np.linalg.norm(ff64))
    b.close()
    alg = sum(c * s["steps"] * (52 if s["mode"] == "tm" else 56) for c, s in zip(cells, spec))
    achieved = alg / dt / 1e9
    line = {
        "metric": "Mcell-updates/s + % HBM roofline; spectra/s at <=1e-4 far-field err",
        "value": round(upd / dt / 1e6, 1), "unit": "Mcell-updates/s", "n_gpus": 1, "steps": max(s["steps"] for s in spec),
        "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"batched sweep of 64 seeded instances (seed {args.seed}): multilayer/zigzag, "
                               "sides 512-2048 + 16-cell UPML, TM/TE, lambda 380-780 nm, 6000-10000 steps",
                   "cells": sum(cells), "cell_updates": upd, "T": T, "R": R,
                   "launch": SHAPES[args.launch] if args.launch >= 2 else "two_streams_pml8",
                   "create_s": round(t_init, 2)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "note": "algorithmic bytes (interior-cell 52/56 B) of the whole sweep over its wall time"},
        "cpu_baseline": None,
        "spectra_per_s": round(len(spec) / dt, 3),
        "wall_s": round(dt, 3),
    }
    line.update(check)
    return line


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
    # the reference's CPU path, timed first: a child process started before anything
    # here has touched the GPU
    cpu_ref = None
    if not args.no_cpu and args.workload == "mie1056":
        log("cpu baseline (the reference, one core)")
        cpu_ref = cpu_reference(args.mode, args.workload)
    import numpy as np

    import mpifdtd_amd as m
    from mpifdtd_amd import autotune
    L = m.lib()
    # the library's own HIP runtime does the synchronising (fdtd_hip_synchronize =
    # hipDeviceSynchronize): a GPU runtime loaded beside it would not see its queues
    assert L.fdtd_hip_device_count() >= 1, "no GPU"
    if args.workload in ("sweep64", "morpho8256_sweep81", "solvers_fp64"):
        line = {"sweep64": run_sweep, "morpho8256_sweep81": run_sweep81,
                "solvers_fp64": run_solvers_fp64}[args.workload](args, m)
        libc.fflush(None)
        sys.stdout.flush()
        os.dup2(real_stdout, 1)
        print(json.dumps(line), flush=True)
        return
    K, W = args.steps, args.warmup
    total = K + W
    wl = args.workload

    T, R = args.T, args.R
    tuned = None
    if args.precision == "fp32" and (T == 0 or R == 0):
        log("autotune")
        probe = 240 if wl == "mie1056" else 48
        tuned = autotune.tune(m, lambda: make_sim(m, "fp32", args.mode, 2000, workload=wl, ntff=args.ntff),
                              probe_steps=probe)
        T = T or tuned["T"]
        R = R or tuned["R"]
    launch = args.launch if args.launch >= 0 else (tuned.get("launch", 0) if tuned else 0)
    log(f"timed run: {wl} {args.precision} T={T} R={R} launch={launch} warmup={W} steps={K}")
    sim = make_sim(m, args.precision, args.mode, total, T or 4, R or 16, workload=wl, ntff=args.ntff)
    if args.precision == "fp32":
        sim.set_launch_mode(launch)
        sim.set_ntff_overlap(args.ntff_overlap)
    cells = sim.N_PX * sim.N_PY
    sim.calc(W)
    if args.precision == "fp32":
        L.fdtd_solver_setBlocking(T, R)  # launches the queued warm-up steps
    L.fdtd_hip_synchronize()
    # events on every launch of a long run of short launches would cost the timed
    # region a few microseconds per launch: then every 8th launch is timed
    nblk = K // max(1, T or 1) if args.precision == "fp32" else 0
    L.fdtd_hip_profile_enable(8 if nblk >= 256 else 1)
    L.fdtd_hip_synchronize()
    t0 = time.perf_counter()
    sim.calc(K)
    if args.precision == "fp32":
        L.fdtd_solver_setBlocking(T, R)  # launches the last partial block inside the timed region
    L.fdtd_hip_synchronize()
    dt = time.perf_counter() - t0
    prof = prof_read(L)
    L.fdtd_hip_profile_enable(0)
    t1 = time.perf_counter()
    ff = sim.far_field()
    t_ff = time.perf_counter() - t1
    f32_0 = sim.fields()[0]  # TM Ez / TE Ex (a copy)
    shape = sim.launch_mode() if args.precision == "fp32" else 0
    shape_name = SHAPES[shape]
    sim.finish()
    ms_step = dt * 1e3 / K
    mcups = cells * K / dt / 1e6

    # roofline of the dominant kernel: the fused launch (UPML + interior tiles, T steps)
    ms_k, _, nlaunch, b_int, b_pml, cupd = prof
    if args.precision == "fp32" and nlaunch > 0 and ms_k > 0:
        achieved = (b_int + b_pml) / (ms_k * 1e-3) / 1e9
        traffic, tsrc = measured_traffic(args.mode, T, R, cells, shape_name)
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": tsrc,
                "kernel": f"k_fast<{args.mode.upper()},MPI,R={R}>" + (
                    "" if shape_name == "merged" else f" interior + UPML launches ({shape_name.replace('_', ' ')})"),
                "launch_avg_us": round(ms_k * 1e3 / nlaunch, 2), "steps_per_launch": T,
                "launches_timed": int(nlaunch), "timed_every": 8 if nblk >= 256 else 1,
                "bytes_per_launch": round((b_int + b_pml) / nlaunch),
                "bytes_per_cell_update": round((b_int + b_pml) / cupd, 2)}
    else:
        roof = {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None, "traffic": None}

    extra = {}
    # spectra/s: one spectrum = a full run of the workload's step count plus its far field
    spec_steps = 8000 if wl == "mie1056" else 20000
    extra["spectra_per_s"] = round(1.0 / (spec_steps * ms_step * 1e-3 + t_ff), 4)
    if not args.no_check and args.precision == "fp32":
        # FP32 far field vs the FP64 reference-order GPU path on the same run length
        log("fp64 reference-order run for the far-field check")
        ref = make_sim(m, "fp64", args.mode, total, workload=wl, ntff=args.ntff)
        t2 = time.perf_counter()
        ref.calc(total)
        ff64 = ref.far_field()
        t64 = time.perf_counter() - t2
        f64 = ref.fields()[0]
        ref.finish()
        nref = float(np.linalg.norm(ff64))
        # a short run of a large grid has no scattered field on the NTFF contour yet
        extra["ff_rel_err_vs_fp64"] = float(np.linalg.norm(ff - ff64) / nref) if nref > 0 else None
        extra["field0_rel_err_vs_fp64"] = float(np.linalg.norm(f32_0 - f64) / max(np.linalg.norm(f64), 1e-300))
        extra["fp64_mcups"] = round(cells * total / t64 / 1e6, 1)
        extra["fp64_spectra_per_s"] = round(1.0 / (spec_steps * t64 / total), 4)
        extra["speedup_vs_fp64_spectra"] = round(extra["spectra_per_s"] / extra["fp64_spectra_per_s"], 2)
    # the CPU leg runs on the default workload only: the one-core oracle needs minutes
    # just to sample the Morpho model's eps on 34 M cells
    run_cpu = not args.no_cpu and wl == "mie1056" and cpu_ref is None
    if run_cpu:
        log("cpu baseline (oracle, one core)")
    cpu = cpu_ref if cpu_ref is not None else (cpu_baseline(args.mode, wl) if run_cpu else None)
    line = {
        "metric": "Mcell-updates/s + % HBM roofline; spectra/s at <=1e-4 far-field err",
        "value": round(mcups, 1), "unit": "Mcell-updates/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": round(ms_step, 5), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f64", "data": "synthetic",
        "config": {"workload": WORKLOADS[wl][8].format(mode=args.mode.upper()), "cells": cells,
                   "precision": args.precision, "T": T, "R": R,
                   "launch": shape_name if args.precision == "fp32" else None,
                   "per_step_ntff": args.ntff, "autotune": tuned},
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
