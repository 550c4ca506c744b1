"""Benchmark of the FP32 fused UPML step through the reference's driver API.

Default workload: the Morpho-scale wavelength sweep on the large grid, the
throughput benchmark.  The Morpho model (morphoScaleModel.c, default
parameters) on 8192 x 4096 cells + a 32-cell UPML (8256 x 4160, 34.3 M cells,
h = 5 nm: the integer-nm twin of 2.5 nm with every length doubled), normal
incidence (+y), 20000 steps per run.  The 81 wavelengths 380-780 nm (every
5 nm) come from one run of the reference's plain pulse solver (fdtdTM_upml.c /
fdtdTE_upml.c: a Gaussian plane-wave pulse whose band covers the sweep) with
the running DFT of the NTFF contour at the 81 wavelengths, instead of 81 CW
runs.  One step = one FDTD time step of the whole grid (H and E halves, UPML
auxiliaries, pulse source) plus the contour samples and their DFT update.  TM
is the headline; TE is reported under "te".

The timed region is K steps (default 20000, a whole run) in whole blocks of T
steps (T divides K), on inputs resident in HBM; the host copies nothing inside
it.  roofline.achieved is the HBM traffic of one block's step launches, from
the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of exactly this launch
configuration under profiles/, divided by the block's event-timed duration in
this run; the algorithmic rate (what an unblocked step would have to move) is
reported beside it as "effective".  The DFT and far-field spectra of the FP32
run are compared with the FP64 reference-order path on the same grid after
`--check-steps` steps.  The CPU baseline is the reference's own plain TM UPML
step, compiled from its sources (oracle/_ref/ref_harness), timed on one host
core and on all host cores before the GPU is touched.

Other workloads (--workload): mie1056 (the Mie cylinder, 1056 x 1056, MPI
solver), morpho8256 (the same Morpho grid with the CW MPI solver), sweep64 (64
seeded batched instances), solvers_fp64 (every solver of enum SOLVER in FP64).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload ...] [--mode tm|te|both]
prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Mcell-updates/s + % HBM roofline; spectra/s at <=1e-4 far-field err"
# launch shapes of the UPML and interior tiles (FAST_LAUNCH_* of fdtd_hip.h)
SHAPES = ["auto", "merged", "split", "two_streams", "split_upml48", "two_streams_upml48", "split_w8", "two_streams_w8",
          "split_p16", "two_streams_p16"]

# name -> (model, width_nm, height_nm, h_nm, pml, lambda_nm, angle_deg, model params, description)
WORKLOADS = {
    "mie1056": ("MIE_CYLINDER", 5280, 5280, 5, 16, 550, 0, {"circle_radius_nm": 500, "circle_n": 1.6},
                "Mie cylinder {mode}, 1024x1024 + 16-cell UPML (1056x1056), h=5 nm, r=500 nm, n=1.6, lambda=550 nm"),
    "morpho8256": ("MORPHO_SCALE", 41280, 20800, 5, 32, 550, 90, {},
                   "Morpho scale {mode}, 8192x4096 + 32-cell UPML (8256x4160), h=5 nm, lambda=550 nm, "
                   "normal incidence (+y)"),
}
SWEEP_LAMBDAS = (380.0, 781.0, 5.0)  # 81 wavelengths
RUN_STEPS = 20000  # steps of one run of the sweep
# Launch configuration of the sweep per mode: (T, R, launch shape), the
# fastest of `--autotune`'s candidates on this grid (interior tiles as 8
# wavefronts of 10 rows; UPML tiles as 16 wavefronts of 3 rows), in the split
# form of the shapes (the two-stream forms time within 2 %, and the split form
# times each launch on its own, as the counter summaries need).  T divides the
# default K.
SWEEP_CONFIG = {"tm": (10, 20, 8), "te": (10, 20, 8)}
T_CANDIDATES = (10, 8, 5, 4, 2, 1)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0,
                    help="timed steps (default: 20000 for the sweep, 8000 mie1056, 400 morpho8256)")
    ap.add_argument("--warmup", type=int, default=-1, help="untimed steps (default 100; mie1056 400)")
    ap.add_argument("--workload", default="morpho8256_sweep81",
                    choices=sorted(WORKLOADS) + ["sweep64", "morpho8256_sweep81", "solvers_fp64"])
    ap.add_argument("--seed", type=int, default=1, help="sweep64: generator seed")
    ap.add_argument("--precision", default="fp32", choices=("fp32", "fp64"))
    ap.add_argument("--mode", default="", choices=("", "tm", "te", "both"),
                    help="TM, TE or both (the sweep's default: both, TM the headline)")
    ap.add_argument("--T", type=int, default=0, help="temporal block depth (0: the workload's default)")
    ap.add_argument("--R", type=int, default=0, help="rows per wavefront (0: the workload's default)")
    ap.add_argument("--launch", type=int, default=-1,
                    help="launch shape of the UPML and interior tiles: 0 auto, 1 merged, 2 split, "
                         "3 split on two streams, 4 / 5: 2 / 3 with the 48-row UPML tiles, "
                         "6 / 7: 4 / 5 with 8-wavefront interior workgroups, 8 / 9: 6 / 7 with "
                         "16-wavefront UPML workgroups (-1: the workload's default)")
    ap.add_argument("--autotune", action="store_true", help="time the candidate configurations first and use the best")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--ntff-overlap", type=int, default=-1,
                    help="contour sample work beside the next block (1), after it (0), by grid size (-1)")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 comparison")
    ap.add_argument("--check-steps", type=int, default=8000,
                    help="sweep: steps of the FP32 / FP64 comparison run (the scattered pulse has crossed every "
                         "contour side by then)")
    ap.add_argument("--ntff", action="store_true",
                    help="mie1056 / morpho8256: the reference's per-step time-domain NTFF inside the timed steps")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the library's streaming device copy of MIB MiB (10 times): the HBM ceiling, "
                         "and the kernel that scales rocprofv3's FETCH_SIZE / WRITE_SIZE counters to bytes")
    a = ap.parse_args()
    sweep = a.workload == "morpho8256_sweep81"
    a.steps_given = a.steps > 0
    if a.steps <= 0:
        a.steps = RUN_STEPS if sweep else (8000 if a.workload == "mie1056" else 400)
    if a.warmup < 0:
        a.warmup = 400 if a.workload == "mie1056" else (100 if sweep else 40)
    if not a.mode:
        a.mode = "both" if sweep else "tm"
    if a.mode == "both" and not sweep:
        sys.exit("--mode both is for the sweep workload")
    if a.R in (18, 20) and a.launch not in (6, 7, 8, 9) and not (sweep and a.launch == -1):
        sys.exit(f"--R {a.R} ({4 * a.R}-row interior tiles as 8 wavefronts) needs --launch 6, 7, 8 or 9")
    return a


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def copy_calibration(mib):
    """The library's streaming copy kernel (k_stream_copy): the achievable HBM
    rate, and the plain copy of known size whose FETCH_SIZE / WRITE_SIZE
    counters scale the step kernels' counters to bytes."""
    import ctypes
    import mpifdtd_amd as m
    from mpifdtd_amd._native import check
    out = (ctypes.c_double * 4)()
    check(m.lib().fdtd_hip_copy_rate(mib << 20, 10, out), "fdtd_hip_copy_rate")
    print(json.dumps({"copy_bytes_read": (mib << 20), "copy_bytes_written": (mib << 20), "launches": 11,
                      "best_GBs": round(out[0], 1), "mean_GBs": round(out[1], 1), "mean_us": round(out[2], 2)}),
          flush=True)


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


# ---------------------------------------------------------------- CPU baseline

def _harness(args, timeout):
    import subprocess
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=timeout, cwd="/tmp")
    return json.loads(r.stdout.strip().splitlines()[-1])


def cpu_reference_sweep():
    """The reference's own plain TM UPML step (fdtdTM_upml.c update: calcMB,
    calcH, calcJD, calcE, field_scatteredPulse, ntffTM_TimeCalc), compiled from
    its sources with gcc -O2 (oracle/_ref/ref_harness, one MPI rank per
    process), Morpho model, h = 5 nm, 32-cell UPML, normal incidence.  Its
    allocation, eps fill and coefficients run before the timed window.

    On the 8256 x 4160 grid itself the reference's per-step NTFF indexes its
    bins before their start (RFperC comes from the y extent only, field.c:140-142,
    so the retarded-time shift of the far x side is about -1700 bins) and the
    process dies with a segmentation fault; the one-core sample therefore runs a
    square 5860 x 5860 grid of the same cell count (34.3 M).  All cores: the
    reference's rank-per-simulation mode (main.c:114-138), one independent
    simulation per core, each of 1/ncores of those cells (square), run
    together.  Runs as child processes before this process touches the GPU.
    Returns None when the harness is not built or does not run."""
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if not os.access(exe, os.X_OK):
        return None
    import subprocess
    try:
        ncores = len(os.sched_getaffinity(0))
    except AttributeError:
        ncores = os.cpu_count() or 1
    ncores = max(1, min(16, ncores))
    base = ["mode=time", "solver=2", "model=3", "h=5", "pml=32", "lambda=550", "angle=90", "warm=1", "ntff=1"]
    side1 = 5860
    try:
        log("cpu baseline: the reference's plain TM UPML step, 5860 x 5860, one core")
        r1 = _harness(base + [f"width={side1 * 5}", f"height={side1 * 5}", "steps=8"], 400)
        one = r1["N_PX"] * r1["N_PY"] * r1["steps"] / r1["seconds"] / 1e6
        side_n = int(round(side1 / math.sqrt(ncores)))
        log(f"cpu baseline: {ncores} ranks of {side_n} x {side_n}, together")
        procs = [subprocess.Popen([exe] + base + [f"width={side_n * 5}", f"height={side_n * 5}", "steps=40"],
                                  stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True, cwd="/tmp")
                 for _ in range(ncores)]
        outs = [p.communicate(timeout=400)[0] for p in procs]
        recs = [json.loads(o.strip().splitlines()[-1]) for o in outs]
        agg = sum(r["N_PX"] * r["N_PY"] * r["steps"] / r["seconds"] for r in recs) / 1e6
        secs = max(r["seconds"] for r in recs)

        # the step without its per-step NTFF: the NTFF timed alone for the same
        # calls on the same fields (ntff=1), subtracted
        def no_ntff(r):
            t = r["seconds"] - r.get("ntff_seconds", -1.0)
            return r["N_PX"] * r["N_PY"] * r["steps"] / t if 0 <= r.get("ntff_seconds", -1.0) and t > 0 else None
        rates = [no_ntff(r) for r in recs]
        agg_no = sum(rates) / 1e6 if all(x is not None for x in rates) else None
        one_no = no_ntff(r1)
    except (OSError, ValueError, IndexError, KeyError, subprocess.SubprocessError) as e:
        log(f"cpu baseline failed: {e}")
        return None
    return {"value": round(agg, 3), "unit": "Mcell-updates/s", "cores": ncores, "kind": "reference",
            "value_1core": round(one, 3),
            "value_no_ntff": round(agg_no, 3) if agg_no is not None else None,
            "value_1core_no_ntff": round(one_no / 1e6, 3) if one_no is not None else None,
            "sample": f"the reference's plain TM UPML step (with its per-step time-domain NTFF) built from its "
                      f"sources with gcc -O2; all cores: {ncores} independent simulations of {side_n}x{side_n} "
                      f"cells together (rank per simulation, main.c:114-138), 40 steps each ({secs:.1f} s); "
                      f"one core: {r1['N_PX']}x{r1['N_PY']} (34.3 M cells, the 8256x4160 grid's count; the "
                      f"reference's NTFF bins overflow on the non-square grid itself), {r1['steps']} steps "
                      f"({r1['seconds']:.1f} s); Morpho model, h=5 nm, 32-cell UPML, eps fill outside the window; "
                      f"*_no_ntff: the same steps less their per-step NTFF (ntffTM_TimeCalc, fdtdTM_upml.c:65), "
                      f"timed alone for the same calls on the same fields and subtracted"}


def cpu_reference(mode, workload, steps=120):
    """mie1056: the reference's MPI-TM / TE UPML solver compiled from its own
    sources (oracle/_ref/ref_harness, one MPI rank, one core), `steps`
    simulator_calc() calls including the per-step NTFF, timed inside the
    harness, as a child process before this process touches the GPU."""
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if not os.access(exe, os.X_OK):
        return None
    model, w, h, hu, pml, lam, ang, params, _ = WORKLOADS[workload]
    if model != "MIE_CYLINDER" or params != {"circle_radius_nm": 500, "circle_n": 1.6}:
        return None  # the reference's cylinder is r = 500 nm, n = 1.6 (circleModel.c:6-10)
    import subprocess
    try:
        res = _harness(["mode=time", "solver=%d" % (4 if mode == "tm" else 5), "model=1", f"width={w}",
                        f"height={h}", f"h={hu}", f"pml={pml}", f"lambda={lam}", f"angle={ang}", f"steps={steps}",
                        "warm=2"], 240)
    except (OSError, ValueError, IndexError, subprocess.SubprocessError):
        return None
    cells = res["N_PX"] * res["N_PY"]
    return {"value": round(cells * res["steps"] / res["seconds"] / 1e6, 3), "unit": "Mcell-updates/s", "cores": 1,
            "kind": "reference",
            "sample": f"{res['steps']} steps of the same {res['N_PX']}x{res['N_PY']} {workload} {mode.upper()} grid, "
                      f"the reference's own MPI {mode.upper()} UPML solver (FP64, per-step NTFF) built from its "
                      f"sources with gcc -O2, one MPI rank ({res['seconds']:.1f} s on one host core)"}


# ---------------------------------------------------------------- roofline

def measured_traffic(key):
    """HBM bytes per block of T steps (the step launches of one block) of the
    launch configuration `key` from the committed rocprofv3 counter summaries
    (FETCH_SIZE / WRITE_SIZE scaled by the same counters on a plain copy of
    known size), and the summary's path; (None, None) when none matches."""
    import glob
    import re
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*traffic*.md"))):
        for block in re.findall(r"```json\n(.*?)\n```", open(path).read(), re.S):
            try:
                rec = json.loads(block)
            except ValueError:
                continue
            if all(rec.get(k) == v for k, v in key.items()):
                return rec["traffic_bytes_per_launch"], os.path.relpath(path, ROOT)
    return None, None


def roofline(prof, key, T, kernel):
    """The roofline object of the step launches: `achieved` = counter HBM bytes
    per block / event-timed block duration (null without a matching counter
    summary), `effective` = the algorithmic bytes of the block's T steps (one
    read and one write of every field per step) over the same duration."""
    ms_k, _, nlaunch, b_int, b_pml, cupd = prof
    if not (nlaunch > 0 and ms_k > 0):
        return {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
                "traffic": None}
    block_s = ms_k * 1e-3 / nlaunch
    alg_block = (b_int + b_pml) / nlaunch  # T steps' algorithmic bytes
    traffic, src = measured_traffic(key)
    achieved = traffic / block_s / 1e9 if traffic else None
    frac = achieved / HBM_PEAK_GBS if achieved else None
    note = None
    if frac is not None and frac > 1.0:  # not a physical rate: do not publish it
        note = f"counter rate {achieved:.0f} GB/s above peak: not published"
        achieved = frac = None
    out = {"bound": "hbm", "achieved": round(achieved, 1) if achieved else None, "peak": HBM_PEAK_GBS,
           "unit": "GB/s", "frac": round(frac, 4) if frac else None, "traffic": traffic, "traffic_source": src,
           "kernel": kernel, "launch_avg_us": round(block_s * 1e6, 2), "steps_per_launch": T,
           "launches_timed": int(nlaunch),
           "compulsory_bytes_per_launch": round(alg_block / T),
           "traffic_over_compulsory": round(traffic / (alg_block / T), 3) if traffic else None,
           "effective": {"GBs": round(alg_block / block_s / 1e9, 1),
                         "bytes_per_launch": round(alg_block),
                         "bytes_per_cell_update": round((b_int + b_pml) / cupd, 2) if cupd else None,
                         "note": "algorithmic bytes of the block's T steps (52/56 B per interior cell-update, "
                                 "148/152 B per UPML cell-update) over its duration: the rate an unblocked "
                                 "kernel would need, not HBM traffic"}}
    if note:
        out["note"] = note
    return out


# ---------------------------------------------------------------- the sweep

def pick_T(T, K):
    """T itself when it divides K, else the deepest candidate that does (the
    timed region holds whole blocks only)."""
    if K % T == 0:
        return T
    for c in T_CANDIDATES:
        if c <= T and K % c == 0:
            return c
    return 1


def tune_sweep(m, mode, lams):
    """Time the candidate (T, R, launch shape) on the sweep grid with the DFT on."""
    L = m.lib()
    model, w, h, hu, pml, lam, ang, params, _ = WORKLOADS["morpho8256"]
    solver = m.TM_UPML_2D if mode == "tm" else m.TE_UPML_2D
    probe = m.Simulator(solver, getattr(m, model), m.field_info(w, h, hu, pml, lam, ang, 4000),
                        params=m.model_params(**params), ntff_time=False, precision="fp32", T=10, R=16)
    tune = {}
    try:
        probe.set_dft(lams)
        probe.calc(40)
        for Tc, Rc in ((10, 20), (10, 18), (10, 16), (8, 16), (10, 8), (8, 8)):
            for sh in (1, 2, 3, 4, 5, 6, 7, 8, 9):
                if Rc in (18, 20) and sh not in (6, 7, 8, 9):
                    continue  # 72- and 80-row interior tiles exist as 8 wavefronts only
                n = 6 * Tc
                probe.set_blocking(Tc, Rc)
                probe.set_launch_mode(sh)
                probe.calc(2 * Tc)
                probe.set_blocking(Tc, Rc)
                L.fdtd_hip_synchronize()
                t0 = time.perf_counter()
                probe.calc(n)
                L.fdtd_hip_synchronize()
                tune[(Tc, Rc, sh)] = (time.perf_counter() - t0) * 1e6 / n
    finally:
        probe.finish()
    return tune


def run_sweep81_mode(args, m, mode):
    import numpy as np
    L = m.lib()
    model, w, h, hu, pml, lam, ang, params, _ = WORKLOADS["morpho8256"]
    lams = np.arange(*SWEEP_LAMBDAS)
    solver = m.TM_UPML_2D if mode == "tm" else m.TE_UPML_2D
    K, W = args.steps, args.warmup
    T, R, launch = SWEEP_CONFIG[mode]
    tune = {}
    if args.autotune:
        log(f"sweep81 {mode}: autotune")
        tune = tune_sweep(m, mode, lams)
        T, R, launch = min(tune, key=tune.get)
    T, R = args.T or T, args.R or R
    launch = args.launch if args.launch >= 0 else launch
    T = pick_T(T, K)

    def make(precision, n):
        return m.Simulator(solver, getattr(m, model), m.field_info(w, h, hu, pml, lam, ang, n),
                           params=m.model_params(**params), ntff_time=False, precision=precision, T=T, R=R)

    log(f"sweep81 {mode}: {len(lams)} wavelengths, warmup {W}, timed {K} steps, T={T} R={R} "
        f"launch={SHAPES[launch]}")
    sim = make("fp32", W + K)
    try:
        cells = sim.N_PX * sim.N_PY
        sim.set_launch_mode(launch)
        sim.set_ntff_overlap(args.ntff_overlap)
        sim.set_dft(lams)
        sim.calc(W)
        sim.set_blocking(T, R)  # launches a partial warm-up block, if any
        L.fdtd_hip_synchronize()
        nblk = K // T
        every = 8 if nblk >= 256 else 1
        L.fdtd_hip_profile_enable(every)
        t0 = time.perf_counter()
        sim.calc(K)  # whole blocks: the last one launches with its T-th step
        L.fdtd_hip_synchronize()
        dt = time.perf_counter() - t0
        prof = prof_read(L)
        L.fdtd_hip_profile_enable(0)
        t1 = time.perf_counter()
        spec = sim.dft_spectra()
        t_spec = time.perf_counter() - t1
        shape = SHAPES[sim.launch_mode()]
    finally:
        sim.finish()
    # a short timed run ends before the pulse reaches the contour (spectra all 0);
    # the comparison run below checks the spectra once they are formed
    assert np.all(np.isfinite(spec)), "sweep spectra"
    ms_step = dt * 1e3 / K
    res = {"value": round(cells * K / dt / 1e6, 1), "ms_per_step": round(ms_step, 5), "steps": K, "warmup": W,
           "T": T, "R": R, "launch": shape, "cells": cells, "wavelengths": int(len(lams))}
    # 81 spectra per run of RUN_STEPS steps (measured when K is a whole run)
    res["spectra_per_s"] = round(len(lams) / (RUN_STEPS * ms_step * 1e-3 + t_spec), 3)
    key = {"workload": "morpho8256_sweep81", "mode": mode, "T": T, "R": R, "launch": shape, "cells": cells}
    res["roofline"] = roofline(prof, key, T, f"{mode.upper()} step launches of one block of T steps: interior "
                                             f"k_fast R={R} + UPML tiles ({shape}), with contour samples")
    res["roofline"]["timed_every"] = every
    if tune:
        res["tune_us_per_step"] = {f"T{k[0]}R{k[1]}{SHAPES[k[2]]}": round(v, 2) for k, v in tune.items()}
    if not args.no_check:
        nchk = args.check_steps
        log(f"sweep81 {mode}: fp32 / fp64 comparison over {nchk} steps")
        out = {}
        for prec in ("fp32", "fp64"):
            c = make(prec, nchk)
            try:
                if prec == "fp32":
                    c.set_launch_mode(launch)
                c.set_dft(lams)
                L.fdtd_hip_synchronize()
                t2 = time.perf_counter()
                c.calc(nchk)
                if prec == "fp32":
                    c.set_blocking(T, R)
                L.fdtd_hip_synchronize()
                tt = time.perf_counter() - t2
                out[prec] = (c.dft(), c.dft_spectra(), tt)
            finally:
                c.finish()
        d32, s32, _ = out["fp32"]
        d64, s64, t64 = out["fp64"]
        assert np.linalg.norm(d64) > 0 and np.abs(s64[:, 180:]).max() > 0, \
            "the comparison run ended before the pulse reached the contour"
        res["check_steps"] = nchk
        res["dft_rel_err_vs_fp64"] = float(np.linalg.norm(d32 - d64) / np.linalg.norm(d64))
        # far-field spectra: all 360 angles, and the 180 back-hemisphere ones (reflectance)
        res["ff_rel_err_vs_fp64"] = float(np.linalg.norm(s32 - s64) / np.linalg.norm(s64))
        res["reflectance_rel_err_vs_fp64"] = float(np.linalg.norm(s32[:, 180:] - s64[:, 180:]) /
                                                   np.linalg.norm(s64[:, 180:]))
        ms64 = t64 * 1e3 / nchk
        res["fp64_mcups"] = round(cells / (ms64 * 1e-3) / 1e6, 1)
        res["fp64_spectra_per_s"] = round(len(lams) / (RUN_STEPS * ms64 * 1e-3), 4)
        res["speedup_vs_fp64_spectra"] = round(res["spectra_per_s"] / res["fp64_spectra_per_s"], 2)
    return res


def run_sweep81(args, m, cpu):
    modes = ["tm", "te"] if args.mode == "both" else [args.mode]
    per = {md: run_sweep81_mode(args, m, md) for md in modes}
    head = per[modes[0]]
    line = {
        "metric": METRIC, "value": head["value"], "unit": "Mcell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"Morpho-scale wavelength sweep {modes[0].upper()}: 8192x4096 + 32-cell UPML "
                               "(8256x4160, 34.3 M cells), h=5 nm (the integer twin of 2.5 nm), default Morpho "
                               "model, plain pulse solver at normal incidence (+y), running DFT of the NTFF contour "
                               "at 380-780 nm every 5 nm (81 wavelengths) in one 20000-step run",
                   "cells": head["cells"], "precision": "fp32", "T": head["T"], "R": head["R"],
                   "launch": head["launch"], "wavelengths": head["wavelengths"], "run_steps": RUN_STEPS},
        "roofline": head["roofline"],
        "cpu_baseline": cpu,
    }
    for k in ("spectra_per_s", "check_steps", "dft_rel_err_vs_fp64", "ff_rel_err_vs_fp64",
              "reflectance_rel_err_vs_fp64", "fp64_mcups", "fp64_spectra_per_s", "speedup_vs_fp64_spectra",
              "tune_us_per_step"):
        if k in head:
            line[k] = head[k]
    if cpu:
        line["speedup_vs_cpu_all_cores"] = round(head["value"] / cpu["value"], 1)
        if cpu.get("value_no_ntff"):
            line["speedup_vs_cpu_all_cores_no_ntff"] = round(head["value"] / cpu["value_no_ntff"], 1)
    for md in modes[1:]:
        line[md] = per[md]
    return line


# ---------------------------------------------------------------- other workloads

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
    line = {
        "metric": METRIC,
        "value": round(upd / dt / 1e6, 1), "unit": "Mcell-updates/s", "n_gpus": 1, "steps": max(s["steps"] for s in spec),
        "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"batched sweep of 64 seeded instances (seed {args.seed}): multilayer/zigzag, "
                               "sides 512-2048 + 16-cell UPML, TM/TE, lambda 380-780 nm, 6000-10000 steps",
                   "cells": sum(cells), "cell_updates": upd, "T": T, "R": R,
                   "launch": SHAPES[args.launch] if args.launch >= 2 else "two_streams_pml8",
                   "create_s": round(t_init, 2)},
        "roofline": {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
                     "traffic": None,
                     "effective": {"GBs": round(alg / dt / 1e9, 1),
                                   "note": "algorithmic bytes (interior-cell 52/56 B) of the whole sweep over its "
                                           "wall time, not HBM traffic"}},
        "cpu_baseline": None,
        "spectra_per_s": round(len(spec) / dt, 3),
        "wall_s": round(dt, 3),
    }
    line.update(check)
    return line


def run_single(args, m, cpu_ref):
    """mie1056 / morpho8256: the MPI solver with its CW scattered-wave source."""
    import numpy as np
    from mpifdtd_amd import autotune
    L = m.lib()
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
    if args.precision == "fp32":
        T = pick_T(T or 4, K)
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
    nblk = K // max(1, T or 1) if args.precision == "fp32" else 0
    every = 8 if nblk >= 256 else 1
    L.fdtd_hip_profile_enable(every)
    t0 = time.perf_counter()
    sim.calc(K)  # whole blocks
    L.fdtd_hip_synchronize()
    dt = time.perf_counter() - t0
    prof = prof_read(L)
    L.fdtd_hip_profile_enable(0)
    t1 = time.perf_counter()
    ff = sim.far_field()
    t_ff = time.perf_counter() - t1
    f32_0 = sim.fields()[0]
    shape_name = SHAPES[sim.launch_mode() if args.precision == "fp32" else 0]
    sim.finish()
    ms_step = dt * 1e3 / K
    if args.precision == "fp32":
        key = {"mode": args.mode, "T": T, "R": R, "cells": cells, "launch": shape_name,
               "workload": wl + ("_ntff" if args.ntff else "")}
        roof = roofline(prof, key, T, f"k_fast<{args.mode.upper()},MPI,R={R}> interior + UPML launches ({shape_name})")
        roof["timed_every"] = every
    else:
        roof = {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None, "traffic": None}
    extra = {}
    spec_steps = 8000 if wl == "mie1056" else 20000
    extra["spectra_per_s"] = round(1.0 / (spec_steps * ms_step * 1e-3 + t_ff), 4)
    if not args.no_check and args.precision == "fp32":
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
    line = {
        "metric": METRIC, "value": round(cells * K / dt / 1e6, 1), "unit": "Mcell-updates/s", "n_gpus": 1,
        "steps": K, "warmup": W, "ms_per_step": round(ms_step, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if args.precision == "fp32" else "f64", "data": "synthetic",
        "config": {"workload": WORKLOADS[wl][8].format(mode=args.mode.upper()), "cells": cells,
                   "precision": args.precision, "T": T, "R": R,
                   "launch": shape_name if args.precision == "fp32" else None,
                   "per_step_ntff": args.ntff, "autotune": tuned},
        "roofline": roof,
        "cpu_baseline": cpu_ref,
    }
    line.update(extra)
    return line


def main():
    args = parse()
    if args.gpus != 1:
        sys.exit(f"bench.py runs on one GPU only (got --gpus {args.gpus}): the path is built and measured on one "
                 "GPU; how it would shard across GPUs is described in the design notes, not built")
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    # the C host prints progress messages like the reference does (printf on stdout);
    # keep stdout for the one JSON line by sending fd 1 to stderr until the end
    import ctypes
    libc = ctypes.CDLL(None)
    real_stdout = os.dup(1)
    os.dup2(2, 1)
    # the reference's CPU path, timed first: child processes started before
    # anything here has touched the GPU
    cpu = None
    if not args.no_cpu:
        if args.workload == "morpho8256_sweep81":
            cpu = cpu_reference_sweep()
        elif args.workload == "mie1056":
            log("cpu baseline (the reference, one core)")
            cpu = cpu_reference(args.mode, args.workload)
    import mpifdtd_amd as m
    L = m.lib()
    # the library's own HIP runtime does the synchronising (fdtd_hip_synchronize =
    # hipDeviceSynchronize): a GPU runtime loaded beside it would not see its queues
    assert L.fdtd_hip_device_count() >= 1, "no GPU"
    if args.workload == "morpho8256_sweep81":
        line = run_sweep81(args, m, cpu)
    elif args.workload == "sweep64":
        line = run_sweep(args, m)
    elif args.workload == "solvers_fp64":
        line = run_solvers_fp64(args, m)
    else:
        line = run_single(args, m, cpu)
    libc.fflush(None)
    sys.stdout.flush()
    os.dup2(real_stdout, 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
