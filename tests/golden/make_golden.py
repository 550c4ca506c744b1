"""Write the golden fixtures that pin the oracle, by running the reference.

Builds oracle/_ref/ref_harness (the reference's own C sources, compiled where
they lie under /root/reference, driven by oracle/ref_harness.c) and runs it on
small configurations; the outputs are stored as compressed .npz fixtures next
to this script.  Run only where /root/reference exists (the fixtures, not
the reference, go wherever the tests run):

    python tests/golden/make_golden.py [fixture names: only these]

It also runs oracle/_ref/ref_main, the reference's own main.c built from all its
sources, in an empty directory and stores the files it writes (main_default).
"""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
REF_MAIN = os.path.join(ROOT, "oracle", "_ref", "ref_main")

# (name, solver, model, sweep, width, height, h, pml, lambda, angle, steps)
RUN_CASES = [
    ("mpitm_mie_a0", 4, 1, 0, 4000, 4000, 50, 10, 500, 0, 300),
    ("mpitm_mie_a30_rect", 4, 1, 0, 4000, 3500, 50, 10, 550, 30, 250),
    ("mpite_mie_a0", 5, 1, 0, 4000, 4000, 50, 10, 500, 0, 300),
    ("mpite_mie_a60_rect", 5, 1, 0, 3500, 4000, 50, 8, 600, 60, 250),
    ("tm_mie_pulse", 2, 1, 0, 4000, 4000, 50, 10, 500, 0, 700),
    ("te_mie_pulse", 3, 1, 0, 4000, 4000, 50, 10, 500, 0, 700),
    ("mpitm_morpho", 4, 3, 0, 1600, 1800, 20, 8, 500, 0, 150),
    ("mpite_zigzag", 5, 5, 0, 1200, 1500, 20, 8, 450, 0, 150),
    ("tm2d_mie_a0", 0, 1, 0, 4000, 4000, 50, 10, 500, 0, 300),      # split-field PML TM (fdtdTM.c)
    ("tm2d_zigzag_a30_rect", 0, 5, 0, 1200, 1500, 20, 8, 450, 30, 150),
    ("te2d_mie_a0", 1, 1, 0, 4000, 4000, 50, 10, 500, 0, 300),       # split-field PML TE (fdtdTE.c)
    ("te2d_layer_a45_rect", 1, 2, 3, 1800, 2000, 20, 8, 550, 45, 150),
    ("nstm_mie_a0", 6, 1, 0, 4000, 4000, 50, 10, 500, 0, 300),        # non-standard FDTD TM (nsFdtdTM.c)
    ("nstm_zigzag_a30_rect", 6, 5, 0, 1400, 1800, 20, 8, 450, 30, 150),
    ("nste_mie_a0", 7, 1, 0, 4000, 4000, 50, 10, 500, 0, 300),        # non-standard FDTD TE (nsFdtdTE.c)
    ("nste_layer_a45_rect", 7, 2, 3, 1400, 1800, 20, 8, 450, 45, 150),
]


def field_size(model_need, h, pml):
    """calcFieldSize (main.c:71-88): model size + (pml + 5 cells) on each side + 200 nm."""
    x_nm, y_nm = model_need
    return (int(x_nm + h * (pml + 5) * 2 + 200), int(y_nm + h * (pml + 5) * 2 + 200))


# (name, model, sweep, width, height, h) -- eps maps of the model callbacks
EPS_CASES = [
    ("eps_circle_h20", 1, 0, 2920, 2920, 20),
    ("eps_circle_h7", 1, 0, 1400, 1400, 7),
    ("eps_layer_s0", 2, 0, 900, 1400, 10),
    ("eps_layer_s7", 2, 7, 900, 1500, 10),
    ("eps_zigzag_s0", 5, 0, 800, 1500, 10),
    ("eps_zigzag_s4", 5, 4, 800, 2200, 10),
    ("eps_morpho_s0", 3, 0, 900, 1600, 10),
    ("eps_morpho_s9", 3, 9, 900, 1900, 10),
]


def run(args, cwd=None):
    r = subprocess.run([HARNESS] + args, capture_output=True, text=True, cwd=cwd)
    if r.returncode != 0:
        raise RuntimeError(f"ref_harness {' '.join(args)} failed ({r.returncode}):\n{r.stdout}\n{r.stderr}")
    return r.stdout


def read_text_rows(path, rows, ncols):
    out = []
    with open(path) as f:
        for i, line in enumerate(f):
            if i in rows:
                out.append(np.array(line.split()[:ncols], dtype=np.float64))
    return np.stack(out)


def write_main_default(tmp):
    """The reference's program as shipped (main.c: MIE_CYLINDER, NS_TE_2D, h = 50 nm,
    pml 15, lambda 500 nm, 400 steps), run in an empty directory: every file it
    writes (relative paths), the ns_te_50nm.txt rows and the image.bmp bytes."""
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "_ref/ref_main"], check=True, capture_output=True)
    cwd = os.path.join(tmp, "main_default")
    os.makedirs(cwd)
    r = subprocess.run([REF_MAIN], cwd=cwd, capture_output=True, text=True, timeout=600)
    if r.returncode != 1:  # main.c ends with `return 1` (main.c:234)
        raise RuntimeError(f"ref_main failed ({r.returncode}): {r.stderr[-2000:]}")
    paths = sorted(os.path.relpath(os.path.join(dp, f), cwd) for dp, _, fs in os.walk(cwd) for f in fs)
    d = dict(paths=np.array(paths), returncode=np.array(r.returncode))
    for p in paths:
        full = os.path.join(cwd, p)
        if p.endswith(".txt"):
            d["elliptic"] = np.loadtxt(full)
        elif p.endswith(".bmp"):
            d["bmp"] = np.frombuffer(open(full, "rb").read(), np.uint8)
    d["stdout"] = np.array([ln for ln in r.stdout.splitlines() if ln.startswith(("field size", "NS TE mode"))])
    np.savez_compressed(os.path.join(HERE, "ref_main_default.npz"), **d)
    print("wrote main_default", {k: v.shape for k, v in d.items()})


def write_field_util(tmp):
    """The reference's field.h / myComplex.h helpers on the harness's fixed inputs
    (mode=util): a 90 x 100 grid (h = 20 nm), lambda 450 nm, angle 30, after 480
    field_nextStep calls (the Gaussian pulse near its peak)."""
    out = os.path.join(tmp, "field_util")
    run(["mode=util", "width=1800", "height=2000", "h=20", "pml=8", "lambda=450", "angle=30", "steps=480",
         f"out={out}"])
    meta = [int(v) for v in open(os.path.join(out, "meta.txt")).read().split()]
    npx, npy = meta[0], meta[1]
    d = dict(cfg=np.array([1800, 2000, 20, 8, 450, 30, 480]), meta=np.array(meta),
             eps=np.fromfile(os.path.join(out, "eps.bin"), np.float64).reshape(npx, npy),
             base=np.fromfile(os.path.join(out, "base.bin"), np.complex128).reshape(npx, npy),
             src=np.fromfile(os.path.join(out, "src.bin"), np.complex128).reshape(4, npx, npy),
             scal=np.fromfile(os.path.join(out, "scal.bin"), np.float64).reshape(9, 4),
             bil=np.fromfile(os.path.join(out, "bil.bin"), np.complex128),
             nrm=np.fromfile(os.path.join(out, "nrm.bin"), np.float64))
    for nm in ("ell", "all_c", "all_d"):
        d[nm] = np.array(open(os.path.join(out, f"{nm}.txt")).read())
    np.savez_compressed(os.path.join(HERE, "ref_field_util.npz"), **d)
    print("wrote field_util", {k: v.shape for k, v in d.items()})


def main(only=()):
    """Writes every fixture, or only the named ones."""
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True, capture_output=True)
    tmp = tempfile.mkdtemp(prefix="golden_")
    try:
        for (name, solver, model, sweep, w, h_nm, hu, pml, lam, ang, steps) in RUN_CASES:
            if only and name not in only:
                continue
            out = os.path.join(tmp, name)
            run([f"mode=run", f"solver={solver}", f"model={model}", f"sweep={sweep}", f"width={w}",
                 f"height={h_nm}", f"h={hu}", f"pml={pml}", f"lambda={lam}", f"angle={ang}",
                 f"steps={steps}", f"out={out}"])
            meta = open(os.path.join(out, "meta.txt")).read().split()
            npx, npy = int(meta[2]), int(meta[3])
            d = dict(cfg=np.array([solver, model, sweep, w, h_nm, hu, pml, lam, ang, steps]),
                     meta=np.array([float(v) for v in meta]))
            for k in range(3):
                d[f"f{k}"] = np.fromfile(os.path.join(out, f"f{k}.bin"), np.complex128).reshape(npx, npy)
            d["eps"] = np.fromfile(os.path.join(out, "eps.bin"), np.float64).reshape(npx, npy)
            if os.path.exists(os.path.join(out, "ff.bin")):
                d["ff"] = np.fromfile(os.path.join(out, "ff.bin"), np.complex128)
            spec = os.path.join(out, f"{ang}[deg]_380nm_700nm_b.dat")
            if os.path.exists(spec):
                full = np.fromfile(spec, np.float64).reshape(321, 360)
                d["spectra"] = full[::4]           # lambda = 380, 384, ..., 700
            for pol in ("tm", "te", "ns_tm", "ns_te"):  # the split-field and NS solvers' reset(): field_outputElliptic
                ell = os.path.join(out, f"{pol}_{hu}nm.txt")
                if os.path.exists(ell):
                    d["elliptic"] = np.loadtxt(ell)[:, 1]
            tdir = os.path.join(out, "MPI_TE_UPML")
            if os.path.isdir(tdir):
                rows = set(range(0, 360, 15))
                for nm in ("Eth", "Eph"):
                    re = read_text_rows(os.path.join(tdir, f"{nm}_r.txt"), rows, steps)
                    im = read_text_rows(os.path.join(tdir, f"{nm}_i.txt"), rows, steps)
                    d[nm] = re + 1j * im
            np.savez_compressed(os.path.join(HERE, f"ref_{name}.npz"), **d)
            print("wrote", name, {k: v.shape for k, v in d.items()})
        if not only or "main_default" in only:
            write_main_default(tmp)
        if not only or "field_util" in only:
            write_field_util(tmp)
        for (name, model, sweep, w, h_nm, hu) in EPS_CASES:
            if only and name not in only:
                continue
            out = os.path.join(tmp, name)
            run([f"mode=eps", f"model={model}", f"sweep={sweep}", f"width={w}", f"height={h_nm}",
                 f"h={hu}", "pml=10", "lambda=500", "steps=10", f"out={out}"])
            npx, npy = [int(v) for v in open(os.path.join(out, "meta.txt")).read().split()]
            d = dict(cfg=np.array([model, sweep, w, h_nm, hu]))
            for nm in ("eps_xy", "eps_dy", "eps_dx"):
                d[nm] = np.fromfile(os.path.join(out, f"{nm}.bin"), np.float64).reshape(npx, npy)
            np.savez_compressed(os.path.join(HERE, f"ref_{name}.npz"), **d)
            print("wrote", name, (npx, npy))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


if __name__ == "__main__":
    sys.exit(main(tuple(sys.argv[1:])))
