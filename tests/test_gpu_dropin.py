"""The reference's own program on the library.

oracle/_ref/main_dropin is the reference's main.c (it hardcodes MIE_CYLINDER and
NS_TE_2D, main.c:155-156; h = 50 nm, pml 15, lambda 500 nm, 400 steps,
main.c:89-107) compiled with parser.c, function.c and drawer.c only: every other
symbol, from simulator_setSolver to the solver's update, comes from
libmpifdtd_amd.so.  Run in an empty directory it must write what the reference
program built from all its sources writes (fixture ref_main_default): the same
directory tree, the same image.bmp bytes (drawn from getEps), ns_te_50nm.txt
within its "%lf" print resolution, and the same exit status (main.c ends with
`return 1`).
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_golden

DROPIN = os.path.join(ROOT, "oracle", "_ref", "main_dropin")


@pytest.mark.gpu
def test_reference_main_runs_unchanged_on_the_library(tmp_path):
    if not os.path.exists(DROPIN):
        pytest.skip("oracle/_ref/main_dropin is built only where the reference sources exist")
    g = load_golden("main_default")
    print("case main_dropin: MIE_CYLINDER, NS_TE_2D, 90x90 cells, 400 steps", flush=True)
    r = subprocess.run([DROPIN], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == int(g["returncode"]), (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    for line in g["stdout"]:
        assert line in r.stdout.splitlines(), line
    paths = sorted(os.path.relpath(os.path.join(dp, f), tmp_path) for dp, _, fs in os.walk(tmp_path) for f in fs)
    assert paths == list(g["paths"]), paths
    for p in paths:
        full = tmp_path / p
        if p.endswith(".bmp"):
            assert np.array_equal(np.frombuffer(full.read_bytes(), np.uint8), g["bmp"]), p
        elif p.endswith(".txt"):
            rows = np.loadtxt(full)
            assert np.array_equal(rows[:, 0], g["elliptic"][:, 0])
            assert np.abs(g["elliptic"][:, 1]).max() > 1e-3
            assert np.max(np.abs(rows[:, 1] - g["elliptic"][:, 1])) <= 1.0000001e-6, p
