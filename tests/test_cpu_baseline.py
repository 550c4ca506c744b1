"""The reference's own step as bench.py's CPU baseline (oracle/_ref/ref_harness
mode=time, built from the reference's sources where they are here): the JSON
line it prints, and the per-step NTFF of the plain TM update (ntffTM_TimeCalc,
fdtdTM_upml.c:65) timed alone with ntff=1, from which the step's rate without
it follows."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")

pytestmark = pytest.mark.skipif(not os.access(HARNESS, os.X_OK), reason="the reference harness is not built")

BASE = ["mode=time", "model=3", "h=5", "pml=10", "lambda=550", "angle=90", "warm=1", "width=1000", "height=800",
        "steps=6"]


def run(*extra):
    out = subprocess.run([HARNESS, *BASE, *extra], capture_output=True, text=True, timeout=120, cwd="/tmp")
    assert out.returncode == 0, out.stderr[-2000:]
    return json.loads(out.stdout.strip().splitlines()[-1])


def test_plain_tm_step_and_its_ntff_timed_alone():
    r = run("solver=2", "ntff=1")
    assert (r["N_PX"], r["N_PY"], r["steps"]) == (200, 160, 6)
    assert r["seconds"] > 0 and r["ntff_seconds"] > 0
    r0 = run("solver=2")
    assert r0["ntff_seconds"] == -1.0  # not asked for


def test_ntff_alone_only_for_plain_tm():
    r = run("solver=3", "ntff=1")  # plain TE: its update's NTFF is another routine
    assert r["seconds"] > 0 and r["ntff_seconds"] == -1.0
