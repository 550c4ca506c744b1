"""bench.py's command-line checks, which run before anything touches the GPU:
--gpus above 1 exits with its message, and R = 18 / 20 (72- and 80-row interior
tiles, which exist only as 8-wavefront bodies) is refused with a launch shape
that has no such tiles."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_exits_with_its_message():
    r = run("--gpus", "2", "--no-cpu")
    assert r.returncode != 0 and "one GPU only" in r.stderr


def test_tall_interior_tiles_need_an_eight_wavefront_shape():
    for R in ("18", "20"):
        r = run("--workload", "mie1056", "--R", R, "--launch", "2", "--no-cpu")
        assert r.returncode != 0 and f"--R {R}" in r.stderr and "needs --launch 6, 7, 8 or 9" in r.stderr
        r = run("--workload", "mie1056", "--R", R, "--no-cpu")  # the workload's default shape: not assumed
        assert r.returncode != 0 and "needs --launch" in r.stderr
