"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer: every
oracle function once on small inputs (all solvers of enum SOLVER, square and
non-square grids, per-step NTFF, far field, spectra, eps maps, FFT) through
oracle/san_driver.c, built by `make -C oracle sanitize`; the sanitizers abort
at the first out-of-bounds access, leak or undefined operation."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("gcc") is None and shutil.which("cc") is None, reason="no C compiler")
def test_oracle_is_clean_under_asan_and_ubsan():
    b = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "sanitize"], capture_output=True, text=True,
                       timeout=300)
    assert b.returncode == 0, b.stderr[-2000:]
    r = subprocess.run([os.path.join(ROOT, "oracle", "_san", "san_driver")], capture_output=True, text=True,
                       timeout=300, env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0"))
    assert r.returncode == 0, r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr
    assert "every case ran clean" in r.stderr
