"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer: every
exported function runs on 4^4 and 4^3x8 lattices (oracle/sanitize_check.cpp),
so an out-of-bounds access, a use of freed memory or undefined arithmetic in
the checker itself aborts the run with the sanitizer's report.  Host code
only; the HIP library is not involved."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None, reason="needs g++ and make")
def test_oracle_runs_clean_under_asan_and_ubsan():
    build = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], capture_output=True,
                           text=True, timeout=600)
    assert build.returncode == 0, build.stderr
    env = dict(os.environ, OMP_NUM_THREADS="2", ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([os.path.join(ROOT, "_build", "sanitize", "sanitize_check")], capture_output=True,
                         text=True, timeout=600, env=env)
    assert run.returncode == 0, run.stdout + run.stderr
    assert "sanitized oracle run ok" in run.stdout
    assert "runtime error" not in run.stderr and "AddressSanitizer" not in run.stderr
