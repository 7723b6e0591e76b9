"""The oracle and the timed CPU baseline under AddressSanitizer and UndefinedBehaviorSanitizer:
their CPU tests run against the `make -C oracle sanitize` builds (any report aborts the run)."""
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ORACLE_TESTS = ["tests/test_oracle.py", "tests/test_lhb.py", "tests/test_montrealer.py", "tests/test_hermite.py"]


def _runtime(name):
    cxx = shutil.which(os.environ.get("CXX", "g++")) or shutil.which("g++")
    if not cxx:
        return None
    out = subprocess.run([cxx, "-print-file-name=" + name], capture_output=True, text=True).stdout.strip()
    return out if os.path.isabs(out) and os.path.exists(out) else None


def test_oracle_tests_pass_under_asan_and_ubsan():
    asan = _runtime("libasan.so")
    if asan is None or shutil.which("make") is None:
        pytest.skip("no g++ AddressSanitizer runtime here")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True, timeout=300)
    san = os.path.join(ROOT, "oracle", "_san")
    env = dict(os.environ,
               LD_PRELOAD=asan,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
               WALRUS_ORACLE_LIB=os.path.join(san, "libwalrus_oracle.so"),
               WALRUS_CPU_TIMING_LIB=os.path.join(san, "libcpu_timing.so"),
               OMP_NUM_THREADS="2")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-p", "no:xdist",
                        "-m", "not gpu"] + ORACLE_TESTS, cwd=ROOT, env=env, capture_output=True, text=True,
                       timeout=900)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    assert " passed" in r.stdout
