"""The oracle's tests again, against a host AddressSanitizer + UndefinedBehavior-
Sanitizer build of the same source (`make -C oracle sanitize`, -O1): no
out-of-bounds access, use-after-free or undefined behaviour on any of their
inputs, and the same results as the -O3 build (the fixture tests are bitwise).
CPU only; runs in a child process that preloads the ASan runtime."""
import os
import shutil
import subprocess
import sys

import pytest

from conftest import REPO

FILES = ["test_oracle_ops.py", "test_oracle_solver.py", "test_oracle_stats.py", "test_oracle_metrics.py",
         "test_oracle_augment.py"]


def _runtime(name):
    cxx = os.environ.get("CXX", "g++")
    if shutil.which(cxx) is None:
        return None
    out = subprocess.run([cxx, f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return out if os.path.isabs(out) and os.path.exists(out) else None


def test_oracle_tests_pass_under_asan_and_ubsan():
    asan = _runtime("libasan.so")
    if asan is None or _runtime("libubsan.so") is None:
        pytest.skip("host compiler has no ASan/UBSan runtime")
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "sanitize"], check=True)
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="print_stacktrace=1",
               SVS_ORACLE_LIB=os.path.join(REPO, "oracle", "_sanitize", "libsvs_oracle.so"))
    cmd = [sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", "-m", "not gpu",
           *[os.path.join(REPO, "tests", f) for f in FILES]]
    r = subprocess.run(cmd, cwd=REPO, env=env, capture_output=True, text=True, timeout=1200)
    tail = (r.stdout + r.stderr)[-4000:]
    assert r.returncode == 0, tail
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, tail
