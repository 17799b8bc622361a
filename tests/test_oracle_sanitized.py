"""The oracle's CPU tests again, against a build of the oracle under
AddressSanitizer and UndefinedBehaviorSanitizer (`make -C oracle
sanitize`; any report aborts), so that an out-of-bounds index or undefined
arithmetic in the checker fails loudly (CPU only)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_TESTS = ["test_oracle.py", "test_oracle_sources.py", "test_oracle_dia.py", "test_oracle_listed.py",
                "test_oracle_pseudomomentum.py", "test_oracle_smoothing.py"]


def _asan_runtime():
    try:
        out = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True,
                             check=True).stdout.strip()
    except (OSError, subprocess.CalledProcessError):
        return None
    return out if os.path.isabs(out) and os.path.exists(out) else None


@pytest.mark.skipif(_asan_runtime() is None, reason="gcc's AddressSanitizer runtime is not installed")
def test_oracle_tests_pass_under_sanitizers():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    env = dict(os.environ)
    env["RIPPLE_ORACLE_LIB"] = os.path.join(ROOT, "oracle", "_san", "liboracle.so")
    env["LD_PRELOAD"] = _asan_runtime()
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"  # the interpreter's own allocations are not ours
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["OMP_NUM_THREADS"] = "2"
    files = [os.path.join(ROOT, "tests", f) for f in ORACLE_TESTS]
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:xdist", "-p", "no:cacheprovider",
                        *files], cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    tail = (r.stdout + r.stderr)[-4000:]
    assert r.returncode == 0, tail
    assert "ERROR: AddressSanitizer" not in tail and "runtime error" not in tail, tail
