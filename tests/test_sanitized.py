"""The CPU test suite again, against the AddressSanitizer / UBSan build of
the library's host code (kbw_amd.build.build_sanitized: the planners
kbwh_reorder_plan and kbwh_batch_plan, the plan checks, argument checking and
the C ABI), in a child process that preloads the sanitizer runtime.  A
memory error or undefined behaviour in host code aborts the child and fails
this test."""
import os
import subprocess
import sys

import pytest

from kbw_amd import build as kb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(os.environ.get("KBWH_LIB") is not None, reason="already running against another library build")
def test_cpu_suite_under_asan_and_ubsan():
    rt = kb.asan_runtime()
    if not os.path.exists(kb.SAN_LIB) or rt is None:
        pytest.skip("sanitized build or sanitizer runtime not present (built by __graft_entry__.build())")
    env = dict(os.environ, KBWH_LIB=kb.SAN_LIB, LD_PRELOAD=rt,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:symbolize=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu",
                        os.path.join(ROOT, "tests"), "--deselect",
                        "tests/test_sanitized.py::test_cpu_suite_under_asan_and_ubsan"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    tail = (r.stdout + r.stderr)[-3000:]
    assert r.returncode == 0, tail
    assert "passed" in r.stdout, tail
