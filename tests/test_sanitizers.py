"""The host code generator under AddressSanitizer + UndefinedBehaviorSanitizer and under
ThreadSanitizer (tests/native/host_check.cpp, built with g++ by `make sanitize`; no device):
LUT mapping, scheduling, register allocation with its threaded output-group refinement,
encoding, template patching and batch code objects run without a report."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "naivesat_amd", "csrc")
BUILD = os.path.join(ROOT, "_build", "csrc")


@pytest.fixture(scope="module")
def built():
    subprocess.run(["make", "-s", "-j8", "-C", CSRC, "sanitize"], check=True, capture_output=True)
    return BUILD


@pytest.mark.parametrize("variant", ["asan", "tsan"])
def test_host_code_generator_is_clean(built, variant):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", TSAN_OPTIONS="halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1")
    p = subprocess.run([os.path.join(built, f"host_check_{variant}")], capture_output=True, text=True,
                       env=env, timeout=600)
    assert p.returncode == 0 and "host_check ok" in p.stdout, p.stderr[-4000:]
    assert "runtime error" not in p.stderr and "WARNING: ThreadSanitizer" not in p.stderr, p.stderr[-4000:]
