"""The CPU suite against an AddressSanitizer + UndefinedBehaviorSanitizer
build of the GL host layer and the oracle pipeline (`make -C oracle
sanitize`): the same tests, run in a child process with the ASan runtime
preloaded and MTGL_ORACLE_LIB pointing at the sanitized library.  Any
out-of-bounds or freed access, or undefined behaviour, ends the child with
the sanitizer's report and fails this test."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN_LIB = os.path.join(ROOT, "oracle", "_build", "liboracle_gl_san.so")


def _runtime():
    try:
        p = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True, check=True)
    except (OSError, subprocess.CalledProcessError):
        return None
    path = p.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


def test_cpu_suite_under_asan_and_ubsan():
    asan = _runtime()
    if not asan:
        pytest.skip("gcc's AddressSanitizer runtime is not installed")
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True, capture_output=True)
    env = dict(os.environ, MTGL_ORACLE_LIB=SAN_LIB, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    tests = [os.path.join(ROOT, "tests", t) for t in (
        "test_oracle_golden.py", "test_oracle_closed_forms.py", "test_oracle_reference.py", "test_host_errors.py",
        "test_diff_oracle.py", "test_batch_oracle.py", "test_list_segments.py", "test_bench_cpu_legs.py")]
    p = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider",
                        *tests], cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    out = p.stdout + p.stderr
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert p.returncode == 0, out[-4000:]
