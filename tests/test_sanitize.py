"""The HIP-free host code (file formats, initial conditions, status strings)
and the CPU oracle built with AddressSanitizer and UndefinedBehaviorSanitizer
and driven on valid and malformed inputs (tests/sanitize/check.cpp).  CPU
only: no GPU code is built or run."""
import os
import shutil
import subprocess

import pytest

from conftest import GOLDEN, ROOT

HERE = os.path.join(ROOT, "tests", "sanitize")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_code_under_asan_ubsan(tmp_path):
    subprocess.run(["make", "-s", "-C", HERE, f"OUT={tmp_path}"], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="1")
    r = subprocess.run([str(tmp_path / "check_sanitized"), GOLDEN, str(tmp_path)], capture_output=True, text=True,
                       env=env, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "sanitized host checks ok" in r.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_threaded_host_code_under_tsan(tmp_path):
    """The CLI's background file writer and the block-parallel snapshot
    formatter under ThreadSanitizer: jobs submitted while the writer works,
    files read back bitwise, and a failed snapshot stops every later job."""
    subprocess.run(["make", "-s", "-C", HERE, f"OUT={tmp_path}", "tsan"], check=True)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([str(tmp_path / "check_tsan"), str(tmp_path)], capture_output=True, text=True, env=env,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "threaded host checks ok" in r.stdout and "ThreadSanitizer" not in r.stderr
