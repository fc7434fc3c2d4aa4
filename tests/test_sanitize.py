"""The oracle and the product's host-side mechanism loader and set-up
thermochemistry under AddressSanitizer and UndefinedBehaviorSanitizer
(oracle/Makefile target `sanitize`, oracle/sanitize_main.cpp): both bundled
mechanisms are parsed, turned into the oracle's tables and run through the
oracle's properties, HLLC, flux stage, chemistry batch and Robertson paths;
any sanitizer report fails the run. Host code only (the GPU kernels are not
built with sanitizers)."""
import shutil
import subprocess

import pytest

from tests.conftest import DME, H2, ROOT


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None, reason="needs g++ and make")
def test_oracle_and_host_loader_under_asan_ubsan(tmp_path):
    exe = tmp_path / "sanitize_check"
    r = subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "sanitize", f"SAN_OUT={exe}"], capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    r = subprocess.run([str(exe), str(H2), str(DME)], capture_output=True, text=True, timeout=600,
                       env={"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0", "UBSAN_OPTIONS": "print_stacktrace=1",
                            "OMP_NUM_THREADS": "1", "PATH": "/usr/bin:/bin"})
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAILED" not in r.stdout and r.stdout.count("ok ") >= 18, r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr
