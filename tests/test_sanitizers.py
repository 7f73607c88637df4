"""CPU builds of the oracle and of the device FFT's thread-by-thread
emulation under AddressSanitizer and UndefinedBehaviorSanitizer (host code
only, gcc/g++): every oracle entry point on a small intense case in both
arithmetic profiles, and every FFT plan against the long-double DFT, must run
clean (any report aborts the program)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0:abort_on_error=1",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="4")

pytestmark = pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("g++") is None,
                                reason="needs gcc/g++")


def _build_and_run(cmd, exe, *args):
    b = subprocess.run(cmd, capture_output=True, text=True)
    assert b.returncode == 0, b.stderr[-3000:]
    r = subprocess.run([exe, *args], capture_output=True, text=True, env=ENV, timeout=600)
    out = r.stdout + r.stderr
    print(out[-2000:])
    assert r.returncode == 0, out[-4000:]
    assert "runtime error" not in out and "AddressSanitizer" not in out, out[-4000:]
    return out


def test_oracle_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "oracle_sanitize")
    src = [os.path.join(ROOT, "oracle", f) for f in ("fft.c", "plasma.c", "propagation.c")]
    cmd = ["gcc", *SAN, "-std=c11", "-D_GNU_SOURCE", "-fopenmp", "-ffp-contract=off", "-o", exe,
           os.path.join(ROOT, "tests", "cpu", "oracle_sanitize.c"), *src, "-lm"]
    out = _build_and_run(cmd, exe)
    assert "finite 1" in out


def test_fft_emulation_under_asan_ubsan(tmp_path):
    """every radix codelet, LDS layout and the cached-twiddle form, at sizes up
    to 1024 (the larger ones add only more of the same stages, and their
    brute-force reference is slow under the sanitizers)"""
    exe = str(tmp_path / "fft_emul")
    cmd = ["g++", *SAN, "-std=c++17", "-o", exe, os.path.join(ROOT, "tests", "cpu", "fft_emul.cpp")]
    out = _build_and_run(cmd, exe, "1024")
    assert out.strip().endswith("OK")
