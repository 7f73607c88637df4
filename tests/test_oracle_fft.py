"""Pins the oracle FFT (the role FFTW/cuFFT play in the reference,
propagation.c:508-613) to the DFT definition and to numpy.fft, and the
device FFT's stage logic (mpc_amd/csrc/fft_core.h) to a brute-force DFT
through a CPU emulation of its threads."""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def brute_dft(x, sign):
    n = x.shape[-1]
    k = np.arange(n)
    W = np.exp(sign * 2j * np.pi * np.outer(k, k) / n)
    return x @ W.T


@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 8, 9, 12, 15, 16, 24, 27, 32, 45, 48, 64])
@pytest.mark.parametrize("sign", [-1, +1])
def test_fft_vs_brute_force(n, sign):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((5, n)) + 1j * rng.standard_normal((5, n))
    y = x.copy()
    O.oracle().oracle_fft_lines(O.vp(y), n, 1, 5, n, 1, 0, sign)
    ref = brute_dft(x, sign)
    assert np.abs(y - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("n", [64, 128, 192, 256, 512, 1024, 2048])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_fft_axis_vs_numpy(n, axis):
    """all three reference geometries (x stride 1, y stride Nx, t stride NxNy)"""
    rng = np.random.default_rng(axis * 7 + n)
    shape = [4, 4, 4]                  # (Nt, Ny, Nx)
    shape[2 - axis] = n
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    dims = O.dims_arr(a.shape)
    for sign, ref in ((-1, np.fft.fft(a, axis=2 - axis)), (+1, np.fft.ifft(a, axis=2 - axis) * n)):
        y = a.copy()
        O.oracle().oracle_fft_axis(O.vp(y), O.ip(dims), axis, sign)
        assert np.abs(y - ref).max() <= 1e-13 * np.abs(ref).max()


def test_fft_thread_count_independent():
    rng = np.random.default_rng(1)
    a = rng.standard_normal((64, 16, 16)) + 1j * rng.standard_normal((64, 16, 16))
    dims = O.dims_arr(a.shape)
    L = O.oracle()
    n0 = L.oracle_get_threads()
    outs = []
    for t in (1, 3):
        L.oracle_set_threads(t)
        y = a.copy()
        L.oracle_fft_axis(O.vp(y), O.ip(dims), 2, -1)
        outs.append(y)
    L.oracle_set_threads(n0)
    assert np.array_equal(outs[0], outs[1])


def test_device_fft_stages_cpu_emulation(tmp_path):
    """fft_core.h stage/codelet code run on the CPU thread by thread (sizes 16..4096)."""
    exe = tmp_path / "fft_emul"
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", str(exe), os.path.join(ROOT, "tests", "cpu", "fft_emul.cpp")],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0 and out.stdout.strip().endswith("OK")
