"""CPU: the FNO's spectral convolution written the way the GPU computes it
(truncated DFTs as products with cos/sin matrices, complex recombination of
the real-block products, per-mode mixing) equals the numpy-FFT form of the
CPU reference network, in FP64."""
import numpy as np

import fno_cpu
from mpc_amd.fno import FnoConfig, dft_mats, init_params


def spectral_as_gemms(v, Rr, Ri, cfg):
    M = dft_mats(cfg)
    B, W, X, Y, T = v.shape
    mx, my, mt = cfg.mx, cfg.my, cfg.mt
    C1 = v @ M["Ft"]                                                   # (B,W,X,Y,2mt)
    P2 = np.einsum("ry,bwxyk->bwxrk", M["Gyf"], C1)                    # (B,W,X,2my,2mt)
    V2 = _comb(P2, my, -1)
    P3 = np.einsum("rx,bwxyk->bwryk", M["Gxf"], V2)                    # (B,W,2mx,my,2mt)
    V3 = _comb(np.moveaxis(P3, 2, 3), mx, -1)                         # rows kx after ky move
    V3 = np.moveaxis(V3, 3, 2)                                          # (B,W,mx,my,2mt)
    Vc = V3[..., :mt] + 1j * V3[..., mt:]
    R = (Rr + 1j * Ri).reshape(W, W, mx, my, mt)
    Uc = np.einsum("bixyk,ioxyk->boxyk", Vc, R)
    U = np.concatenate([Uc.real, Uc.imag], axis=-1)
    P5 = np.einsum("xr,bwryk->bwxyk", M["Gxi"], U)                     # (B,W,2X,my,2mt)
    V5 = np.moveaxis(_comb(np.moveaxis(P5, 2, 3), X, +1), 3, 2)       # (B,W,X,my,2mt)
    P6 = np.einsum("yr,bwxrk->bwxyk", M["Gyi"], V5)                    # (B,W,X,2Y,2mt)
    V6 = _comb(P6, Y, +1)
    return V6 @ M["Gt"]


def _comb(P, m, sign):
    """P (..., 2m, 2h) -> (..., m, 2h): the library's k_combine on one block"""
    h = P.shape[-1] // 2
    c, s = P[..., :m, :], P[..., m:, :]
    return np.concatenate([c[..., :h] - sign * s[..., h:], c[..., h:] + sign * s[..., :h]], axis=-1)


def test_gemm_form_spectral_conv_equals_fft_form():
    cfg = FnoConfig(X=16, Y=8, T=32, width=4, layers=1, mx=6, my=4, mt=7, hidden=8)
    p = init_params(cfg, seed=3)
    rng = np.random.default_rng(1)
    v = rng.standard_normal((2, cfg.width, cfg.X, cfg.Y, cfg.T))
    Rr, Ri = p["Rr0"] * 50, p["Ri0"] * 50
    a = fno_cpu.spectral(v, Rr, Ri, cfg)
    b = spectral_as_gemms(v, Rr, Ri, cfg)
    assert np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(a)


def test_spectral_conv_is_a_mode_filter():
    """with R = identity on every kept mode a kept plane wave passes unchanged
    and one outside the kept modes is removed"""
    cfg = FnoConfig(X=16, Y=16, T=32, width=2, layers=1, mx=4, my=4, mt=5, hidden=4)
    nm = cfg.mx * cfg.my * cfg.mt
    Rr = np.zeros((2, 2, nm))
    Rr[0, 0] = Rr[1, 1] = 1.0
    Ri = np.zeros_like(Rr)
    x, y, t = np.meshgrid(np.arange(16), np.arange(16), np.arange(32), indexing="ij")
    wave = np.cos(2 * np.pi * (x / 16 + 2 * t / 32))          # kx = 1, kt = 2: kept
    w2 = np.stack([wave, 0 * wave])[None]
    assert np.allclose(spectral_as_gemms(w2, Rr, Ri, cfg), w2, atol=1e-12)
    outside = np.cos(2 * np.pi * (3 * x / 16 + 2 * t / 32))   # kx = 3: not kept
    o2 = np.stack([outside, 0 * outside])[None]
    assert np.abs(spectral_as_gemms(o2, Rr, Ri, cfg)).max() < 1e-12
