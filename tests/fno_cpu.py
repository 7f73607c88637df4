"""FP64 CPU restatement of the FNO pass surrogate (test infrastructure).

The same network as mpc_amd/fno.py, written independently with FFTs: the
spectral convolution is irfftn(R . rfftn(v) on the kept modes), the pointwise
maps are matrix products over the points.  The GPU network (DFTs as BF16 GEMMs, hand-written
backward) is checked against it: outputs to BF16 tolerance, gradients against
central finite differences of this network's loss.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import scipy.fft

from mpc_amd.fno import mode_index

THREADS = min(16, os.cpu_count() or 1)


def gelu(x):
    from scipy.special import erf
    out = np.empty_like(x)
    flat, fo = x.reshape(-1), out.reshape(-1)
    bounds = np.linspace(0, flat.size, THREADS + 1).astype(int)

    def part(i):
        a, b = bounds[i], bounds[i + 1]
        fo[a:b] = 0.5 * flat[a:b] * (1.0 + erf(flat[a:b] / np.sqrt(2.0)))
    with ThreadPoolExecutor(THREADS) as ex:
        list(ex.map(part, range(THREADS)))
    return out


def spectral(v, Rr, Ri, cfg):
    B, W = v.shape[:2]
    V = scipy.fft.rfftn(v, axes=(2, 3, 4), workers=THREADS)
    kx, ky = mode_index(cfg.X, cfg.mx), mode_index(cfg.Y, cfg.my)
    Vm = V[:, :, kx][:, :, :, ky][..., :cfg.mt]
    R = (Rr + 1j * Ri).reshape(W, W, cfg.mx, cfg.my, cfg.mt)
    Um = np.einsum("bixyk,ioxyk->boxyk", Vm, R)
    U = np.zeros_like(V)
    U[:, :, kx[:, None], ky[None, :], :cfg.mt] = Um
    return scipy.fft.irfftn(U, s=(cfg.X, cfg.Y, cfg.T), axes=(2, 3, 4), workers=THREADS)


def forward(p, cfg, x):
    def pw(Wm, v, b):
        # 1x1 channel map: (o, c) x (batch, c, points), as a BLAS product
        y = np.matmul(Wm, v.reshape(v.shape[0], v.shape[1], -1)).reshape((v.shape[0], Wm.shape[0]) + v.shape[2:])
        return y + b[None, :, None, None, None]
    v = pw(p["P"], x, p["bP"])
    for l in range(cfg.layers):
        z = spectral(v, p[f"Rr{l}"], p[f"Ri{l}"], cfg) + pw(p[f"W{l}"], v, p[f"b{l}"])
        v = gelu(z) if l < cfg.layers - 1 else z
    h = gelu(pw(p["Q1"], v, p["q1"]))
    return x[:, :cfg.cout] + pw(p["Q2"], h, p["q2"])


def loss(p, cfg, x, tgt):
    pred = forward(p, cfg, x)
    d = (pred - tgt).reshape(x.shape[0], -1)
    t = tgt.reshape(x.shape[0], -1)
    return float(np.mean(np.sum(d * d, axis=1) / np.sum(t * t, axis=1)))
