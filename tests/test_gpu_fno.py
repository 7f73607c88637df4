"""GPU: the FNO surrogate's kernels and network.

* mpc_gemm_bf16 against the FP64 product of the BF16-rounded operands, over
  transposed operands, two-level batches, alpha/beta and split K (FP32
  accumulation: 1e-4 of the result's norm).
* k_combine forward against numpy and its bwd map as the exact transpose.
* the fused pointwise channel map (bias, residual, GELU, pre-activation
  store, transposed weights) and the per-mode complex mixing with its
  backward against numpy (FP32: 1e-5 of the largest term).
* the network's forward against the FP64 CPU network (fno_cpu) to BF16
  tolerance (3e-2), and the hand-written backward's gradient of every
  parameter tensor against central finite differences of the CPU loss along
  a random direction (5e-2 of the larger of the difference quotient and the
  gradient's norm).
* Adam on the GPU lowers the loss of a learnable map.
"""
import numpy as np
import pytest

import fno_cpu

pytestmark = pytest.mark.gpu


def _bf16(a):
    """round float32 values to BF16 (nearest even), kept as float64"""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _t(a):
    from mpc_amd.fno import DArr
    return DArr.of(a)


@pytest.mark.parametrize("M,N,K,ta,tb", [(64, 64, 32, False, False), (70, 33, 45, True, False),
                                         (17, 130, 64, False, True), (100, 90, 200, True, True),
                                         (32, 32, 8192, False, True),
                                         # ragged M and N, aligned unit-stride rows
                                         (40, 96, 100, False, False)])
def test_gemm_matches_bf16_product(gpu_lib, M, N, K, ta, tb):
    from mpc_amd.fno import Ops
    print(f"case gemm M={M} N={N} K={K} ta={ta} tb={tb}", flush=True)
    rng = np.random.default_rng(M * 7 + N + K)
    nb1, nb2 = (2, 3) if K < 1000 else (1, 1)
    A = rng.standard_normal((nb1, nb2, M, K)).astype(np.float32)
    B = rng.standard_normal((nb1, nb2, K, N)).astype(np.float32)
    C0 = rng.standard_normal((nb1, nb2, M, N)).astype(np.float32)
    alpha, beta = 0.75, 0.5
    ref = alpha * np.einsum("abik,abkj->abij", _bf16(A), _bf16(B)) + beta * C0
    As = np.swapaxes(A, 2, 3).copy() if ta else A       # stored transposed: element (i,k) at k*M + i
    Bs = np.swapaxes(B, 2, 3).copy() if tb else B
    dA, dB, dC = _t(As), _t(Bs), _t(C0)
    rsA, csA = (1, M) if ta else (K, 1)
    rsB, csB = (1, K) if tb else (N, 1)
    ops = Ops()
    ops.gemm(M, N, K, dA, rsA, csA, dB, rsB, csB, dC, N, 1, alpha=alpha, beta=beta,
             b1=(nb1, nb2 * M * K, nb2 * K * N, nb2 * M * N), b2=(nb2, M * K, K * N, M * N))
    got = dC.numpy().astype(np.float64)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"  rel err {err:.2e}", flush=True)
    assert err <= 1e-4


def test_combine_and_its_transpose(gpu_lib):
    from mpc_amd.fno import DArr, Ops
    rng = np.random.default_rng(5)
    ns, m, nb, h = 3, 4, 5, 6
    P = rng.standard_normal((ns, 2 * m, nb, 2 * h)).astype(np.float32)
    dV = rng.standard_normal((ns, m, nb, 2 * h)).astype(np.float32)
    ops = Ops()
    for sign in (-1, 1):
        V = DArr((ns, m, nb, 2 * h))
        ops.combine(_t(P), V, ns, m, nb, h, sign)
        dP = DArr((ns, 2 * m, nb, 2 * h))
        ops.combine(_t(dV), dP, ns, m, nb, h, sign, bwd=True)
        c, s = P[:, :m], P[:, m:]
        ref = np.concatenate([c[..., :h] - sign * s[..., h:], c[..., h:] + sign * s[..., :h]], axis=-1)
        assert np.allclose(V.numpy(), ref, atol=1e-6)
        lhs = float(np.sum(V.numpy().astype(np.float64) * dV))
        rhs = float(np.sum(P.astype(np.float64) * dP.numpy()))
        assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def _gelu(z):
    from scipy.special import erf
    return 0.5 * z * (1 + erf(z / np.sqrt(2)))


@pytest.mark.parametrize("cin,cout,act,npts", [
    (4, 32, False, 1000), (32, 32, True, 1000), (64, 2, False, 1000), (2, 64, True, 1000),
    # matrix-core widths: two output tiles per wave, ragged and short point
    # counts, three output groups
    (16, 64, True, 1000), (64, 32, False, 77), (32, 64, True, 20), (16, 96, False, 4096)])
def test_pointwise_channel_map(gpu_lib, cin, cout, act, npts):
    """fused 1x1 map against numpy: plain, transposed weights, residual with
    its own batch stride, pre-activation store, in-place output"""
    from mpc_amd.fno import DArr, Ops
    print(f"case pointwise cin={cin} cout={cout} act={act} npts={npts}", flush=True)
    rng = np.random.default_rng(cin * 100 + cout)
    nb = 3
    x = rng.standard_normal((nb, cin, npts)).astype(np.float32)
    Wm = rng.standard_normal((cout, cin)).astype(np.float32) / np.sqrt(cin)
    bias = rng.standard_normal(cout).astype(np.float32)
    zin = rng.standard_normal((nb, cout + 1, npts)).astype(np.float32)   # batch stride (cout+1)*npts
    ops = Ops()
    f = (lambda z: _gelu(z)) if act else (lambda z: z)
    z_ref = np.einsum("oi,bip->bop", Wm.astype(np.float64), x) + 0.5 * zin[:, :cout] + bias[None, :, None]
    out, zs = DArr((nb, cout, npts)), DArr((nb, cout, npts))
    ops.pointwise(_t(x), cin, _t(Wm), cin, 1, out, cout, nb, npts, bias=_t(bias), zin=_t(zin), beta=0.5, zs=zs,
                  act=act, zin_bs=(cout + 1) * npts)
    assert np.abs(zs.numpy() - z_ref).max() <= 1e-5 * np.abs(z_ref).max()
    assert np.abs(out.numpy() - f(z_ref)).max() <= 1e-5 * np.abs(z_ref).max()
    if cout > 64:       # the map takes at most 64 input channels
        return
    # transposed weights (the backward's W^T: weight of input o for output i
    # at o*cin + i), no bias, output in place of zin
    y = rng.standard_normal((nb, cout, npts)).astype(np.float32)
    acc = rng.standard_normal((nb, cin, npts)).astype(np.float32)
    d = _t(acc)
    ops.pointwise(_t(y), cout, _t(Wm), 1, cin, d, cin, nb, npts, zin=d, beta=1.0)
    ref = np.einsum("oi,bop->bip", Wm.astype(np.float64), y) + acc
    assert np.abs(d.numpy() - ref).max() <= 1e-5 * np.abs(ref).max()
    WT = DArr((cin, cout))
    ops.transpose(_t(Wm), WT, cout, cin)
    assert np.array_equal(WT.numpy(), Wm.astype(np.float32).T)


def test_mode_mixing_and_its_adjoint(gpu_lib):
    """per-mode complex mixing against numpy; backward as the exact adjoint
    (<U(V), dU> = <V, dV>) and the weight gradient against numpy, with more
    samples than one launch takes"""
    from mpc_amd.fno import DArr, Ops
    rng = np.random.default_rng(3)
    nb, w, nxy, mt = 10, 6, 5, 7
    V = rng.standard_normal((nb, w, nxy, 2 * mt)).astype(np.float32)
    dU = rng.standard_normal((nb, w, nxy, 2 * mt)).astype(np.float32)
    Rr = rng.standard_normal((w, w, nxy, mt)).astype(np.float32)
    Ri = rng.standard_normal((w, w, nxy, mt)).astype(np.float32)
    ops = Ops()
    U = DArr((nb, w, nxy, 2 * mt))
    ops.mix(_t(V), _t(Rr), _t(Ri), U, nb, w, nxy, mt)
    Vc = V[..., :mt].astype(np.float64) + 1j * V[..., mt:]
    R = Rr.astype(np.float64) + 1j * Ri
    Uc = np.einsum("bikm,iokm->bokm", Vc, R)
    got = U.numpy()
    assert np.abs(got[..., :mt] - Uc.real).max() <= 1e-5 * np.abs(Uc).max()
    assert np.abs(got[..., mt:] - Uc.imag).max() <= 1e-5 * np.abs(Uc).max()
    g0 = rng.standard_normal(Rr.shape).astype(np.float32)
    dV, dRr, dRi = DArr(V.shape), _t(g0), _t(g0)
    ops.mix_bwd(_t(V), _t(dU), _t(Rr), _t(Ri), dV, dRr, dRi, nb, w, nxy, mt)
    lhs = float(np.sum(got.astype(np.float64) * dU))
    rhs = float(np.sum(V.astype(np.float64) * dV.numpy()))
    assert abs(lhs - rhs) <= 1e-5 * np.abs(got * dU).sum()
    dUc = dU[..., :mt].astype(np.float64) + 1j * dU[..., mt:]
    dR = np.einsum("bikm,bokm->iokm", np.conj(Vc), dUc)
    assert np.abs(dRr.numpy() - g0 - dR.real).max() <= 1e-4 * np.abs(dR).max()
    assert np.abs(dRi.numpy() - g0 - dR.imag).max() <= 1e-4 * np.abs(dR).max()


def _small():
    from mpc_amd.fno import FnoConfig, init_params
    cfg = FnoConfig(X=16, Y=8, T=32, width=8, layers=2, mx=4, my=4, mt=6, hidden=16)
    p = init_params(cfg, seed=4)
    for l in range(cfg.layers):        # spectral weights of a size that matters next to the pointwise path
        p[f"Rr{l}"] *= 20
        p[f"Ri{l}"] *= 20
    rng = np.random.default_rng(9)
    x = rng.standard_normal((2, cfg.cin, cfg.X, cfg.Y, cfg.T))
    tgt = 0.8 * x[:, :2] + 0.1 * np.roll(x[:, :2], 1, axis=-1)
    return cfg, p, x, tgt


def test_forward_matches_cpu_network(gpu_lib):
    from mpc_amd.fno import FnoGpu
    cfg, p, x, _ = _small()
    net = FnoGpu(cfg, p)
    got = net.forward(_t(x)).numpy()
    ref = fno_cpu.forward(p, cfg, x)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref - x[:, :2])
    print(f"case fno forward: rel err of the change {err:.2e}", flush=True)
    assert err <= 3e-2


def test_backward_matches_finite_differences(gpu_lib):
    from mpc_amd.fno import FnoGpu
    cfg, p, x, tgt = _small()
    net = FnoGpu(cfg, p)
    loss_gpu = net.loss_and_grad(_t(x), _t(tgt))
    grads = net.grads()
    loss_cpu = fno_cpu.loss(p, cfg, x, tgt)
    print(f"case fno gradients: loss gpu {loss_gpu:.6f} cpu {loss_cpu:.6f}", flush=True)
    assert abs(loss_gpu - loss_cpu) <= 2e-2 * loss_cpu
    rng = np.random.default_rng(11)
    worst = 0.0
    for k in p:
        d = rng.standard_normal(p[k].shape)
        eps = 1e-4 * max(1.0, float(np.abs(p[k]).max())) / max(1.0, float(np.abs(d).max()))
        pp = {kk: vv.copy() for kk, vv in p.items()}
        pm = {kk: vv.copy() for kk, vv in p.items()}
        pp[k] = p[k] + eps * d
        pm[k] = p[k] - eps * d
        fd = (fno_cpu.loss(pp, cfg, x, tgt) - fno_cpu.loss(pm, cfg, x, tgt)) / (2 * eps)
        an = float(np.sum(grads[k] * d))
        # measured against the larger of |fd| and |g_k|, the rms of <g_k, d>
        # over random directions d: a small projection is held to the
        # gradient's own scale, not to its accidental smallness
        rel = abs(an - fd) / max(abs(fd), float(np.linalg.norm(grads[k])), 1e-12)
        worst = max(worst, rel)
        print(f"  {k:4s} fd {fd:+.6e} gpu {an:+.6e} rel {rel:.2e}", flush=True)
        assert rel <= 5e-2, k
    print(f"  worst {worst:.2e}", flush=True)


def test_adam_training_lowers_the_loss(gpu_lib):
    from mpc_amd.fno import FnoGpu
    cfg, p, x, tgt = _small()
    net = FnoGpu(cfg, p)
    X, T = _t(x), _t(tgt)
    first = net.loss_and_grad(X, T)
    for _ in range(60):
        net.adam_step(lr=3e-3)
        last = net.loss_and_grad(X, T)
    print(f"case fno adam: loss {first:.4f} -> {last:.4f}", flush=True)
    assert last <= 0.5 * first


def _configured():
    """the configured network (64x64x256, width 32, 4 layers, 16x16x32
    modes) on a smooth input whose modes all lie inside the kept window, with
    the spectral weights scaled so the spectral path carries about as much as
    the 1x1 path"""
    from mpc_amd.fno import FnoConfig, init_params
    import scipy.fft
    cfg = FnoConfig()
    p = init_params(cfg, seed=4)
    for l in range(cfg.layers):
        p[f"Rr{l}"] *= 300
        p[f"Ri{l}"] *= 300
    rng = np.random.default_rng(21)
    spec = np.zeros((1, 2, cfg.X, cfg.Y, cfg.T // 2 + 1), dtype=np.complex128)
    for kx in list(range(0, 6)) + list(range(cfg.X - 5, cfg.X)):
        for ky in list(range(0, 6)) + list(range(cfg.Y - 5, cfg.Y)):
            spec[:, :, kx, ky, :12] = rng.standard_normal((1, 2, 12)) + 1j * rng.standard_normal((1, 2, 12))
    f = scipy.fft.irfftn(spec, s=(cfg.X, cfg.Y, cfg.T), axes=(2, 3, 4))
    f /= f.std()
    x = np.concatenate([f, np.full((1, 1, cfg.X, cfg.Y, cfg.T), 0.3), np.full((1, 1, cfg.X, cfg.Y, cfg.T), -0.5)], axis=1)
    tgt = 0.8 * x[:, :2] + 0.1 * np.roll(x[:, :2], 3, axis=-1)
    return cfg, p, x, tgt


def test_configured_network_matches_cpu_network(gpu_lib):
    """64x64x256, width 32: forward against the FP64 CPU network (BF16
    tolerance on the change the network makes), and the analytic gradient
    against central finite differences of the CPU loss along two random
    directions: all parameters at once, and the last layer's spectral
    weights alone"""
    from mpc_amd.fno import FnoGpu
    cfg, p, x, tgt = _configured()
    net = FnoGpu(cfg, p)
    got = net.forward(_t(x)).numpy()
    print("case fno configured size: CPU forward", flush=True)
    ref = fno_cpu.forward(p, cfg, x)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref - x[:, :2])
    print(f"  forward rel err of the change {err:.2e}", flush=True)
    assert err <= 3e-2
    loss_gpu = net.loss_and_grad(_t(x), _t(tgt))
    grads = net.grads()
    rng = np.random.default_rng(23)
    for label, keys in (("all parameters", list(p)), ("Rr3, Ri3", ["Rr3", "Ri3"])):
        d = {k: rng.standard_normal(p[k].shape) * (float(np.abs(p[k]).std()) or 1.0) for k in keys}
        eps = 1e-4
        pp = {k: (v + eps * d[k] if k in d else v) for k, v in p.items()}
        pm = {k: (v - eps * d[k] if k in d else v) for k, v in p.items()}
        lp = fno_cpu.loss(pp, cfg, x, tgt)
        lm = fno_cpu.loss(pm, cfg, x, tgt)
        fd = (lp - lm) / (2 * eps)
        an = float(sum(np.sum(grads[k] * d[k]) for k in keys))
        scale = np.sqrt(sum(float(np.sum((grads[k] * np.abs(p[k]).std()) ** 2)) for k in keys))
        rel = abs(an - fd) / max(abs(fd), scale, 1e-12)
        print(f"  {label}: loss gpu {loss_gpu:.6f} fd {fd:+.6e} gpu {an:+.6e} rel {rel:.2e}", flush=True)
        assert rel <= 5e-2, label
