"""Fourier-neural-operator pass surrogate (no counterpart in the reference).

The network maps the field at one mirror plane, downsampled to X x Y x T, to
the field after one cell pass.  Input channels: Re A, Im A (scaled by the
pulse's own peak), the pulse energy and the gas pressure as constant
channels.  Layers:

    v0 = P x + bP                                       (lift, 4 -> W channels)
    v_{l+1} = act(S_l(v_l) + W_l v_l + b_l)             (L Fourier layers; no act on the last)
    S_l(v)  = irfftn( R_l . rfftn(v) restricted to modes Kx x Ky x Kt )
    y = Q2 GELU(Q1 v_L + q1) + q2                       (project, W -> H -> 2)
    pred = x[:, :2] + y                                 (the network learns the change over a pass)

Every dense contraction -- the truncated DFTs along t, y and x written as
real GEMMs against cos/sin matrices, the per-mode complex channel mixing,
the pointwise channel maps and every transpose of these in the backward
pass -- runs through mpc_gemm_bf16 (BF16 operands on the matrix cores, FP32
accumulation).  The backward pass is written out by hand below; there is no
autograd.  Device buffers come from the library (mpc_dev_*), and every kernel
runs on the default stream.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from ._lib import check, lib


class MpcGemm(C.Structure):
    _fields_ = [("M", C.c_int), ("N", C.c_int), ("K", C.c_int), ("nb1", C.c_int), ("nb2", C.c_int),
                ("ksplit", C.c_int), ("A", C.c_void_p), ("B", C.c_void_p), ("C", C.c_void_p), ("W", C.c_void_p),
                ("rsA", C.c_long), ("csA", C.c_long), ("rsB", C.c_long), ("csB", C.c_long),
                ("rsC", C.c_long), ("csC", C.c_long),
                ("sA1", C.c_long), ("sB1", C.c_long), ("sC1", C.c_long),
                ("sA2", C.c_long), ("sB2", C.c_long), ("sC2", C.c_long),
                ("alpha", C.c_float), ("beta", C.c_float)]


class MpcPw(C.Structure):
    _fields_ = [("nb", C.c_int), ("cin", C.c_int), ("cout", C.c_int), ("np", C.c_long),
                ("x", C.c_void_p), ("x_bs", C.c_long), ("w", C.c_void_p), ("ldo", C.c_long), ("ldi", C.c_long),
                ("bias", C.c_void_p), ("zin", C.c_void_p), ("zin_bs", C.c_long), ("beta", C.c_float),
                ("zs", C.c_void_p), ("out", C.c_void_p), ("act", C.c_int)]


MIX_BMAX = 8          # samples per mode-mixing launch


class DArr:
    """A float32 device array owned by this object (mpc_dev_alloc)."""

    def __init__(self, shape):
        self.shape = tuple(int(d) for d in shape)
        self.size = int(np.prod(self.shape))
        self._p = C.c_void_p()
        check(lib().mpc_dev_alloc(4 * max(1, self.size), C.byref(self._p)), "mpc_dev_alloc")

    @property
    def ptr(self):
        return self._p.value

    def numel(self):
        return self.size

    @classmethod
    def of(cls, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        d = cls(a.shape)
        check(lib().mpc_dev_upload(C.c_void_p(d.ptr), a.ctypes.data_as(C.c_void_p), 4 * a.size), "mpc_dev_upload")
        return d

    def numpy(self):
        out = np.empty(self.shape, dtype=np.float32)
        lib().mpc_dev_sync()
        check(lib().mpc_dev_download(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), 4 * self.size),
              "mpc_dev_download")
        return out

    def fill(self, v=0.0):
        check(lib().mpc_dev_fill(C.c_void_p(self.ptr), self.size, v, None), "mpc_dev_fill")
        return self

    def copy_from(self, src, count, dst_off=0, src_off=0):
        check(lib().mpc_dev_copy(C.c_void_p(self.ptr + 4 * dst_off), C.c_void_p(src.ptr + 4 * src_off), 4 * count,
                                 None), "mpc_dev_copy")

    def __del__(self):
        try:
            if self._p:
                lib().mpc_dev_free(self._p)
                self._p = C.c_void_p()
        except Exception:
            pass


def zeros(shape):
    return DArr(shape).fill(0.0)


def sync():
    check(lib().mpc_dev_sync(), "mpc_dev_sync")


def _ptr(t, off=0):
    return C.c_void_p(t.ptr + 4 * off)


def _stream():
    return None


class Ops:
    """Thin wrappers over the library's FNO kernels (default stream)."""

    def __init__(self):
        if lib().mpc_device_count() < 1:
            raise RuntimeError("the FNO surrogate runs on the GPU only (no device visible)")
        if lib().mpc_sizeof_gemm() != C.sizeof(MpcGemm):
            raise RuntimeError("mpc_gemm layout mismatch between library and binding")
        if lib().mpc_sizeof_pw() != C.sizeof(MpcPw):
            raise RuntimeError("mpc_pw layout mismatch between library and binding")
        self.ws = DArr((1 << 22,))                # split-K workspace (16 MiB)
        self.gemm_flops = 0                       # 2 M N K summed over the matrix-core products issued

    def gemm(self, M, N, K, A, rsA, csA, B, rsB, csB, Cm, rsC, csC, alpha=1.0, beta=0.0,
             b1=(1, 0, 0, 0), b2=(1, 0, 0, 0), offA=0, offB=0, offC=0):
        """C = alpha A B + beta C; A, B, C device arrays with element strides;
        b1 / b2 = (count, strideA, strideB, strideC)"""
        g = MpcGemm()
        g.M, g.N, g.K = M, N, K
        g.nb1, g.nb2 = b1[0], b2[0]
        g.A, g.B, g.C = _ptr(A, offA), _ptr(B, offB), _ptr(Cm, offC)
        g.rsA, g.csA, g.rsB, g.csB, g.rsC, g.csC = rsA, csA, rsB, csB, rsC, csC
        g.sA1, g.sB1, g.sC1 = b1[1], b1[2], b1[3]
        g.sA2, g.sB2, g.sC2 = b2[1], b2[2], b2[3]
        g.alpha, g.beta = alpha, beta
        # long reductions with few output tiles: split K over workgroups
        tiles = ((M + 63) // 64) * ((N + 63) // 64) * b1[0] * b2[0]
        ks = 1
        if tiles < 512 and K >= 4096:
            ks = min(max(1, 2048 // tiles), max(1, K // 1024), max(1, self.ws.numel() // (M * N * b1[0] * b2[0])))
        g.ksplit = ks
        g.W = _ptr(self.ws) if ks > 1 else None
        self.gemm_flops += 2 * M * N * K * b1[0] * b2[0]
        check(lib().mpc_gemm_bf16(C.cast(C.pointer(g), C.c_void_p), _stream()), "mpc_gemm_bf16")

    def combine(self, src, dst, nslice, m, nb, h, sign, bwd=False):
        check(lib().mpc_fno_combine(_ptr(src), _ptr(dst), nslice, m, nb, h, sign, 1 if bwd else 0, _stream()),
              "mpc_fno_combine")

    def pointwise(self, x, cin, w, ldo, ldi, out, cout, nb, npts, bias=None, zin=None, beta=0.0, zs=None,
                  act=False, x_bs=None, zin_bs=None):
        """out[b][o] = act(sum_i w[o*ldo + i*ldi] x[b][i] + beta zin[b][o] + bias[o]) per point; zs gets the
        pre-activation"""
        a = MpcPw()
        a.nb, a.cin, a.cout, a.np = nb, cin, cout, npts
        a.x, a.x_bs = _ptr(x), cin * npts if x_bs is None else x_bs
        a.w, a.ldo, a.ldi = _ptr(w), ldo, ldi
        a.bias = _ptr(bias) if bias is not None else None
        a.zin = _ptr(zin) if zin is not None else None
        a.zin_bs = cout * npts if zin_bs is None else zin_bs
        a.beta = beta
        a.zs = _ptr(zs) if zs is not None else None
        a.out = _ptr(out)
        a.act = 1 if act else 0
        check(lib().mpc_fno_pointwise(C.cast(C.pointer(a), C.c_void_p), _stream()), "mpc_fno_pointwise")

    def transpose(self, src, dst, rows, cols):
        check(lib().mpc_fno_transpose(_ptr(src), _ptr(dst), rows, cols, _stream()), "mpc_fno_transpose")

    def mix(self, V, Rr, Ri, U, nb, w, nxy, mt):
        """U = V R per mode (complex), in launches of at most MIX_BMAX samples"""
        per = w * nxy * 2 * mt
        for b0 in range(0, nb, MIX_BMAX):
            n = min(MIX_BMAX, nb - b0)
            check(lib().mpc_fno_mix(_ptr(V, b0 * per), _ptr(Rr), _ptr(Ri), _ptr(U, b0 * per), n, w, nxy, mt,
                                    _stream()), "mpc_fno_mix")

    def mix_bwd(self, V, dU, Rr, Ri, dV, dRr, dRi, nb, w, nxy, mt):
        """dV = dU conj(R) per mode; dR += conj(V) dU summed over samples"""
        per = w * nxy * 2 * mt
        for b0 in range(0, nb, MIX_BMAX):
            n = min(MIX_BMAX, nb - b0)
            check(lib().mpc_fno_mix_bwd(_ptr(V, b0 * per), _ptr(dU, b0 * per), _ptr(Rr), _ptr(Ri),
                                        _ptr(dV, b0 * per), _ptr(dRr), _ptr(dRi), n, w, nxy, mt, _stream()),
                  "mpc_fno_mix_bwd")

    def bias_act(self, z, bias, out, nb, nc, npts, act):
        check(lib().mpc_fno_bias_act(_ptr(z), _ptr(bias) if bias is not None else None, _ptr(out), nb, nc, npts,
                                     1 if act else 0, _stream()), "mpc_fno_bias_act")

    def act_bwd(self, z, dout, dz, act):
        check(lib().mpc_fno_act_bwd(_ptr(z), _ptr(dout), _ptr(dz), z.numel(), 1 if act else 0, _stream()),
              "mpc_fno_act_bwd")

    def bias_grad(self, d, db, nb, nc, npts, accumulate=False):
        check(lib().mpc_fno_bias_grad(_ptr(d), _ptr(db), nb, nc, npts, 1 if accumulate else 0, _stream()),
              "mpc_fno_bias_grad")

    def adam(self, w, g, m, v, lr, b1, b2, eps, wd):
        check(lib().mpc_fno_adam(_ptr(w), _ptr(g), _ptr(m), _ptr(v), w.numel(), lr, b1, b2, eps, wd, _stream()),
              "mpc_fno_adam")

    def rel_l2(self, pred, tgt, dpred, loss):
        check(lib().mpc_fno_rel_l2(_ptr(pred), _ptr(tgt), _ptr(dpred), _ptr(loss), pred.shape[0],
                                   pred.size // pred.shape[0], _stream()), "mpc_fno_rel_l2")


@dataclass
class FnoConfig:
    X: int = 64
    Y: int = 64
    T: int = 256
    width: int = 32       # the configured surrogate; 16 trains faster for the same held-out error
    layers: int = 4
    mx: int = 16          # x modes: mx/2 lowest positive and mx/2 negative
    my: int = 16
    mt: int = 32          # rfft modes 0 .. mt-1 along t
    hidden: int = 64      # projection width
    cin: int = 4          # Re, Im, energy, pressure
    cout: int = 2


def mode_index(n, m):
    """FFT-order indices of the m kept modes of an n-point full transform"""
    h = m // 2
    return np.concatenate([np.arange(h), np.arange(n - h, n)])


def dft_mats(cfg: FnoConfig):
    """cos/sin matrices of the truncated transforms (float64, host)"""
    X, Y, T, mx, my, mt = cfg.X, cfg.Y, cfg.T, cfg.mx, cfg.my, cfg.mt
    t = np.arange(T)
    kt = np.arange(mt)
    th = 2 * np.pi * np.outer(t, kt) / T                       # T x mt
    Ft = np.concatenate([np.cos(th), -np.sin(th)], axis=1)     # T x 2mt : rfft, [re | im]
    w = np.where(kt == 0, 1.0, 2.0)[:, None]
    Gt = np.concatenate([w * np.cos(th.T), -w * np.sin(th.T)], axis=0) / (X * Y * T)   # 2mt x T : irfft

    def fwd(n, m):
        k = mode_index(n, m)
        a = 2 * np.pi * np.outer(k, np.arange(n)) / n         # m x n
        return np.concatenate([np.cos(a), np.sin(a)], axis=0)   # 2m x n, combined with sign -1

    def inv(n, m):
        k = mode_index(n, m)
        a = 2 * np.pi * np.outer(np.arange(n), k) / n         # n x m
        return np.concatenate([np.cos(a), np.sin(a)], axis=0)   # 2n x m, combined with sign +1

    return {"Ft": Ft, "Gt": Gt, "Gyf": fwd(Y, my), "Gxf": fwd(X, mx), "Gyi": inv(Y, my), "Gxi": inv(X, mx)}


def init_params(cfg: FnoConfig, seed: int = 0) -> dict:
    """float64 host parameters (the CPU network in tests reads the same dict)"""
    rng = np.random.default_rng(seed)
    W, H = cfg.width, cfg.hidden

    def lin(o, i):
        return rng.uniform(-1, 1, (o, i)) / math.sqrt(i)

    p = {"P": lin(W, cfg.cin), "bP": np.zeros(W)}
    nm = cfg.mx * cfg.my * cfg.mt
    for l in range(cfg.layers):
        s = 1.0 / (W * W)
        p[f"Rr{l}"] = s * rng.uniform(-1, 1, (W, W, nm))
        p[f"Ri{l}"] = s * rng.uniform(-1, 1, (W, W, nm))
        p[f"W{l}"] = lin(W, W)
        p[f"b{l}"] = np.zeros(W)
    p["Q1"] = lin(H, W)
    p["q1"] = np.zeros(H)
    p["Q2"] = lin(cfg.cout, H) * 0.1
    p["q2"] = np.zeros(cfg.cout)
    return p


class FnoGpu:
    """The network on the GPU: parameters, forward, hand-written backward, Adam."""

    def __init__(self, cfg: FnoConfig, params: dict):
        self.cfg = cfg
        self.ops = Ops()
        self.p = {k: DArr.of(v) for k, v in params.items()}
        self.g = {k: zeros(v.shape) for k, v in self.p.items()}
        self.m = {k: zeros(v.shape) for k, v in self.p.items()}
        self.v = {k: zeros(v.shape) for k, v in self.p.items()}
        self.mats = {k: DArr.of(v) for k, v in dft_mats(cfg).items()}
        self.step = 0
        self._buf = {}

    def host_params(self) -> dict:
        return {k: v.numpy().astype(np.float64) for k, v in self.p.items()}

    def grads(self) -> dict:
        return {k: v.numpy().astype(np.float64) for k, v in self.g.items()}

    def _b(self, name, shape):
        t = self._buf.get(name)
        if t is None or t.shape != tuple(shape):
            t = DArr(shape)
            self._buf[name] = t
        return t

    # ---- spectral convolution ------------------------------------------------
    def _spectral_fwd(self, v, l, out, save):
        """out = S_l(v); v, out: (B, W, X, Y, T).  Saves the mode coefficients for the backward."""
        c, o = self.cfg, self.ops
        B, W, X, Y, T, mx, my, mt = v.shape[0], c.width, c.X, c.Y, c.T, c.mx, c.my, c.mt
        M = self.mats
        n2 = 2 * mt
        # t: rows (b,w,x,y) x T -> 2mt
        C1 = self._b(f"C1", (B * W * X * Y, n2))
        o.gemm(B * W * X * Y, n2, T, v, T, 1, M["Ft"], n2, 1, C1, n2, 1)
        # y: per (b,w,x): [C;S] (2my x Y) . C1 slice (Y x 2mt)
        P2 = self._b("P2", (B * W * X, 2 * my, n2))
        o.gemm(2 * my, n2, Y, M["Gyf"], Y, 1, C1, n2, 1, P2, n2, 1, b1=(B * W * X, 0, Y * n2, 2 * my * n2))
        V2 = self._b("V2", (B * W * X, my, n2))
        o.combine(P2, V2, B * W * X, my, 1, mt, -1)
        # x: per (b,w): [C;S] (2mx x X) . V2 slice (X x my*2mt)
        row = my * n2
        P3 = self._b("P3", (B * W, 2 * mx, row))
        o.gemm(2 * mx, row, X, M["Gxf"], X, 1, V2, row, 1, P3, row, 1, b1=(B * W, 0, X * row, 2 * mx * row))
        V3 = save if save is not None else self._b("V3", (B, W, mx, my, n2))
        o.combine(P3, V3, B * W, mx, my, mt, -1)
        # mode mixing, per mode (kx,ky | kt): (B x W) . (W x W), complex via four real products
        U = self._b("U", (B, W, mx, my, n2))
        self._mix(V3, self.p[f"Rr{l}"], self.p[f"Ri{l}"], U, B)
        # inverse x: per (b,o): [C';S'] (2X x mx) . U slice (mx x my*2mt)
        P5 = self._b("P5", (B * W, 2 * X, row))
        o.gemm(2 * X, row, mx, M["Gxi"], mx, 1, U, row, 1, P5, row, 1, b1=(B * W, 0, mx * row, 2 * X * row))
        V5 = self._b("V5", (B * W, X, my, n2))
        o.combine(P5, V5, B * W, X, my, mt, +1)
        # inverse y: per (b,o,x): [C';S'] (2Y x my) . V5 slice (my x 2mt)
        P6 = self._b("P6", (B * W * X, 2 * Y, n2))
        o.gemm(2 * Y, n2, my, M["Gyi"], my, 1, V5, n2, 1, P6, n2, 1, b1=(B * W * X, 0, my * n2, 2 * Y * n2))
        V6 = self._b("V6", (B * W * X * Y, n2))
        o.combine(P6, V6, B * W * X, Y, 1, mt, +1)
        # inverse t
        o.gemm(B * W * X * Y, T, n2, V6, n2, 1, M["Gt"], T, 1, out, T, 1)

    def _mix(self, V, Rr, Ri, U, B):
        """U = V . R per mode (complex), V/U: (B, W, mx, my, 2mt), R: (W, W, mx*my, mt)"""
        c = self.cfg
        self.ops.mix(V, Rr, Ri, U, B, c.width, c.mx * c.my, c.mt)

    def _spectral_bwd(self, dz, l, V3, dv):
        """dv += S_l^T(dz); accumulates dRr_l, dRi_l from the saved mode coefficients V3"""
        c, o = self.cfg, self.ops
        B, W, X, Y, T, mx, my, mt = dz.shape[0], c.width, c.X, c.Y, c.T, c.mx, c.my, c.mt
        M = self.mats
        n2 = 2 * mt
        row = my * n2
        # inverse t^T
        dV6 = self._b("V6", (B * W * X * Y, n2))
        o.gemm(B * W * X * Y, n2, T, dz, T, 1, M["Gt"], 1, T, dV6, n2, 1)
        # inverse y^T
        dP6 = self._b("P6", (B * W * X, 2 * Y, n2))
        o.combine(dV6, dP6, B * W * X, Y, 1, mt, +1, bwd=True)
        dV5 = self._b("V5", (B * W, X, my, n2))
        o.gemm(my, n2, 2 * Y, M["Gyi"], 1, my, dP6, n2, 1, dV5, n2, 1, b1=(B * W * X, 0, 2 * Y * n2, my * n2))
        # inverse x^T
        dP5 = self._b("P5", (B * W, 2 * X, row))
        o.combine(dV5, dP5, B * W, X, my, mt, +1, bwd=True)
        dU = self._b("U", (B, W, mx, my, n2))
        o.gemm(mx, row, 2 * X, M["Gxi"], 1, mx, dP5, row, 1, dU, row, 1, b1=(B * W, 0, 2 * X * row, mx * row))
        # mixing^T and the weight gradients
        dV3 = self._b("dV3", (B, W, mx, my, n2))
        self._mix_bwd(V3, dU, self.p[f"Rr{l}"], self.p[f"Ri{l}"], dV3, self.g[f"Rr{l}"], self.g[f"Ri{l}"], B)
        # x^T
        dP3 = self._b("P3", (B * W, 2 * mx, row))
        o.combine(dV3, dP3, B * W, mx, my, mt, -1, bwd=True)
        dV2 = self._b("V2", (B * W * X, my, n2))
        o.gemm(X, row, 2 * mx, M["Gxf"], 1, X, dP3, row, 1, dV2, row, 1, b1=(B * W, 0, 2 * mx * row, X * row))
        # y^T
        dP2 = self._b("P2", (B * W * X, 2 * my, n2))
        o.combine(dV2, dP2, B * W * X, my, 1, mt, -1, bwd=True)
        dC1 = self._b("C1", (B * W * X * Y, n2))
        o.gemm(Y, n2, 2 * my, M["Gyf"], 1, Y, dP2, n2, 1, dC1, n2, 1, b1=(B * W * X, 0, 2 * my * n2, Y * n2))
        # t^T
        o.gemm(B * W * X * Y, T, n2, dC1, n2, 1, M["Ft"], 1, n2, dv, T, 1, beta=1.0)

    def _mix_bwd(self, V, dU, Rr, Ri, dV, dRr, dRi, B):
        """dV = dU . R^H per mode; dR += V^H . dU"""
        c = self.cfg
        self.ops.mix_bwd(V, dU, Rr, Ri, dV, dRr, dRi, B, c.width, c.mx * c.my, c.mt)

    # ---- pointwise maps --------------------------------------------------------
    def _pw_bwd_input(self, Wm, dout, dx, nout, nin, npts, B):
        """dx[b] = Wm^T . dout[b]  (Wm: nout x nin, row-major; its transpose is
        formed first so the map reads unit-stride weight rows)"""
        WT = self._b(f"WT{nout}x{nin}", (nin, nout))
        self.ops.transpose(Wm, WT, nout, nin)
        self.ops.pointwise(dout, nout, WT, nout, 1, dx, nin, B, npts)

    def _pw_bwd_weight(self, dout, x, dW, nout, nin, npts, B):
        """dW += sum_b dout[b] . x[b]^T"""
        for b in range(B):
            self.ops.gemm(nout, nin, npts, dout, npts, 1, x, 1, npts, dW, nin, 1, beta=1.0,
                          offA=b * nout * npts, offB=b * nin * npts)

    # ---- network ---------------------------------------------------------------
    def forward(self, x, keep=False):
        """x: (B, cin, X, Y, T) float32 cuda -> pred (B, cout, X, Y, T).  keep: save activations for backward"""
        c = self.cfg
        B = x.shape[0]
        npts = c.X * c.Y * c.T
        W, H = c.width, c.hidden
        # activations live in named buffers that persist between calls (no
        # allocation per call); with keep they are per layer, for the backward
        st = {"x": x}
        shp = (B, W, c.X, c.Y, c.T)
        v = self._b("v0", shp)
        o = self.ops
        o.pointwise(x, c.cin, self.p["P"], c.cin, 1, v, W, B, npts, bias=self.p["bP"])
        for l in range(c.layers):
            z = self._b(f"z{l}" if keep else "z", shp)
            V3 = self._b(f"V3{l}", (B, W, c.mx, c.my, 2 * c.mt)) if keep else None
            self._spectral_fwd(v, l, z, V3)
            vn = self._b(f"v{l + 1}" if keep else f"v{(l + 1) % 2}", shp)
            act = l < c.layers - 1
            o.pointwise(v, W, self.p[f"W{l}"], W, 1, vn, W, B, npts, bias=self.p[f"b{l}"], zin=z, beta=1.0,
                        zs=z if keep else None, act=act)
            if keep:
                st[f"v{l}"], st[f"z{l}"], st[f"V3{l}"] = v, z, V3
            v = vn
        hz = self._b("hz", (B, H, c.X, c.Y, c.T)) if keep else None
        h = self._b("h", (B, H, c.X, c.Y, c.T))
        o.pointwise(v, W, self.p["Q1"], W, 1, h, H, B, npts, bias=self.p["q1"], zs=hz if keep else None, act=True)
        y = self._b("y", (B, c.cout, c.X, c.Y, c.T))
        # pred = x[:, :cout] + Q2 h + q2
        o.pointwise(h, H, self.p["Q2"], H, 1, y, c.cout, B, npts, bias=self.p["q2"], zin=x, beta=1.0,
                    zin_bs=c.cin * npts)
        if keep:
            st["vL"], st["hz"], st["h"] = v, hz, h
            self._st = st
        return y

    def backward(self, dpred):
        """gradients of the loss whose d/dpred is dpred (after forward(keep=True)) into self.g"""
        c, o, st = self.cfg, self.ops, self._st
        B = dpred.shape[0]
        npts = c.X * c.Y * c.T
        W, H = c.width, c.hidden
        for g in self.g.values():
            g.fill(0.0)
        # head
        self._pw_bwd_weight(dpred, st["h"], self.g["Q2"], c.cout, H, npts, B)
        o.bias_grad(dpred, self.g["q2"], B, c.cout, npts)
        dh = self._b("dh", st["h"].shape)
        self._pw_bwd_input(self.p["Q2"], dpred, dh, c.cout, H, npts, B)
        o.act_bwd(st["hz"], dh, dh, True)
        self._pw_bwd_weight(dh, st["vL"], self.g["Q1"], H, W, npts, B)
        o.bias_grad(dh, self.g["q1"], B, H, npts)
        dv = self._b("dv", st["vL"].shape)
        self._pw_bwd_input(self.p["Q1"], dh, dv, H, W, npts, B)
        # Fourier layers, last to first
        for l in reversed(range(c.layers)):
            act = l < c.layers - 1
            dz = self._b("dz", dv.shape)
            o.act_bwd(st[f"z{l}"], dv, dz, act)
            o.bias_grad(dz, self.g[f"b{l}"], B, W, npts)
            self._pw_bwd_weight(dz, st[f"v{l}"], self.g[f"W{l}"], W, W, npts, B)
            self._pw_bwd_input(self.p[f"W{l}"], dz, dv, W, W, npts, B)
            self._spectral_bwd(dz, l, st[f"V3{l}"], dv)
        # lift (dz0 = dv, no activation)
        self._pw_bwd_weight(dv, st["x"], self.g["P"], W, c.cin, npts, B)
        o.bias_grad(dv, self.g["bP"], B, W, npts)
        self._st = None

    def loss_and_grad(self, x, tgt):
        """relative L2 loss mean_b |pred-tgt|^2/|tgt|^2; gradients into self.g"""
        pred = self.forward(x, keep=True)
        dpred = self._b("dpred", pred.shape)
        loss = self._b("loss", (1 + 2 * x.shape[0],))
        self.ops.rel_l2(pred, tgt, dpred, loss)
        self.backward(dpred)
        return float(loss.numpy()[0])

    def adam_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
        self.step += 1
        lr_t = lr * math.sqrt(1 - b2 ** self.step) / (1 - b1 ** self.step)
        for k in self.p:
            self.ops.adam(self.p[k], self.g[k], self.m[k], self.v[k], lr_t, b1, b2, eps, wd)


def time_network(batch: int = 1, reps: int = 20, train_steps: int = 0, seed: int = 1, width: int = 32) -> dict:
    """host-clock timing of the network at its full size (synthetic input):
    forward passes at `batch`, then `train_steps` training steps (forward,
    backward, Adam) at the same batch"""
    import time
    cfg = FnoConfig(width=width, hidden=2 * width)
    net = FnoGpu(cfg, init_params(cfg, seed))
    n = cfg.X * cfg.Y * cfg.T
    rng = np.random.default_rng(seed)
    x = DArr.of(rng.standard_normal((batch, cfg.cin, n)).astype(np.float32) * 0.1)
    net.forward(x)
    sync()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        net.forward(x)
        sync()
        ts.append(time.perf_counter() - t0)
    f0 = net.ops.gemm_flops
    net.forward(x)
    sync()
    fwd_flops = net.ops.gemm_flops - f0
    fwd_s = float(np.median(ts))
    # the BF16 matrix-core share of the forward: its GEMM flops over the whole
    # forward time, against the dense BF16 peak (2.5 PFLOP/s)
    out = {"width": width, "batch": batch, "forward_ms": 1e3 * fwd_s, "forward_gemm_gflop": fwd_flops / 1e9,
           "forward_bf16_tflops": fwd_flops / fwd_s / 1e12, "bf16_peak_frac": fwd_flops / fwd_s / 2.5e15}
    if train_steps:
        tgt = DArr.of(rng.standard_normal((batch, cfg.cout, n)).astype(np.float32) * 0.1)
        net.loss_and_grad(x, tgt)
        net.adam_step()
        sync()
        ts = []
        for _ in range(train_steps):
            t0 = time.perf_counter()
            net.loss_and_grad(x, tgt)
            net.adam_step()
            sync()
            ts.append(time.perf_counter() - t0)
        out["train_step_ms"] = 1e3 * float(np.median(ts))
    return out


if __name__ == "__main__":
    import argparse
    import json
    ap = argparse.ArgumentParser(description="time the surrogate network at its full size")
    ap.add_argument("--batch", type=int, default=1)
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--train-steps", type=int, default=0)
    ap.add_argument("--width", type=int, default=32, help="channels (the head is twice as wide)")
    a = ap.parse_args()
    print(json.dumps(time_network(a.batch, a.reps, a.train_steps, width=a.width)), flush=True)
