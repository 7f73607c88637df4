"""Pass surrogate: training data from the solver, training, and surrogate-
assisted rollouts with the plain split step as fallback and as the measure
of their error.

Workload (the precision-form ranges around the 64x64 workload): pulses on a
64 x 64 x 256 grid over 4 mm x 2 ps, 0.5 m cell, 200 z-steps per pass,
energies 5-20 uJ, argon 0.5-2 bar, waists 0.4-0.6 mm, durations 150-250 fs.
A training pair is (field at a mirror plane, field one pass later), both
taken on the device from the solver's resident field (mpc_export_planar) and
scaled by the same per-pulse factor (1/16 of the field's rms), so the network
works in units where the pulse peak is O(1).

Physics prior: the network does not map a field to the next pass's field
directly.  Its input is the field after the pass's linear propagation
(mpc_linear: dispersion and diffraction over the cell length, then the
mirror; a few sweeps instead of the pass's hundreds), and it predicts the
correction that the nonlinear steps add to it.  The method without the
network is then that linear map, and the held-out error of each is reported.

Rollout: every plain_every-th pass, and every pass whose prediction fails
the energy check (|E_pred / E_in - 1| > energy_tol) or the spectral-width
check (|w_pred / w_in - 1| > width_tol), or is not finite, is run by the
plain split step; the other passes are the network's.  The checks are the
solver guard's verdict (mpc_guard_verdict; a tolerance < 0 skips its check).
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass

import numpy as np

from . import Batch, Propagator, configs, make_inputs
from ._lib import check, lib
from .fno import DArr, FnoConfig, FnoGpu, init_params, sync

FS = 1e-15


def pulse_workloads(n: int, seed: int, passes: int, nt: int = 256, nx: int = 64) -> list:
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        e = float(10 ** rng.uniform(np.log10(5e-6), np.log10(20e-6)))
        p = float(rng.uniform(0.5, 2.0))
        w0 = float(rng.uniform(0.4e-3, 0.6e-3))
        tau = float(rng.uniform(150, 250)) * FS
        out.append(configs.Workload(f"surrogate pulse {i}", nx, nt, 4e-3, 2e-12, w0, tau, e, p, 20.0, passes, 200,
                                    seed=seed * 1000 + i))
    return out


def scalars_of(wl) -> tuple:
    """the two constant input channels: log energy and pressure, O(1)"""
    return (math.log(wl.energy / 10e-6), wl.pressure - 1.0)


def field_scale(prop) -> float:
    d = prop.diagnostics()
    n = prop.nx * prop.ny * prop.nt
    return 1.0 / (16.0 * math.sqrt(d["energy_sum"] / n))


def linear_pass(prop, wl):
    """the linear part of one pass (curved mirror: diffraction sign +1) and the
    mirror after it"""
    prop.linear(wl.cell_L, 1, mirror=True)


def export(prop, dst: DArr, off: int, scale: float):
    check(lib().mpc_export_planar(prop.ctx, C.c_void_p(dst.ptr + 4 * off), scale), "mpc_export_planar")


def import_(prop, src: DArr, off: int, scale: float):
    sync()
    check(lib().mpc_import_planar(prop.ctx, C.c_void_p(src.ptr + 4 * off), scale), "mpc_import_planar")


@dataclass
class PairSet:
    states: DArr          # (pulses, passes + 1, 2, N) on the device
    lin: DArr             # (pulses, passes, 2, N): the linear pass map of states[:, k]
    scalars: np.ndarray   # (pulses, 2)
    npulses: int
    passes: int
    npts: int
    gen_seconds: float


def generate(n_pulses: int, passes: int, seed: int, chunk: int = 64, prec: str = "f32",
             log=print) -> PairSet:
    """run the batched solver pass by pass and keep every mirror-plane field"""
    wls = pulse_workloads(n_pulses, seed, passes + 1)
    npts = wls[0].nx * wls[0].nx * wls[0].nt
    states = DArr((n_pulses, passes + 1, 2, npts))
    lin = DArr((n_pulses, passes, 2, npts))
    scal = np.array([scalars_of(w) for w in wls])
    t0 = time.perf_counter()
    S = wls[0].steps_per_pass
    for c0 in range(0, n_pulses, chunk):
        props, scales = [], []
        for i in range(c0, min(n_pulses, c0 + chunk)):
            inp = make_inputs(wls[i])
            p = Propagator(inp, prec)
            p.upload(inp.A)
            p.calibrate()
            inp.A = None
            props.append(p)
            scales.append(field_scale(p))
        for j, p in enumerate(props):
            export(p, states, ((c0 + j) * (passes + 1) + 0) * 2 * npts, scales[j])
        with Batch(props) as b:
            for k in range(passes):
                b.run(k * S, S, graph=True)
                for j, p in enumerate(props):
                    export(p, states, ((c0 + j) * (passes + 1) + k + 1) * 2 * npts, scales[j])
        for j, p in enumerate(props):               # the linear map of every pass's input
            for k in range(passes):
                i = c0 + j
                import_(p, states, (i * (passes + 1) + k) * 2 * npts, scales[j])
                linear_pass(p, wls[i])
                export(p, lin, (i * passes + k) * 2 * npts, scales[j])
        for p in props:
            p.close()
        log(f"[surrogate] pulses {c0}..{c0 + len(props) - 1}: {passes} passes")
    return PairSet(states, lin, scal, n_pulses, passes, npts, time.perf_counter() - t0)


class Trainer:
    """batches of (state k -> state k+1) pairs assembled on the device"""

    def __init__(self, cfg: FnoConfig, data: PairSet, batch: int = 4, seed: int = 0, train_pulses=None):
        self.cfg, self.data, self.B = cfg, data, batch
        self.net = FnoGpu(cfg, init_params(cfg, seed))
        self.rng = np.random.default_rng(seed + 1)
        n = data.npts
        self.x = DArr((batch, cfg.cin, n))
        self.y = DArr((batch, cfg.cout, n))
        self.pulses = np.arange(data.npulses) if train_pulses is None else np.asarray(train_pulses)

    def _fill(self, picks):
        d, n = self.data, self.data.npts
        for b, (p, k) in enumerate(picks):
            src = (p * (d.passes + 1) + k) * 2 * n
            self.x.copy_from(d.lin, 2 * n, dst_off=b * self.cfg.cin * n, src_off=(p * d.passes + k) * 2 * n)
            for c in range(2):
                check(lib().mpc_dev_fill(C.c_void_p(self.x.ptr + 4 * (b * self.cfg.cin * n + (2 + c) * n)), n,
                                         float(d.scalars[p, c]), None), "mpc_dev_fill")
            self.y.copy_from(d.states, 2 * n, dst_off=b * 2 * n, src_off=src + 2 * n)

    def step(self, lr: float) -> float:
        picks = [(int(self.rng.choice(self.pulses)), int(self.rng.integers(self.data.passes)))
                 for _ in range(self.B)]
        self._fill(picks)
        loss = self.net.loss_and_grad(self.x, self.y)
        self.net.adam_step(lr=lr)
        return loss

    def eval_loss(self, pulses) -> float:
        """mean relative L2 (not squared) over every pair of the given pulses"""
        errs = []
        for p in pulses:
            for k in range(self.data.passes):
                self._fill([(p, k)] * self.B)
                pred = self.net.forward(self.x).numpy()[0].ravel()
                tgt = self.y.numpy()[0].ravel()
                errs.append(float(np.linalg.norm(pred - tgt) / np.linalg.norm(tgt)))
        return float(np.mean(errs))


def rollout(net: FnoGpu, wl, passes: int, plain_every: int = 5, energy_tol: float = 0.02,
            width_tol: float = 0.05, prec: str = "f32") -> dict:
    """surrogate-assisted rollout of one pulse against the plain trajectory;
    times cover the pass work only (solver passes, or network + checks).
    The plain pass is timed both ways the library offers: host-driven (one
    launch call per sweep) and as a HIP graph of the pass's launches
    captured before the timed region.  The faster of the two is the plain
    baseline, and the rollout's own plain passes run that way."""
    cfg = net.cfg
    S = wl.steps_per_pass
    inp = make_inputs(wl)
    n = wl.nx * wl.nx * wl.nt
    # plain trajectory
    ref = Propagator(inp, prec)
    ref_b = Batch([ref])
    ref_b.run(0, S, graph=True)                  # captures the pass graph (the state is reset below)
    ref.upload(inp.A)
    ref.calibrate()
    scale = field_scale(ref)
    ref_states = DArr((passes + 1, 2, n))
    export(ref, ref_states, 0, scale)
    t_plain = 0.0
    for k in range(passes):
        t0 = time.perf_counter()
        ref_b.run(k * S, S, graph=True)
        t_plain += time.perf_counter() - t0
        export(ref, ref_states, (k + 1) * 2 * n, scale)
    # the same passes driven from the host without a graph (same results)
    ref.upload(inp.A)
    ref.calibrate()
    t_eager = 0.0
    for k in range(passes):
        t0 = time.perf_counter()
        ref.run(k * S, S)
        t_eager += time.perf_counter() - t0
    ref_b.close()
    ref.close()
    refs = ref_states.numpy()
    use_graph = t_plain < t_eager
    t_best = min(t_plain, t_eager)
    # assisted
    p = Propagator(inp, prec)
    pb = Batch([p])
    pb.run(0, S, graph=True)                     # capture, then back to the input
    p.upload(inp.A)
    p.calibrate()
    x = DArr((1, cfg.cin, n))
    cur = DArr((2, n))
    start = DArr((2, n))
    sc = scalars_of(wl)
    x.fill(0.0)
    net.forward(x)                               # allocates the network's buffers before any timing
    for c in range(2):                           # the constant channels; exports only write channels 0-1
        check(lib().mpc_dev_fill(C.c_void_p(x.ptr + 4 * (2 + c) * n), n, float(sc[c]), None), "mpc_dev_fill")
    sync()
    t_net = []
    errs, kinds = [], []
    t_assist = 0.0
    t_netpass = []                               # whole network passes: linear map, network, import, checks
    d_cur = None                                 # diagnostics of the current state, when known
    for k in range(passes):
        t0 = time.perf_counter()
        kind = "plain" if k % plain_every == plain_every - 1 else "net"
        if kind == "net":
            d_in = d_cur if d_cur is not None else p.diagnostics()
            export(p, start, 0, scale)
            linear_pass(p, wl)
            export(p, x, 0, scale)
            sync()
            t1 = time.perf_counter()
            pred = net.forward(x)
            sync()
            t_net.append(time.perf_counter() - t1)
            import_(p, pred, 0, scale)
            d_out = p.diagnostics()
            # the solver guard's verdict, two-sided on energy and width
            verdict = lib().mpc_guard_verdict(d_in["energy_sum"], d_out["energy_sum"], d_in["spec_rms_width"],
                                              d_out["spec_rms_width"], -1.0, energy_tol, width_tol)
            d_cur = d_out
            if verdict != 0:
                import_(p, start, 0, scale)      # back to the pass's input, then the plain pass
                kind = "fallback"
        if kind == "net":
            t_netpass.append(time.perf_counter() - t0)
        if kind != "net":
            if use_graph:
                pb.run(k * S, S, graph=True)
            else:
                p.run(k * S, S)
            d_cur = None
        t_assist += time.perf_counter() - t0
        kinds.append(kind)
        export(p, cur, 0, scale)
        a = cur.numpy()
        errs.append(float(np.linalg.norm(a - refs[k + 1]) / np.linalg.norm(refs[k + 1])))
    pb.close()
    p.close()
    return {"errs": errs, "kinds": kinds, "t_plain": t_best, "t_assist": t_assist,
            "passes_per_s_plain": passes / t_best, "passes_per_s_assist": passes / t_assist,
            "passes_per_s_plain_graph": passes / t_plain, "passes_per_s_plain_host_driven": passes / t_eager,
            "plain_uses_graph": use_graph,
            "net_forward_ms": 1e3 * float(np.median(t_net)) if t_net else None,
            "net_pass_ms": 1e3 * float(np.median(t_netpass)) if t_netpass else None,
            "plain_pass_ms": 1e3 * t_best / passes}


def baseline_errors(data: PairSet, pulses) -> tuple:
    """mean rel L2 over the given pulses' pairs of the two methods without a
    network: 'no change over a pass' and the linear pass map alone (what the
    network's correction has to improve on)"""
    ident, linear = [], []
    n = data.npts
    buf = DArr((data.passes + 1, 2, n))
    lb = DArr((data.passes, 2, n))
    for p in pulses:
        buf.copy_from(data.states, (data.passes + 1) * 2 * n, src_off=p * (data.passes + 1) * 2 * n)
        lb.copy_from(data.lin, data.passes * 2 * n, src_off=p * data.passes * 2 * n)
        a, l = buf.numpy(), lb.numpy()
        for k in range(data.passes):
            nrm = np.linalg.norm(a[k + 1])
            ident.append(float(np.linalg.norm(a[k + 1] - a[k]) / nrm))
            linear.append(float(np.linalg.norm(a[k + 1] - l[k]) / nrm))
    return float(np.mean(ident)), float(np.mean(linear))


def run_experiment(n_pulses: int = 512, passes: int = 8, train_seconds: float = 240.0, batch: int = 4,
                   lr: float = 2e-3, test_pulses: int = 4, rollout_passes: int = 30, plain_every: int = 5,
                   cfg: FnoConfig | None = None, log=print) -> dict:
    """generate pairs, train for a wall-clock budget, evaluate held-out pairs
    and surrogate-assisted rollouts of unseen pulses"""
    cfg = cfg or FnoConfig()
    data = generate(n_pulses, passes, seed=20261202 + 4, log=log)
    gen_s = data.gen_seconds
    log(f"[surrogate] {n_pulses * passes} pairs generated in {gen_s:.1f}s")
    held = list(range(n_pulses - max(1, n_pulses // 16), n_pulses))
    tr = Trainer(cfg, data, batch=batch, train_pulses=[p for p in range(n_pulses) if p not in held])
    t0 = time.perf_counter()
    steps, losses = 0, []
    while time.perf_counter() - t0 < train_seconds:
        frac = (time.perf_counter() - t0) / train_seconds
        losses.append(tr.step(lr * 0.5 * (1 + math.cos(math.pi * min(1.0, frac)))))
        steps += 1
        if steps % 200 == 0:
            log(f"[surrogate] step {steps}: loss {np.mean(losses[-200:]):.4e}")
    t_train = time.perf_counter() - t0
    held_err = tr.eval_loss(held[:4])
    identity_err, linear_err = baseline_errors(data, held[:4])
    log(f"[surrogate] trained {steps} steps in {t_train:.0f}s; held-out pair rel L2 {held_err:.3e} "
        f"(linear map alone {linear_err:.3e}, no change {identity_err:.3e})")
    tr.data = data = None                      # free the pair set before the rollouts
    wls = pulse_workloads(test_pulses, seed=777, passes=rollout_passes + 1)
    rolls = [rollout(tr.net, w, rollout_passes, plain_every=plain_every) for w in wls]
    for i, r in enumerate(rolls):
        log(f"[surrogate] rollout {i}: max rel L2 {max(r['errs']):.3e}, final {r['errs'][-1]:.3e}, "
            f"{r['kinds'].count('net')} net / {r['kinds'].count('fallback')} fallback / "
            f"{r['kinds'].count('plain')} plain passes; {r['passes_per_s_assist']:.1f} vs "
            f"{r['passes_per_s_plain']:.1f} passes/s plain (the faster of a HIP graph per pass, "
            f"{r['passes_per_s_plain_graph']:.1f}, and host-driven launches, "
            f"{r['passes_per_s_plain_host_driven']:.1f})")
    per_pass = [e for r in rolls for e in r["errs"]]
    return {"network": {"width": cfg.width, "layers": cfg.layers, "modes": [cfg.mx, cfg.my, cfg.mt],
                        "hidden": cfg.hidden},
            "pairs": n_pulses * passes, "gen_seconds": round(gen_s, 1),
            "train_steps": steps, "train_seconds": round(t_train, 1), "batch": batch,
            "final_train_loss": float(np.mean(losses[-100:])) if losses else None,
            "held_out_pair_rel_l2": held_err, "identity_pair_rel_l2": identity_err,
            "linear_pair_rel_l2": linear_err,
            "rollout_max_rel_l2": max(per_pass), "rollout_mean_rel_l2": float(np.mean(per_pass)),
            "net_passes": sum(r["kinds"].count("net") for r in rolls),
            "fallback_passes": sum(r["kinds"].count("fallback") for r in rolls),
            "plain_passes": sum(r["kinds"].count("plain") for r in rolls),
            "passes_per_s_assisted": float(np.mean([r["passes_per_s_assist"] for r in rolls])),
            "passes_per_s_plain": float(np.mean([r["passes_per_s_plain"] for r in rolls])),
            "passes_per_s_plain_graph": float(np.mean([r["passes_per_s_plain_graph"] for r in rolls])),
            "passes_per_s_plain_host_driven": float(np.mean([r["passes_per_s_plain_host_driven"] for r in rolls])),
            "net_forward_ms": float(np.median([r["net_forward_ms"] for r in rolls if r["net_forward_ms"]]))
            if any(r["net_forward_ms"] for r in rolls) else None,
            "net_pass_ms": float(np.median([r["net_pass_ms"] for r in rolls if r["net_pass_ms"]]))
            if any(r["net_pass_ms"] for r in rolls) else None,
            "plain_pass_ms": float(np.mean([r["plain_pass_ms"] for r in rolls])),
            # with every plain_every-th pass plain, the assisted rate is below
            # plain_every x the plain rate however fast the network passes are
            "assisted_speedup_bound": float(plain_every)}
