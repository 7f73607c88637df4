#!/usr/bin/env python3
"""Benchmark: split-step steps/s of the 256x256x2048 argon workload on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--no-batch]
    python bench.py --batch        (the batched workloads alone)
    python bench.py --surrogate    (the pass surrogate experiment instead)
    python bench.py --adaptive     (adaptive step doubling + Richardson instead)

One "step" is one z-step L(dz/2) N(dz) L(dz/2) of the whole 256x256x2048 field
(Kerr + self-steepening + ADK/avalanche plasma + diffraction + dispersion),
run through libmpc_hip.so with the field resident in HBM.  `value` is the FP32
path (FP64 kept for the propagator phase, the ADK exponent and the density
scan); the plain FP64 path is timed on the same inputs, and the FP32 field is
compared with the FP64 one after the same steps (rel L2).

The default run also
* compares the GPU's FP64 and FP32 fields, energies, rms spectral widths and
  peak densities with the FP64 CPU oracle's after the steps its cpu_baseline
  leg ran (config.vs_oracle), and
* times the 48-pulse batched workload and 48 small (64x64x256) pulses, batched
  against one by one (config.batched).

Prints ONE JSON line.  The roofline object is for the dominant sweep kernel,
timed with HIP events on the library's stream during a separate pass over the
same steps; cpu_baseline is the FP64 CPU oracle on a bounded sample of the
same workload (with the host's core count and CPU model).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "split-step steps/s at 256²×2048 (FP32/batched vs plain FP64, ≤1e-4 rel L2)"
HBM_PEAK_GBS = 8000.0
KERNELS = ["t-line (nonlinear + t-FFTs)", "x-line (FFT x D x IFFT)", "y-line (FFT x D x IFFT)"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def variant_sweep(inp, prec, steps=3):
    """per-variant kernel times (ms): the numbers the autotuner chooses from"""
    from mpc_amd import Propagator, variant_counts
    out = {}
    p = Propagator(inp, prec)
    ntv, nav = variant_counts()
    for tv in range(ntv):
        for av in range(nav):
            p.upload(inp.A)
            p.calibrate()
            p.set_variants(tv, av)
            p.set_timing(True)
            p.run(0, steps)
            out[f"t{tv}a{av}"] = [round(p.kernel_ms(i), 4) for i in range(3)]
            p.set_timing(False)
    p.close()
    log(f"[bench] {prec} variant sweep (t-line, x-line, y-line ms): {out}")
    return out


def timed_run(inp, prec, warmup, steps, with_kernel_timing=True):
    from mpc_amd import Propagator
    p = Propagator(inp, prec)
    p.upload(inp.A)
    p.calibrate()
    variants = p.autotune()
    log(f"[bench] {prec}: autotuned variants (t-line, x-line, y-line) = {variants}")
    t0 = time.perf_counter()
    p.run(0, warmup)
    log(f"[bench] {prec}: warmup {warmup} steps {time.perf_counter() - t0:.2f}s")
    t0 = time.perf_counter()
    p.run(warmup, steps)                     # returns after the library stream drained
    dt = time.perf_counter() - t0
    log(f"[bench] {prec}: timed {steps} steps {dt:.3f}s ({steps / dt:.1f} steps/s)")
    A = p.download()
    kms = None
    if with_kernel_timing:
        p.upload(inp.A)
        p.calibrate()
        p.set_timing(True)
        p.run(0, steps)
        kms = [p.kernel_ms(i) for i in range(3)]
        p.set_timing(False)
        log(f"[bench] {prec}: kernel ms t/x/y = {kms}")
    bytes_field = p.inp.A.size * (8 if prec == "f32" else 16)
    p.close()
    return steps / dt, dt, A, kms, bytes_field, variants


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(inp, max_seconds=30.0):
    """FP64 CPU oracle (test infrastructure) on the same workload: 1 warm-up
    step, then steps until ~10 s have elapsed (at most 3).  Returns the
    baseline record, the oracle's field after those steps and their count."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    O.oracle()
    A = inp.calibrated()
    threads = O.oracle().oracle_get_threads()
    t0 = time.perf_counter()
    A, _ = O.run_fixed(inp, A, 0, 1)
    log(f"[bench] cpu oracle warm-up step {time.perf_counter() - t0:.1f}s on {threads} threads")
    n, t0 = 0, time.perf_counter()
    rho = 0.0
    while n < 3:
        A, rho = O.run_fixed(inp, A, 1 + n, 1)
        n += 1
        if time.perf_counter() - t0 > min(10.0, max_seconds):
            break
    dt = time.perf_counter() - t0
    rec = {"value": n / dt, "unit": "steps/s", "cores": threads, "kind": "port",
           "nproc": os.cpu_count(), "cpu_model": _cpu_model(),
           "sample": f"{n} z-step(s) of the full 256x256x2048 workload after 1 warm-up step, "
                     f"FP64 oracle (oracle/, OpenMP {threads} threads), {dt:.1f} s"}
    return rec, A, 1 + n, rho


def against_oracle(inp, A_ref, nsteps, rho_ref):
    """the GPU's FP64 and FP32 fields after the oracle's steps, from the same
    calibrated input: field rel L2 and the observables' relative errors"""
    import oracle_lib as O
    from mpc_amd import Propagator
    ob = O.observables(inp, A_ref)
    out = {"steps": nsteps}
    for prec in ("f64", "f32"):
        with Propagator(inp, prec) as p:
            p.upload(inp.calibrated())
            p.run(0, nsteps)
            A = p.download()
            d = p.diagnostics()
        out[prec] = {"field_rel_l2": O.rel_l2(A, A_ref),
                     "energy_rel_err": abs(d["energy_sum"] - ob["energy"]) / ob["energy"],
                     "rms_width_rel_err": abs(d["spec_rms_width"] - ob["width"]) / ob["width"],
                     "rho_peak_rel_err": abs(d["rho_peak"] - rho_ref) / rho_ref if rho_ref > 0 else None}
        log(f"[bench] {prec} vs oracle after {nsteps} steps: {out[prec]}")
        del A
    return out


def _batch_vs_one_by_one(wls, steps, label):
    """pulses advancing `steps` z-steps together (mpc_batch, one launch per
    sweep and launch group, a HIP graph per pass, variants tuned on the
    batch's own launches) against the same pulses run one after another with
    their own tuned variants; the four calls of `steps` steps (two warm-up,
    the timed batched one, the one-by-one one) stay inside the first pass"""
    from mpc_amd import Batch, Propagator, make_inputs
    steps = max(1, min(steps, (min(w.steps_per_pass for w in wls) - 1) // 4))
    props, npts = [], []
    tuned = {}
    t0 = time.perf_counter()
    for wl in wls:
        inp = make_inputs(wl)
        p = Propagator(inp, "f32")
        p.upload(inp.A)
        p.calibrate()
        inp.A = None                             # host copy no longer needed
        key = (wl.nx, wl.nt)
        if key not in tuned:
            tuned[key] = p.autotune()
        else:
            p.set_variants(*tuned[key])
        props.append(p)
        npts.append(wl.nx * wl.nx * wl.nt)
    log(f"[bench] {label}: {len(props)} pulses set up in {time.perf_counter() - t0:.1f}s; "
        f"grids {sorted(set((w.nx, w.nt) for w in wls))}; variants {tuned}")
    b = Batch(props)
    b.autotune()
    forms = {key: next(p.variants() for p, w in zip(props, wls) if (w.nx, w.nt) == key) for key in tuned}
    log(f"[bench] {label}: variants tuned on the batch {forms}")
    # a run leaves the fields t-spectral, so the second warm-up call has the
    # timed call's launch list (t-spectral start, same length, no mirror
    # inside) and its pass graph is captured before the timed region
    b.run(0, steps, graph=True)
    b.run(steps, steps, graph=True)
    t0 = time.perf_counter()
    b.run(2 * steps, steps, graph=True)
    dt_b = time.perf_counter() - t0
    graphs = b.graphs()
    b.close()
    for p, w in zip(props, wls):                  # one by one: each pulse's own tuned variants
        p.set_variants(*tuned[(w.nx, w.nt)])
    t0 = time.perf_counter()
    for p in props:
        p.run(3 * steps, steps)
    dt_s = time.perf_counter() - t0
    for p in props:
        p.close()
    nref = 256 * 256 * 2048
    work = sum(npts) * steps / nref
    log(f"[bench] {label}: batched {dt_b:.3f}s ({work / dt_b:.1f} equivalent steps/s, graphs {graphs}); "
        f"one by one {dt_s:.3f}s ({work / dt_s:.1f}); ratio {dt_s / dt_b:.3f}")
    return {"batched_equivalent_steps_per_s": round(work / dt_b, 3),
            "one_by_one_equivalent_steps_per_s": round(work / dt_s, 3),
            "batched_over_one_by_one": round(dt_s / dt_b, 4),
            "pulse_steps_per_s_batched": round(len(wls) * steps / dt_b, 1),
            "pulse_steps_per_s_one_by_one": round(len(wls) * steps / dt_s, 1),
            "pulses": len(wls), "steps": steps, "points_total": int(sum(npts)),
            "variants_batch_tuned": {f"{k[0]}^2x{k[1]}": list(v) for k, v in forms.items()},
            "unit": "256x256x2048-equivalent steps/s", "dtype": "f32"}


def batch_bench(steps, warmup):
    """the 48-pulse batched workload in FP32 (rate in 256x256x2048-equivalent
    steps/s: sum_p N_p * steps / N_ref per s), plus 48 small pulses
    (64x64x256, the pass surrogate's solver workload) where launches
    dominate a step"""
    from mpc_amd import configs, surrogate
    big = _batch_vs_one_by_one(configs.cfg4_batch(), steps, "batch")
    small = _batch_vs_one_by_one(surrogate.pulse_workloads(48, seed=20261209, passes=10), min(3 * steps, 60),
                                 "small-pulse batch")
    return {"metric": "batched split-step, 256x256x2048-equivalent steps/s (48 pulses, FP32)",
            "value": big["batched_equivalent_steps_per_s"], "unit": "steps/s", "n_gpus": 1, "steps": steps,
            "warmup": steps, "higher_is_better": True, "dtype": "f32",
            "data": "synthetic (seeded pulses as drawn below)",
            "config": {"workload": "48 pulses drawn from {128^2x512, 192^2x1024, 256^2x2048}, 0.1-2 mJ, "
                                   "100-500 fs, Ar 0.5-3 bar, w0 0.5-1 mm; 10 passes x 300 steps",
                       "one_by_one_steps_per_s": big["one_by_one_equivalent_steps_per_s"],
                       "points_total": big["points_total"], "graphs": True, "cfg4": big,
                       "small_pulses_64x64x256": small}}


def load_traffic(prec):
    path = os.path.join(ROOT, "profiles", f"traffic_tline_256x256x2048_{prec}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def load_valu(prec, variant):
    """VALU wave-instructions per launch of the t-line variant in use (from the
    SQ-counter run) and the issue floor they set: one wave-instruction per
    SIMD every 4 cycles on 1024 SIMDs at 2.4 GHz"""
    path = os.path.join(ROOT, "profiles", f"valu_tline_256x256x2048_{prec}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    v = d["by_variant"].get(str(variant))
    if v is None:
        return None
    floor_ms = v["valu_instr"] * d["cycles_per_instr"] / (d["simds"] * d["clock_ghz"] * 1e9) * 1e3
    return {"kernel": v["kernel"], "valu_instr_per_launch": v["valu_instr"], "issue_floor_ms": round(floor_ms, 4),
            "source": d["source"]}


def adaptive_bench(steps, warmup):
    """adaptive step doubling + Richardson at 256x256x2048 (FP32 and FP64):
    each attempt is one 2dz step (C1) and two dz steps (C2), i.e. three
    isolated split steps, plus the error norm, the beam-width cap and the
    Richardson combination; an accepted step advances z by 2dz"""
    from mpc_amd import Propagator, configs, make_inputs
    wl = configs.cfg2_production_256()
    inp = make_inputs(wl)
    out = {}
    for prec in ("f32", "f64"):
        p = Propagator(inp, prec)
        p.upload(inp.A)
        p.calibrate()
        p.autotune()
        r = p.run_adaptive(0.0, inp.params.dzz, 1, warmup)
        t0 = time.perf_counter()
        r2 = p.run_adaptive(r["z"], r["dzz"], r["signmem"], steps)
        dt = time.perf_counter() - t0
        n = r2["accepted"]
        out[prec] = {"accepted_steps_per_s": round(n / dt, 3), "accepted": n, "seconds": round(dt, 4),
                     "z_advanced_m": r2["z"] - r["z"], "dzz_end": r2["dzz"],
                     "isolated_split_steps_per_s_lower_bound": round(3 * n / dt, 3)}
        log(f"[bench] adaptive {prec}: {n} accepted steps in {dt:.3f}s ({n / dt:.1f}/s), "
            f"z {r['z']:.4f} -> {r2['z']:.4f} m")
        p.close()
    return {"metric": "adaptive split-step (step doubling + Richardson), accepted steps/s at 256x256x2048",
            "value": out["f32"]["accepted_steps_per_s"], "unit": "accepted steps/s", "n_gpus": 1,
            "steps": steps, "warmup": warmup, "higher_is_better": True, "dtype": "f32",
            "data": "synthetic (seeded Gaussian pulse + 1e-6 complex noise, argon 2 bar)",
            "config": {"workload": "256x256x2048 adaptive mode, def_err 1e-3, start dzz = 1 mm",
                       "per_precision": out}}


def drift_bench(every=100):
    """FP32 against FP64 over the first pass of the 256x256x2048 workload (500
    steps and the mirror): rel L2 of the FP32 field against the FP64 field
    every `every` steps, both from the same input with the radix-16 x/y plans
    the tuner picks at this size (the FP64 path is the one the oracle pins)"""
    import numpy as np
    from mpc_amd import Propagator, configs, make_inputs
    wl = configs.cfg2_production_256()
    inp = make_inputs(wl)
    A0 = inp.calibrated()
    fields = {}
    for prec, var in (("f64", (1, 16, 28)), ("f32", (1, 16, 24))):
        with Propagator(inp, prec) as p:
            p.upload(A0)
            p.set_variants(*var)
            for k in range(0, wl.steps_per_pass, every):
                p.run(k, min(every, wl.steps_per_pass - k))
                fields[(prec, k + every)] = p.download()
        log(f"[bench] drift: {prec} pass done")
    out = {}
    for n in range(every, wl.steps_per_pass + 1, every):
        a, b = fields[("f32", n)], fields[("f64", n)]
        out[n] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
        log(f"[bench] drift: after {n} steps fp32 vs fp64 rel L2 {out[n]:.3e}")
    return {"metric": "FP32 vs FP64 field rel L2 over the first pass at 256x256x2048",
            "value": out[wl.steps_per_pass], "unit": "rel L2", "n_gpus": 1, "higher_is_better": False,
            "dtype": "f32 vs f64", "data": "synthetic (the 256x256x2048 input)", "vs_baseline": None,
            "config": {"workload": "256x256x2048, 500 steps per pass, mirror after step 499",
                       "rel_l2_by_steps": out}}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the 48-pulse batched leg of the default run")
    ap.add_argument("--sweep-variants", action="store_true", help="also time every kernel variant")
    ap.add_argument("--batch", action="store_true",
                    help="instead: the 48-pulse batched workload (FP32), batched vs one by one")
    ap.add_argument("--adaptive", action="store_true",
                    help="instead: adaptive mode (step doubling + Richardson), accepted steps/s, FP32 and FP64")
    ap.add_argument("--surrogate", action="store_true",
                    help="instead: the pass-surrogate experiment (pairs from the solver, training on a time "
                         "budget, held-out errors, assisted rollouts against plain ones)")
    ap.add_argument("--surrogate-pulses", type=int, default=512, help=argparse.SUPPRESS)
    ap.add_argument("--surrogate-passes", type=int, default=8, help=argparse.SUPPRESS)
    ap.add_argument("--train-seconds", type=float, default=240.0, help=argparse.SUPPRESS)
    ap.add_argument("--surrogate-width", type=int, default=32, help=argparse.SUPPRESS)
    ap.add_argument("--surrogate-layers", type=int, default=4, help=argparse.SUPPRESS)
    ap.add_argument("--drift", action="store_true",
                    help="instead: FP32 against FP64 every 100 steps over the first 500-step pass")
    ap.add_argument("--prec", choices=["both", "f32", "f64"], default="both",
                    help="f32 / f64 alone skips the other path and the fp32-vs-fp64 comparison (profiling runs)")
    ap.add_argument("--nt", type=int, default=2048, help=argparse.SUPPRESS)
    ap.add_argument("--nx", type=int, default=256, help=argparse.SUPPRESS)
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              f"(multi-GPU sharding is designed, not built: pulses shard with no exchange).", file=sys.stderr)
        return 2

    from mpc_amd import configs, lib, make_inputs
    if lib().mpc_device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        return 1
    if a.batch:
        print(json.dumps(batch_bench(a.steps, a.warmup)))
        return 0
    if a.adaptive:
        print(json.dumps(adaptive_bench(a.steps, a.warmup)))
        return 0
    if a.drift:
        print(json.dumps(drift_bench()))
        return 0
    if a.surrogate:
        from mpc_amd import surrogate
        from mpc_amd.fno import FnoConfig
        cfg = FnoConfig(width=a.surrogate_width, layers=a.surrogate_layers, hidden=2 * a.surrogate_width)
        r = surrogate.run_experiment(n_pulses=a.surrogate_pulses, passes=a.surrogate_passes,
                                     train_seconds=a.train_seconds, cfg=cfg, log=log)
        print(json.dumps({"metric": "surrogate-assisted passes/s at 64x64x256 (rel L2 per pass vs plain FP32)",
                          "value": r["passes_per_s_assisted"], "unit": "passes/s", "n_gpus": 1,
                          "higher_is_better": True, "dtype": "bf16 (network), f32 (solver)",
                          "data": "synthetic pulses drawn by seed", "vs_baseline": None,
                          "config": {"workload": "pass surrogate, 64x64x256, 200 z-steps per pass"},
                          "result": r}))
        return 0
    wl = configs.cfg2_production_256()
    wl.nx, wl.nt = a.nx, a.nt
    t0 = time.perf_counter()
    inp = make_inputs(wl)
    log(f"[bench] inputs {wl.nx}x{wl.nx}x{wl.nt} generated in {time.perf_counter() - t0:.1f}s")

    sweep = {}
    if a.sweep_variants:
        sweep = {"f32": variant_sweep(inp, "f32"), "f64": variant_sweep(inp, "f64")}
    sps64 = kms64 = var64 = rel = None
    if a.prec in ("both", "f64"):
        sps64, dt64, A64, kms64, bf64, var64 = timed_run(inp, "f64", a.warmup, a.steps)
    if a.prec == "f64":                       # profiling run of the FP64 path alone
        sps32, dt32, A32, kms32, bf32, var32 = sps64, dt64, None, kms64, bf64, var64
    else:
        sps32, dt32, A32, kms32, bf32, var32 = timed_run(inp, "f32", a.warmup, a.steps)
    if a.prec == "both":
        rel = float(np.linalg.norm((A32 - A64).ravel()) / np.linalg.norm(A64.ravel()))
        log(f"[bench] fp32 vs fp64 rel L2 after {a.warmup + a.steps} steps: {rel:.3e}")
        del A64
    del A32
    dtype = "f64" if a.prec == "f64" else "f32"

    # dominant kernel of the measured path: the largest share of step time
    dom = int(np.argmax(kms32))
    per_launch_bytes = 2 * bf32           # read + write of the field
    achieved = per_launch_bytes / (kms32[dom] * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": load_traffic(dtype),
            "kernel": KERNELS[dom], "kernel_ms": round(kms32[dom], 4),
            "all_kernels_ms": {KERNELS[i]: round(kms32[i], 4) for i in range(3)},
            "algorithmic_bytes_per_launch": per_launch_bytes}
    # the t-line is bound by its vector ALUs, not by HBM: its VALU issue floor
    # against the measured launch time
    valu = load_valu(dtype, var32[0]) if dom == 0 else None
    if valu is not None:
        valu["frac"] = round(valu["issue_floor_ms"] / kms32[dom], 4)
        roof["valu"] = valu

    cpu = vs_oracle = None
    if not a.no_cpu_baseline and int(os.environ.get("RANK", "0")) == 0:
        cpu, A_ref, n_ref, rho_ref = cpu_baseline(inp)
        if a.prec == "both":
            vs_oracle = against_oracle(inp, A_ref, n_ref, rho_ref)
        del A_ref
    batched = None
    if not a.no_batch and a.prec == "both" and (wl.nx, wl.nt) == (256, 2048):
        # the 48-pulse workload, batched and one by one (four calls of
        # bsteps steps stay inside the first 300-step pass)
        bsteps = max(1, min(a.steps, 60))
        br = batch_bench(bsteps, bsteps)
        batched = {"equivalent_steps_per_s": br["value"],
                   "one_by_one_equivalent_steps_per_s": br["config"]["one_by_one_steps_per_s"],
                   "batched_over_one_by_one": br["config"]["cfg4"]["batched_over_one_by_one"],
                   "steps": bsteps, "pulses": 48, "points_total": br["config"]["points_total"],
                   "workload": br["config"]["workload"], "unit": "256x256x2048-equivalent steps/s",
                   "dtype": "f32", "small_pulses_64x64x256": br["config"]["small_pulses_64x64x256"]}

    out = {
        "metric": METRIC, "value": round(sps32, 3), "unit": "steps/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(1e3 * dt32 / a.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
        "data": "synthetic (seeded Gaussian pulse + 1e-6 complex noise, argon 2 bar)",
        "config": {"workload": f"{wl.nx}x{wl.nx}x{wl.nt} split-step, 1 mJ 300 fs, Ar 2 bar, ADK + SST on, "
                               f"dz = 1 mm (500 steps/pass), curved mirror R = 0.3 m",
                   "precision": ("fp64 complex" if dtype == "f64" else
                                 "fp32 complex, fp64 propagator phase / ADK exponent / density scan"),
                   "plain_fp64_steps_per_s": round(sps64, 3) if sps64 else None,
                   "passes_per_s": {"f32": round(sps32 / wl.steps_per_pass, 4),
                                    "f64": round(sps64 / wl.steps_per_pass, 4) if sps64 else None,
                                    "steps_per_pass": wl.steps_per_pass},
                   "fp32_over_fp64": round(sps32 / sps64, 3) if a.prec == "both" else None,
                   "fp32_vs_fp64_rel_l2": rel,
                   "vs_oracle": vs_oracle,
                   "batched": batched,
                   "fp64_kernel_ms": {KERNELS[i]: round(kms64[i], 4) for i in range(3)} if kms64 else None,
                   "autotuned_variants": {"f32": var32, "f64": var64}},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if sweep:
        out["variant_sweep_ms"] = sweep
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
