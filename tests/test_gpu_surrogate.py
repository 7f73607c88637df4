"""GPU: the pass surrogate's data and its fallback.

* Training pairs are the solver's passes: each stored mirror-plane field
  equals a per-pulse run of the same pulse (FP32 bar, 1e-4 rel L2: the batch
  and the per-pulse run may take different kernel variants), and each stored
  linear-map input is mpc_linear applied to the stored field (bitwise).
* Fallback: with checks that always fail, every network pass is replaced by
  the plain pass, so the assisted trajectory is the plain one (up to the
  float round trip of export / import, 1e-5); with checks that always pass,
  the network takes every pass but the plain_every-th.
* Training gain: 400 Adam steps on 4 solver pairs bring a width-32
  network's error on them below 0.7 of the linear pass map's (the method
  without the network).
"""
import math

import numpy as np
import pytest

from mpc_amd import Propagator, make_inputs, surrogate
from mpc_amd.fno import DArr, FnoConfig, FnoGpu, init_params

pytestmark = pytest.mark.gpu


def test_pairs_are_solver_passes(gpu_lib):
    n_p, passes, seed = 2, 2, 9
    print("case surrogate pairs", flush=True)
    data = surrogate.generate(n_p, passes, seed=seed, log=lambda *a: None)
    wls = surrogate.pulse_workloads(n_p, seed, passes + 1)
    n = data.npts
    S = wls[0].steps_per_pass
    for i, wl in enumerate(wls):
        inp = make_inputs(wl)
        p = Propagator(inp, "f32")
        p.upload(inp.A)
        p.calibrate()
        sc = surrogate.field_scale(p)
        own = DArr((passes + 1, 2, n))
        surrogate.export(p, own, 0, sc)
        for k in range(passes):
            p.run(k * S, S)
            surrogate.export(p, own, (k + 1) * 2 * n, sc)
        stored = DArr((passes + 1, 2, n))
        stored.copy_from(data.states, (passes + 1) * 2 * n, src_off=i * (passes + 1) * 2 * n)
        a, b = own.numpy(), stored.numpy()
        for k in range(passes + 1):
            err = float(np.linalg.norm(a[k] - b[k]) / np.linalg.norm(a[k]))
            print(f"  pulse {i} state {k}: rel err {err:.2e}", flush=True)
            assert err <= 1e-4
        lin = DArr((2, n))
        want = DArr((passes, 2, n))
        want.copy_from(data.lin, passes * 2 * n, src_off=i * passes * 2 * n)
        w = want.numpy()
        for k in range(passes):
            surrogate.import_(p, data.states, (i * (passes + 1) + k) * 2 * n, sc)
            surrogate.linear_pass(p, wl)
            surrogate.export(p, lin, 0, sc)
            assert np.array_equal(lin.numpy(), w[k])
        p.close()


def _net():
    cfg = FnoConfig()
    return FnoGpu(cfg, init_params(cfg, seed=3))


def test_failed_checks_fall_back_to_the_plain_pass(gpu_lib):
    print("case surrogate fallback", flush=True)
    wl = surrogate.pulse_workloads(1, seed=5, passes=7)[0]
    # zero tolerances: no network pass keeps energy and width exactly
    r = surrogate.rollout(_net(), wl, 6, plain_every=3, energy_tol=0.0, width_tol=0.0)
    print(f"  kinds {r['kinds']} max err {max(r['errs']):.2e}", flush=True)
    assert r["kinds"] == ["fallback", "fallback", "plain"] * 2
    assert max(r["errs"]) <= 1e-5


def test_passed_checks_keep_the_network_pass(gpu_lib):
    print("case surrogate accepted", flush=True)
    wl = surrogate.pulse_workloads(1, seed=5, passes=7)[0]
    r = surrogate.rollout(_net(), wl, 6, plain_every=3, energy_tol=math.inf, width_tol=math.inf)
    print(f"  kinds {r['kinds']} errs {['%.2e' % e for e in r['errs']]}", flush=True)
    assert r["kinds"] == ["net", "net", "plain"] * 2
    assert all(np.isfinite(r["errs"]))
    assert r["net_forward_ms"] is not None and r["net_forward_ms"] > 0


def test_training_beats_the_linear_map_on_its_pairs(gpu_lib):
    """a few hundred Adam steps on a handful of solver pairs bring the
    network's error on those pairs below the linear pass map's (the method
    without the network); generalisation to unseen pulses is measured by
    bench.py --surrogate, not here"""
    print("case surrogate training gain", flush=True)
    data = surrogate.generate(2, 2, seed=13, log=lambda *a: None)
    ident, linear = surrogate.baseline_errors(data, [0, 1])
    # width 32, the configured network (width 16 needs more steps to pass the
    # linear map: about 800 at this learning rate in the full experiment)
    tr = surrogate.Trainer(FnoConfig(width=32, hidden=64), data, batch=4, seed=2)
    first = tr.eval_loss([0, 1])
    for i in range(400):
        tr.step(2e-3 * (1.0 - i / 400))
    last = tr.eval_loss([0, 1])
    print(f"  pair rel L2: no change {ident:.3e}, linear map {linear:.3e}, network {first:.3e} -> {last:.3e}",
          flush=True)
    assert last < 0.7 * linear        # measured: 0.51x after 300 steps
