"""GPU: batched runs (many pulses, one launch per sweep and launch group of
same-size pulses, a HIP graph per pass) against per-pulse runs and the oracle.

Bar: each pulse's batched result is bit-identical to mpc_run on its own
context with the same kernel variants (the batch takes each size class's
first pulse's: persistent variants walk the line groups or tiles of every
pulse in the launch, the others take one workgroup per group or tile); one
pulse is also checked against the FP64 oracle at 1e-10.
Graph replays are interleaved with host downloads and eager work, and every
replay's result is compared.
"""
import numpy as np
import pytest

import oracle_lib as O
from mpc_amd import Batch, Propagator, configs, make_inputs

pytestmark = pytest.mark.gpu

S, PASSES = 3, 3
# uneven grids, energies, pressures and waists; two pulses share a size class
SPECS = [dict(nx=32, nt=512, energy=1e-3, pressure=2.0, seed=11),
         dict(nx=64, nt=512, energy=3e-4, pressure=1.0, seed=12, waist=0.3e-3),
         dict(nx=32, nt=1024, energy=2e-4, pressure=2.5, seed=13),
         dict(nx=32, nt=512, energy=2e-4, pressure=0.5, seed=14, waist=0.25e-3)]


def _inputs(mirror_R=0.3):
    out = []
    for s in SPECS:
        inp = make_inputs(configs.small(steps_per_pass=S, passes=PASSES, mirror_R=mirror_R, **s))
        out.append((inp, inp.calibrated()))
    return out


def _props(cases, prec, variants):
    props = []
    for inp, A0 in cases:
        p = Propagator(inp, prec)
        p.upload(A0)
        p.set_variants(*variants)
        props.append(p)
    return props


def _close(props):
    for p in props:
        p.close()


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("mirror_R", [0.3, 0.0])
@pytest.mark.parametrize("var", [(1, 2, 3), (0, 18, 19), (0, 2, 27)])
def test_batch_is_bitwise_per_pulse(gpu_lib, prec, mirror_R, var):
    """one call over all passes, with and without graphs, against per-pulse runs"""
    nsteps = S * PASSES
    cases = _inputs(mirror_R)
    print(f"case batch one-shot prec={prec} mirror_R={mirror_R} variants={var}", flush=True)
    # per-pulse runs with the variants under test
    ref = []
    for p in _props(cases, prec, var):
        p.run(0, nsteps)
        ref.append(p.download())
        assert np.isfinite(ref[-1]).all()
        d = p.diagnostics()
        p.close()
    for graph in (False, True):
        props = _props(cases, prec, var)
        with Batch(props) as b:
            b.run(0, nsteps, graph=graph)
            for i, p in enumerate(props):
                got = p.download()
                assert np.array_equal(got, ref[i]), (prec, mirror_R, var, graph, i)
            if graph:
                print(f"  graphs captured: {b.graphs()}", flush=True)
                assert 1 <= b.graphs() <= PASSES
        # the batch leaves each context usable: diagnostics read its last step
        dd = props[-1].diagnostics()
        assert dd["rho_peak"] == d["rho_peak"]
        _close(props)


@pytest.mark.parametrize("prec,tol", [("f64", 1e-10), ("f32", 1e-4)])
@pytest.mark.parametrize("var", [(0, 18, 19), (1, 2, 3)])
def test_batch_pass_by_pass_graph_replays(gpu_lib, prec, tol, var):
    """one call per pass: interior passes replay the same graph; every replay
    is compared with per-pulse runs split the same way, and pulse 0 with the
    oracle"""
    cases = _inputs(0.3)
    print(f"case batch pass-by-pass prec={prec} variants={var}", flush=True)
    ref_props = _props(cases, prec, var)
    props = _props(cases, prec, var)
    oracle_A = cases[0][1]
    with Batch(props) as b:
        for k in range(PASSES):
            for p in ref_props:
                p.run(k * S, S)
            b.run(k * S, S, graph=True)
            oracle_A, _ = O.run_fixed(cases[0][0], oracle_A, k * S, S)
            for i, p in enumerate(props):
                got = p.download()          # host work between replays
                assert np.array_equal(got, ref_props[i].download()), (prec, var, k, i)
            err = O.rel_l2(props[0].download(), oracle_A)
            print(f"  pass {k}: graphs={b.graphs()} pulse-0 rel_l2 vs oracle {err:.3e}", flush=True)
            assert err <= tol
        # pass 0 starts from the uploaded (time-domain) fields, passes 1 and 2
        # from the t-spectral fields the pass before left, and the last pass
        # has no mirror after it: at most three graphs, pass 1's launch list
        # being the one every further interior pass would replay
        assert b.graphs() <= 3
    _close(props)
    _close(ref_props)


# every sweep kind groups these pulses the same way, so each launch group runs
# its sweeps back to back; the three 128^2 x 1024 fields (134 MB each in FP32,
# 268 MB in FP64) also split their class into several launch groups
GROUPED_SPECS = [dict(nx=128, nt=1024, energy=1e-3, pressure=2.0, seed=21),
                 dict(nx=32, nt=512, energy=3e-4, pressure=1.0, seed=22),
                 dict(nx=128, nt=1024, energy=2e-4, pressure=2.5, seed=23, waist=0.3e-3),
                 dict(nx=128, nt=1024, energy=5e-4, pressure=0.5, seed=24)]


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("var", [(0, 18, 19), (1, 2, 3)])
def test_group_major_batch_is_bitwise_per_pulse(gpu_lib, prec, var):
    """launch groups run one after another within a pass (the launch order
    differs from the sweep-major one), with the mirror inside: each pulse
    still equals its own per-pulse run bit for bit"""
    nsteps = S * PASSES
    cases = []
    for s in GROUPED_SPECS:
        inp = make_inputs(configs.small(steps_per_pass=S, passes=PASSES, mirror_R=0.3, **s))
        cases.append((inp, inp.calibrated()))
    print(f"case batch group-major prec={prec} variants={var}", flush=True)
    ref = []
    for p in _props(cases, prec, var):
        p.run(0, nsteps)
        ref.append(p.download())
        p.close()
    for graph in (False, True):
        props = _props(cases, prec, var)
        with Batch(props) as b:
            b.run(0, nsteps, graph=graph)
            for i, p in enumerate(props):
                assert np.array_equal(p.download(), ref[i]), (prec, var, graph, i)
        _close(props)


def test_batch_rejects_mixed_schedules(gpu_lib):
    a = make_inputs(configs.small(32, 512, steps_per_pass=3, passes=3))
    c = make_inputs(configs.small(32, 512, steps_per_pass=4, passes=3))
    p1, p2 = Propagator(a, "f32"), Propagator(c, "f32")
    with pytest.raises(RuntimeError):
        Batch([p1, p2])
    p3 = Propagator(a, "f64")
    with pytest.raises(RuntimeError):
        Batch([p1, p3])
    _close([p1, p2, p3])


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_batch_autotune_leaves_fields_and_keeps_results(gpu_lib, prec):
    """mpc_batch_autotune times the batch's own launches on scratch copies:
    the fields and the plasma maxima are left as they were, each launch class
    gets one variant, and the batched run with the chosen variants is still
    bit-identical to per-pulse runs with them"""
    print(f"case batch autotune prec={prec}", flush=True)
    cases = _inputs()
    props = _props(cases, prec, (0, 2, 3))
    for p in props:
        p.run(0, 2)
    before = [p.download() for p in props]
    rho = [p.diagnostics()["rho_peak"] for p in props]
    with Batch(props) as b:
        b.autotune()
        chosen = [p.variants() for p in props]
        print(f"  variants chosen {chosen}", flush=True)
        for i, p in enumerate(props):
            assert np.array_equal(p.download(), before[i])
            assert p.diagnostics()["rho_peak"] == rho[i]
        # one t-line form per (t length, plasma) class: pulses 0 and 3 share
        # one; pulses 0, 2 and 3 (32-point lines) share x and y variants;
        # pulse 1 is alone in its x/y class and keeps its own
        assert chosen[0][0] == chosen[3][0]
        assert chosen[0][1:] == chosen[2][1:] == chosen[3][1:]
        assert chosen[1][1:] == (2, 3)
        b.run(2, 4, graph=True)
        got = [p.download() for p in props]
    _close(props)
    for i, (inp, A0) in enumerate(cases):
        with Propagator(inp, prec) as p:
            p.upload(A0)
            p.set_variants(0, 2, 3)
            p.run(0, 2)
            p.set_variants(*chosen[i])
            p.run(2, 4)
            assert np.array_equal(p.download(), got[i]), (prec, i)
