"""GPU: mpc_linear, the linear part of propagation in one go, against the
fixed-step schedule (mpc_run) of the same pulse with the nonlinear step
switched off (n2 = 0, no plasma).  The nonlinear step then multiplies by
exp(0) = 1, so the two routes differ only by rounding: the merged tables
exp(i beta L) and exp(-i L k^2 / 2k0) against products of per-step factors.

Bar: 1e-10 rel L2 (FP64) and 1e-4 (FP32), within a pass (no mirror) and
over a whole pass with the mirror after it, for the curved mirror and for
the reference's sign-flip mirror (diffraction sign -1 on odd passes).
"""
import numpy as np
import pytest

from mpc_amd import Propagator, configs, make_inputs

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-10, "f32": 1e-4}


def _linear_inputs(mirror_R, S=6):
    wl = configs.small(nx=32, nt=256, steps_per_pass=S, passes=3, mirror_R=mirror_R, seed=21)
    wl.n2_per_bar = 0.0
    wl.plasm = 0
    inp = make_inputs(wl)
    return inp, inp.calibrated()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("mirror_R", [0.3, 0.0])
def test_linear_equals_run_without_nonlinearity(gpu_lib, prec, mirror_R):
    S = 6
    inp, A0 = _linear_inputs(mirror_R, S)
    dz = inp.params.dzz
    print(f"case linear prec={prec} mirror_R={mirror_R}", flush=True)
    cases = [(0, S // 2, 1, False),          # part of pass 0, no mirror
             (0, S, 1, True),                # all of pass 0, mirror after it
             (S, S, -1 if mirror_R == 0.0 else 1, True)]   # pass 1: sign -1 for the sign-flip mirror
    for step0, n, sign, mirror in cases:
        ref = Propagator(inp, prec)
        ref.upload(A0)
        if step0:
            ref.run(0, step0)
        start = ref.download()
        ref.run(step0, n)
        want = ref.download()
        ref.close()
        p = Propagator(inp, prec)
        p.upload(start)
        p.linear(n * dz, sign, mirror)
        got = p.download()
        p.close()
        err = _rel(got, want)
        change = _rel(start, want)
        print(f"  steps {step0}+{n} sign {sign:+d} mirror {mirror}: rel err {err:.2e} "
              f"(field change {change:.2e})", flush=True)
        assert change > 1e-3                  # the map is not trivially the identity
        assert err <= TOL[prec]


def test_linear_rejects_bad_arguments(gpu_lib):
    inp, A0 = _linear_inputs(0.3)
    p = Propagator(inp, "f32")
    p.upload(A0)
    with pytest.raises(RuntimeError):
        p.linear(-1.0)
    with pytest.raises(RuntimeError):
        p.linear(0.01, sign=2)
    p.close()
