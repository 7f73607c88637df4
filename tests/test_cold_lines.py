"""The bounds behind the t-line's cold lines, checked on the FP64 oracle.

The t-line stops a line's density scans when two bounds make its plasma part
smaller than the field's rounding:
* the recurrence R_k = (R_{k+1} + b_k)(1 - a_k) + a_k, R_NT = 0, with
  a, b >= 0 and S = sum_k X_k <= 1, X_k = b_k (1 - a_k) + a_k, gives (without
  the second species) ionised fractions u_k = 1 - exp(-R_k) <= S, largest at
  k = 0, with u_0 >= S (1 - 3S/2);
* sigma(tau) (1 + w0 tau) <= K4 / (2 sqrt K5) + w0 K4 / K5 for every collision
  time tau, with sigma = K4 tau / (1 + K5 tau^2).
Both are checked here against the oracle's recurrence (plasma.c:218-241) and
its cross-section (plasma.c:207-210), over magnitudes from far below to near
the thresholds the kernel uses.
"""
import numpy as np
import pytest

import oracle_lib as O
from mpc_amd import configs, make_inputs

C0 = 299792458.0


def _line(b, a, c, sig):
    nt = len(b)
    F, rf = np.zeros(nt), np.zeros(nt)
    O.oracle().oracle_density_line(O.dp(b), O.dp(a), O.dp(c), O.dp(sig), nt, O.dp(F), O.dp(rf), O.OQ_PHYSICAL)
    return F, rf


@pytest.mark.parametrize("S_target", [1e-20, 1e-13, 1e-7, 1e-3, 0.05])
@pytest.mark.parametrize("shape", ["uniform", "gaussian", "spike"])
def test_ionised_fraction_is_bounded_by_the_sum_of_scan_constants(S_target, shape):
    rng = np.random.default_rng(11)
    nt = 512
    t = np.arange(nt)
    env = {"uniform": np.ones(nt),
           "gaussian": np.exp(-((t - 300) / 40.0) ** 2),
           "spike": (t == 17).astype(float) + 1e-9}[shape]
    a = env * rng.uniform(0.2, 1.0, nt)
    b = env * rng.uniform(0.0, 3.0, nt) * (t % 3 == 0)
    X = b * (1 - a) + a
    k = S_target / X.sum()
    a, b = a * k, b * k
    X = b * (1 - a) + a
    S = X.sum()
    sig = rng.uniform(0.5, 2.0, nt)
    F, u = _line(b, a, np.zeros(nt), sig)
    assert u.max() == u[0]
    assert u[0] <= S * (1 + 1e-13)
    assert u[0] >= S * (1 - 1.5 * S) * (1 - 1e-13)
    # the plasma part per element: F = u sigma <= S sigma
    assert np.all(F <= S * sig * (1 + 1e-13))


def test_the_sum_of_a_alone_is_not_a_bound():
    """the field-ionisation term b must be in the sum: with b carrying the
    line, the fraction exceeds sum a by orders of magnitude"""
    nt = 256
    a = np.full(nt, 1e-15)
    b = np.zeros(nt)
    b[::8] = 1e-10
    _, u = _line(b, a, np.zeros(nt), np.ones(nt))
    assert u[0] > 1e3 * a.sum()
    assert u[0] <= (b * (1 - a) + a).sum() * (1 + 1e-13)


@pytest.mark.parametrize("energy", [3e-3, 1e-6, 1e-12])
def test_cross_section_bound_holds_on_the_oracle(energy):
    """sigma (1 + w0 tau) from oracle_ioniz's outputs (sigma = F / rho_frac)
    stays below the closed-form bound, which weak fields approach"""
    inp = make_inputs(configs.small(8, 64, energy=energy))
    A = inp.calibrated()
    P = inp.params
    N = A.size
    tau, F, rf = np.zeros(N), np.zeros(N), np.zeros(N)
    O.oracle().oracle_ioniz(O.vp(A), O.ip(O.dims_arr(A.shape)), P.w0, P.deltat, P.gas, P.rho_c, O.dp(inp.w),
                            P.rho_nt, P.n0, O.dp(tau), O.dp(F), O.dp(rf), O.OQ_PHYSICAL)
    K4 = (P.w0 / C0) * P.w0 / (P.n0 * P.rho_c)
    K5 = P.w0 ** 2
    bound = K4 / (2 * np.sqrt(K5)) + P.w0 * K4 / K5
    ok = rf > 0
    assert ok.sum() > N // 2
    m = F[ok] / rf[ok] * (1 + P.w0 * tau[ok])
    assert m.max() <= bound * (1 + 1e-12)
    if energy <= 1e-6:
        assert m.max() >= 0.5 * bound


def test_cross_section_bound_closed_form():
    """the bound is the maximum over tau: K4/(2 sqrt K5) at tau = 1/sqrt K5
    for sigma, K4/K5 as tau -> inf for sigma tau"""
    w0 = 2 * np.pi * C0 / 1030e-9
    K4, K5 = 3.7e-3, w0 ** 2
    tau = np.logspace(-22, -6, 20001)
    sig = K4 * tau / (1 + K5 * tau ** 2)
    bound = K4 / (2 * np.sqrt(K5)) + w0 * K4 / K5
    assert np.all(sig * (1 + w0 * tau) <= bound * (1 + 1e-12))
    assert sig.max() == pytest.approx(K4 / (2 * np.sqrt(K5)), rel=1e-6)
    assert (sig * tau)[-1] == pytest.approx(K4 / K5, rel=1e-9)
