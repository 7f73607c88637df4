"""Known-answer tests that pin the oracle's plasma restatement (plasma.c):
constants, the ADK closed form, collision time / cross-section closed forms,
and the reverse-t density recurrence by brute force and special cases."""
import ctypes as C
from math import gamma, pi, sqrt

import numpy as np
import pytest

import oracle_lib as O
from mpc_amd import configs, make_inputs

IH = 27.21
GASES = {  # plasma.c:4-40 and :61-115 (I1, I2 in eV; sigma_coll)
    "Argon": (15.75962, 27.62967, 1.57e-20), "Helium": (24.58741, 54.41778, 6.11e-20),
    "Neon": (21.5646, 40.96328, 1.65e-20), "Nitrogen": (14.53414, 29.6013, 10.2e-20),
    "Xenon": (12.1298, 21.20979, 3.55e-20), "Krypton": (13.99961, 24.35985, 1.15e-20),
}


@pytest.mark.parametrize("gas", sorted(GASES))
def test_gas_tables(gas):
    I1, I2, s = C.c_double(), C.c_double(), C.c_double()
    O.oracle().oracle_ionizpot(gas.encode(), C.byref(I1), C.byref(I2))
    n, sc = C.c_double(), C.c_double()
    O.oracle().oracle_natoms(gas.encode(), C.byref(n), C.byref(sc))
    assert (I1.value, I2.value, sc.value) == GASES[gas]


def _kon(q):
    k = np.zeros(6)
    O.oracle().oracle_adk_consts(b"Argon", q, O.dp(k))
    return k


def test_adk_constants_physical_closed_form():
    Ei0, Ei1 = 15.75962 / IH, 27.62967 / IH
    n0, n1 = (2 * Ei0) ** -0.5, 2 * (2 * Ei1) ** -0.5
    c0 = sqrt(2 ** (2 * n0) / (n0 * gamma(n0 + 1) * gamma(n0)))
    c1 = sqrt(2 ** (2 * n1) / (n1 * gamma(n1 + 1) * gamma(n1)))
    ref = [c0 ** 2 * sqrt(6 / pi) * Ei0, c1 ** 2 * sqrt(6 / pi) * Ei1, 2 * (2 * Ei0) ** 1.5, 2 * (2 * Ei1) ** 1.5,
           2 * n0 - 1.5, 2 * n1 - 1.5]
    np.testing.assert_allclose(_kon(O.OQ_PHYSICAL), ref, rtol=1e-14)
    assert abs(2 * n0 - 1.5 - 0.358) < 1e-3          # n* of argon ~ 0.929


def test_adk_constants_literal_quirks():
    k = _kon(O.OQ_LITERAL)
    nq0 = (2 * 27.62967 / IH) ** -0.5               # Q5: Iion[1] used for the first ionisation
    assert abs(k[4] - (2 * nq0 - 1.5)) < 1e-15 and k[4] < 0
    c0 = sqrt(2 ** (2 * nq0) / (nq0 * gamma(nq0 + 1) * gamma(nq0)))
    assert 1 < c0 < 2                                # Q10: int abs truncates to 1
    assert abs(k[0] - 1.0 * sqrt(6 / pi) * 15.75962 / IH) < 1e-15
    assert k[1] == 0.0                               # Q6: W2 := 0
    # Q7 alone: C_nl[1] uses 2^n* instead of 2^(2 n*)
    k7 = _kon(O.Q["Q7"])
    kp = _kon(O.OQ_PHYSICAL)
    n1 = 2 * (2 * 27.62967 / IH) ** -0.5
    assert abs(k7[1] / kp[1] - 2.0 ** (-n1)) < 1e-14


def test_adk_rate_closed_form_and_nan_rule():
    k = _kon(O.OQ_PHYSICAL)
    E = np.array([1e-3, 3e-3, 0.01, 0.03, 0.1, 0.3])
    W1, W2 = np.zeros(6), np.zeros(6)
    O.oracle().oracle_adk(O.dp(W1), O.dp(W2), 6, O.dp(k), O.dp(E))
    ref1 = k[0] * np.exp(k[4] * np.log(k[2] / E) - k[2] / (3 * E))
    ref2 = k[1] * np.exp(k[5] * np.log(k[3] / E) - k[3] / (3 * E))
    np.testing.assert_allclose(W1, ref1, rtol=1e-12)
    np.testing.assert_allclose(W2, ref2, rtol=1e-12)
    # E = 0: pow(inf, >0) * exp(-inf) is NaN -> 0 (plasma.c:184-192)
    Z = np.zeros(1)
    w1, w2 = np.ones(1), np.ones(1)
    O.oracle().oracle_adk(O.dp(w1), O.dp(w2), 1, O.dp(k), O.dp(Z))
    assert w1[0] == 0.0 and w2[0] == 0.0


def _recurrence_brute(b, a, c, sig, q_literal):
    nt = len(b)
    if q_literal:
        ions1, ions2 = np.zeros(nt + 1), np.zeros(nt + 1)
        R = 0.0
        for k in range(nt - 1, -1, -1):
            R = R + b[k]
            R = R + (1 - R) * a[k]
            ions1[k + 1] = 1 - np.exp(-R) - ions2[k]
            ions2[k + 1] = ions2[k] + c[k] * ions1[k]
        fr = ions1[:nt] + 2 * ions2[:nt]
    else:
        R, i1, i2 = 0.0, 0.0, 0.0
        fr = np.zeros(nt)
        for k in range(nt - 1, -1, -1):
            R = R + b[k]
            R = R + (1 - R) * a[k]
            u = -np.expm1(-R)
            i2n = i2 + c[k] * i1
            i1n = u - i2n
            fr[k] = i1n + 2 * i2n
            i1, i2 = i1n, i2n
    return fr * sig, fr


def _line(b, a, c, sig, q):
    nt = len(b)
    F, rf = np.zeros(nt), np.zeros(nt)
    O.oracle().oracle_density_line(O.dp(b), O.dp(a), O.dp(c), O.dp(sig), nt, O.dp(F), O.dp(rf), q)
    return F, rf


@pytest.mark.parametrize("q", [O.OQ_PHYSICAL, O.Q["Q8"]])
def test_recurrence_brute_force(q):
    rng = np.random.default_rng(4)
    nt = 200
    b, a, c = rng.uniform(0, 1e-3, nt), rng.uniform(0, 1e-3, nt), rng.uniform(0, 1e-2, nt)
    sig = rng.uniform(1, 2, nt)
    F, rf = _line(b, a, c, sig, q)
    Fr, rfr = _recurrence_brute(b, a, c, sig, q != O.OQ_PHYSICAL)
    np.testing.assert_allclose(F, Fr, rtol=1e-13, atol=1e-300)
    np.testing.assert_allclose(rf, rfr, rtol=1e-13, atol=1e-300)


def test_recurrence_special_cases():
    nt = 64
    b = np.full(nt, 1e-12)
    z = np.zeros(nt)
    one = np.ones(nt)
    # a = c = 0: R_k = b (Nt - k); PHYSICAL rho_k = 1 - exp(-R_k) ~ R_k
    _, rf = _line(b, z, z, one, O.OQ_PHYSICAL)
    R = 1e-12 * (nt - np.arange(nt))
    np.testing.assert_allclose(rf, -np.expm1(-R), rtol=1e-14)
    # LITERAL (Q8): rho_k uses R through sample k-1 and rho_0 = 0; 1-exp(-R) cancels for tiny R
    bl = np.full(nt, 1e-6)            # large enough that 1 - exp(-R) does not cancel
    Rl = 1e-6 * (nt - np.arange(nt))
    _, rfl = _line(bl, z, z, one, O.Q["Q8"])
    assert rfl[0] == 0.0
    np.testing.assert_allclose(rfl[1:], 1 - np.exp(-Rl[:-1]), rtol=1e-8)
    _, tiny = _line(np.full(nt, 1e-20), z, z, one, O.Q["Q8"])
    assert not tiny.any()                              # cancels to exactly 0 in the reference
    _, tinyp = _line(np.full(nt, 1e-20), z, z, one, O.OQ_PHYSICAL)
    np.testing.assert_allclose(tinyp, 1e-20 * (nt - np.arange(nt)), rtol=1e-12)
    # b = 0: R_k = 1 - prod_{j>=k} (1 - a_j)  (avalanche term alone)
    a = np.linspace(1e-4, 2e-4, nt)
    _, rf = _line(z, a, z, one, O.OQ_PHYSICAL)
    Ra = 1 - np.cumprod((1 - a)[::-1])[::-1]
    np.testing.assert_allclose(rf, -np.expm1(-Ra), rtol=1e-12)


def test_ioniz_tau_sigma_closed_forms():
    inp = make_inputs(configs.small(8, 64, energy=3e-3))
    A = inp.calibrated()
    P = inp.params
    N = A.size
    tau, F, rf = np.zeros(N), np.zeros(N), np.zeros(N)
    dims = O.dims_arr(A.shape)
    O.oracle().oracle_ioniz(O.vp(A), O.ip(dims), P.w0, P.deltat, P.gas, P.rho_c, O.dp(inp.w), P.rho_nt, P.n0,
                            O.dp(tau), O.dp(F), O.dp(rf), O.OQ_PHYSICAL)
    hbar, e, me = 1.054560652926899e-34, 1.602e-19, 9.1094e-31
    I = np.abs(A.ravel()) ** 2 * hbar * P.w0
    E0 = 0.89 * np.sqrt(2 * 377 * I)
    ve = e * E0 / (np.sqrt(2) * me * P.w0)          # quiver velocity at the carrier
    tref = 1 / (1.57e-20 * P.rho_nt * ve)
    np.testing.assert_allclose(tau, tref, rtol=1e-12)
    sig = (P.w0 / 299792458.0) * P.w0 / (P.n0 * P.rho_c) * tref / (1 + P.w0 ** 2 * tref ** 2)
    np.testing.assert_allclose(F, rf * sig, rtol=1e-12)
    assert rf.max() > 0


@pytest.mark.parametrize("bit", ["Q5", "Q10", "Q11", "Q12", "Q21"])
def test_plasma_quirk_bits_touch_only_plasma_terms(bit):
    """each plasma quirk bit changes the step when plasma is on and nothing when it is off"""
    for plasm in (0, 1):
        inp = make_inputs(configs.small(8, 64, energy=3e-3, plasm=plasm))
        A = inp.calibrated()
        outs = []
        for q in (O.OQ_PHYSICAL, O.Q[bit]):
            P = O.params_from(inp, q)
            B = A.copy()
            O.oracle().oracle_nonlinear(O.vp(B), C.byref(P), inp.params.dzz, None)
            outs.append(B)
        d = O.rel_l2(outs[1], outs[0])
        if plasm:
            assert d > 0
        else:
            assert d == 0
