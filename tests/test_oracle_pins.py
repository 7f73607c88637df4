"""Independent pins of the oracle's plasma and nonlinear-step compositions.

The pieces are pinned elsewhere (ADK constants, tau_c / sigma closed forms,
the recurrence by brute force, S by the analytic derivative).  These tests pin
what glues them together, each against numbers built here from the physics
formulas of the model, not from the oracle's code:

* oracle_ioniz (plasma.c:118-246): ionised fraction and F from the ADK rate,
  the Drude collision time and cross-section, the avalanche rate and the
  reverse-t two-species recurrence, with b = W1 dt/T0, a = sigma I dt/(I1 e),
  c = W2 dt/T0;
* oracle_nonlinear (propagation.c:164-327): the self-steepening closed form of
  a real Gaussian, the Drude ratio of plasma phase to plasma absorption, and
  the whole A_nl = i gamma |A|^2 - S - (1 + i w0 tau_c) rho_nt F / 2 assembly.

Each test also shows that it catches a plausible mistake: a quirk bit of the
oracle switched on (Q11, Q12, Q21), the sign of S flipped through a negative
w0, or a mis-scaled coefficient planted in the comparison numbers.
"""
import ctypes as C
from math import gamma as Gamma

import numpy as np
import pytest

import oracle_lib as O
from mpc_amd import configs, make_inputs

# physical constants of the model (constants.h:1-12)
HBAR = 1.054560652926899e-34
EL = 1.602e-19
ME = 9.1094e-31
C0 = 299792458.0
T0AU = 2.4188843265e-17
EF = 5.14220642e11
IH = 27.21
FCTR = 0.89
ZW0 = 377.0
ARGON = dict(I1=15.75962, I2=27.62967, sigma_coll=1.57e-20)   # eV, eV, m^2


def _intense(nx=8, nt=64, **kw):
    kw.setdefault("energy", 3e-3)
    inp = make_inputs(configs.small(nx, nt, **kw))
    return inp, inp.calibrated()


def _adk_rate(Ea, Ip_eV, Z):
    """ADK rate (atomic units) of an ion of charge Z, l = m = 0, with the
    effective quantum number n* = Z / sqrt(2 Ip):
    W = |C_n*|^2 sqrt(6/pi) Ip (2 F0/Ea)^(2n* - 3/2) exp(-2 F0 / (3 Ea)),
    |C_n*|^2 = 2^(2n*) / (n* Gamma(n*+1) Gamma(n*)), F0 = (2 Ip)^(3/2)"""
    Ip = Ip_eV / IH
    ns = Z / np.sqrt(2 * Ip)
    c2 = 2.0 ** (2 * ns) / (ns * Gamma(ns + 1) * Gamma(ns))
    F0 = (2 * Ip) ** 1.5
    with np.errstate(all="ignore"):
        W = c2 * np.sqrt(6 / np.pi) * Ip * (2 * F0 / Ea) ** (2 * ns - 1.5) * np.exp(-2 * F0 / (3 * Ea))
    return np.where(np.isnan(W), 0.0, W)


def _physics_ioniz(A, P, dt_scale=None, b_scale=None, c_scale=None):
    """tau_c, sigma, ionised fraction and F of the Drude + ADK + avalanche
    model for a field A (reference layout [t][y][x]), as numpy arrays of the
    same shape.  b_scale / dt_scale replace the coefficients' time factors
    (planted mistakes)."""
    w0, dt = P.w0, P.deltat
    I = np.abs(A) ** 2 * HBAR * w0
    I = np.where(I == 0, 1e-30, I)
    E0 = FCTR * np.sqrt(2 * ZW0 * I)
    Ea = E0 / EF
    W1 = _adk_rate(Ea, ARGON["I1"], 1)
    W2 = _adk_rate(Ea, ARGON["I2"], 2)
    ve = EL * E0 / (np.sqrt(2) * ME * w0)           # quiver velocity amplitude at w0
    tau = 1 / (ARGON["sigma_coll"] * P.rho_nt * ve)
    k0 = w0 / C0
    sig = k0 * w0 / (P.n0 * P.rho_c) * tau / (1 + (w0 * tau) ** 2)
    b = W1 * (b_scale if b_scale is not None else dt / T0AU)          # field ionisation per sample
    a = sig * I * (dt_scale if dt_scale is not None else dt) / (ARGON["I1"] * EL)   # avalanche per sample
    c = W2 * (c_scale if c_scale is not None else dt / T0AU)           # second ionisation per sample
    nt = A.shape[0]
    err = np.seterr(all="ignore")      # planted mistakes may overflow
    rho = np.zeros(A.shape)
    R = np.zeros(A.shape[1:])
    i1 = np.zeros(A.shape[1:])
    i2 = np.zeros(A.shape[1:])
    for k in range(nt - 1, -1, -1):                  # the pulse tail first (plasma.c:220)
        R = R + b[k]
        R = R + (1 - R) * a[k]
        u = -np.expm1(-R)                             # fraction ionised at least once
        i2 = i2 + c[k] * i1                           # doubly ionised, from the singly ionised after k
        i1 = u - i2
        rho[k] = i1 + 2 * i2                          # electrons per atom
    np.seterr(**err)
    return tau, sig, rho, rho * sig


def _oracle_ioniz(A, P, inp, q=O.OQ_PHYSICAL):
    N = A.size
    tau, F, rf = np.zeros(N), np.zeros(N), np.zeros(N)
    O.oracle().oracle_ioniz(O.vp(A), O.ip(O.dims_arr(A.shape)), P.w0, P.deltat, P.gas, P.rho_c, O.dp(inp.w),
                            P.rho_nt, P.n0, O.dp(tau), O.dp(F), O.dp(rf), q)
    return tau.reshape(A.shape), F.reshape(A.shape), rf.reshape(A.shape)


def _rel(a, b):
    """max |a - b| / max |b|; inf if either holds a non-finite value"""
    if not (np.isfinite(a).all() and np.isfinite(b).all()):
        return float("inf")
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("energy", [3e-3, 3e-2])
def test_ioniz_fraction_and_F_from_the_physics_formulas(energy):
    """weak (peak fraction 3e-2) and strong (0.89, double ionisation visible)
    ionisation"""
    inp, A = _intense(8, 64, energy=energy)
    P = inp.params
    tau_p, sig_p, rho_p, F_p = _physics_ioniz(A, P)
    tau, F, rf = _oracle_ioniz(A, P, inp)
    assert rho_p.max() > 1e-6                          # the check sees real ionisation ...
    assert (rho_p > 1e-12).sum() > 100                 # ... over many samples
    assert _rel(tau, tau_p) < 1e-13
    assert _rel(rf, rho_p) < 1e-13
    assert _rel(F, F_p) < 1e-13
    if energy > 1e-2:
        # the second species is seen by the pin: without it, or with its
        # coefficient scaled by dt instead of dt/T0, the fraction differs
        _, _, rho1, _ = _physics_ioniz(A, P, c_scale=0.0)
        _, _, rho2, _ = _physics_ioniz(A, P, c_scale=P.deltat)
        assert _rel(rf, rho1) > 1e-4
        assert _rel(rf, rho2) > 1e-4


@pytest.mark.parametrize("bit", ["Q11", "Q12", "Q5", "Q8"])
def test_ioniz_pin_catches_quirk_bits(bit):
    """each plasma quirk bit (avalanche units, quiver-velocity grid, n* of the
    first ionisation, recurrence order) moves the oracle far outside the pin
    (at strong ionisation: near 3e-3 the Q5 rate curve happens to cross the
    correct one at the field's peak)"""
    inp, A = _intense(8, 64, energy=3e-2)
    P = inp.params
    _, _, rho_p, F_p = _physics_ioniz(A, P)
    _, F, rf = _oracle_ioniz(A, P, inp, O.Q[bit])
    assert _rel(rf, rho_p) > 1e-6 or _rel(F, F_p) > 1e-6


def test_ioniz_pin_catches_misscaled_coefficients():
    """b scaled by dt instead of dt/T0, or the avalanche by dt/T0 instead of
    dt: such an oracle would disagree with the physics numbers (strong field:
    at 3e-3 the avalanche term, which seeds itself through (1 - R) a, makes
    almost all of the fraction)"""
    inp, A = _intense(8, 64, energy=3e-2)
    P = inp.params
    _, F, rf = _oracle_ioniz(A, P, inp)
    _, _, rho_b, _ = _physics_ioniz(A, P, b_scale=P.deltat)
    _, _, rho_a, _ = _physics_ioniz(A, P, dt_scale=P.deltat / T0AU)
    assert _rel(rf, rho_b) > 1e-3
    assert _rel(rf, rho_a) > 1e-3


def _gauss_case(nt=512, T=40e-15):
    """a real Gaussian a(t) = exp(-t^2 / (2 T^2)) on a 4x4 grid"""
    inp = make_inputs(configs.small(4, nt, plasm=0, sst_on=1))
    t = (np.arange(nt) - nt // 2) * inp.params.deltat
    a = np.exp(-t ** 2 / (2 * T ** 2))
    A = np.broadcast_to(a[:, None, None], (nt, 4, 4)).astype(np.complex128).copy()
    return inp, A, a, t


def _nonlinear(inp, A, dz, **over):
    P = O.params_from(inp)
    for k, v in over.items():
        setattr(P, k, v)
    B = A.copy()
    O.oracle().oracle_nonlinear(O.vp(B), C.byref(P), dz, None)
    return B


def test_self_steepening_step_closed_form():
    """SST on, plasma off, real Gaussian: S = 3 (gamma/w0) a^2 (t/T^2) is
    real, so A' = a exp(-dz S) exp(i dz gamma a^2): the amplitude is
    steepened towards the trailing edge (t > 0 loses, t < 0 gains)"""
    T = 40e-15
    inp, A, a, t = _gauss_case(T=T)
    g, w0, dz = 10.0, inp.params.w0, 1.0
    S = 3 * (g / w0) * a ** 2 * t / T ** 2
    ref = a * np.exp(-dz * S) * np.exp(1j * dz * g * a ** 2)
    assert np.abs(dz * S).max() > 0.1                 # the SST term is not negligible here
    B = _nonlinear(inp, A, dz, gamma2=g, sst_on=1, plasm=0)
    err = np.abs(B[:, 1, 2] - ref).max()
    assert err < 1e-10, err
    assert np.abs(B[t > T, 0, 0]).max() < a[t > T].max()  # trailing edge loses
    # the same check on an oracle whose S had the opposite sign (w0 -> -w0
    # flips gamma/w0 and nothing else with the plasma off) fails by far
    Bf = _nonlinear(inp, A, dz, gamma2=g, sst_on=1, plasm=0, w0=-w0)
    assert np.abs(Bf[:, 1, 2] - ref).max() > 1e-2


def test_plasma_phase_to_absorption_ratio_is_w0_tau():
    """Kerr and SST off: A' = A exp(-dz (1 + i w0 tau_c) rho_nt F / 2), so
    ln|A'/A| = -dz rho_nt F / 2 and arg(A'/A) / ln|A'/A| = w0 tau_c (Drude),
    with tau_c and F from the physics numbers above"""
    inp, A = _intense(8, 64, energy=3e-2)
    P = inp.params
    tau_p, _, _, F_p = _physics_ioniz(A, P)
    g = P.rho_nt * F_p / 2
    dz = 0.5 / g.max()                                 # absorption exponent up to 0.5
    B = _nonlinear(inp, A, dz, gamma2=0.0, sst_on=0, plasm=1)
    m = dz * g > 1e-6                                  # samples with visible plasma terms
    assert m.sum() > 50
    r = B[m] / A[m]
    lnmag, ph = np.log(np.abs(r)), np.angle(r)
    assert np.all(lnmag < 0)                           # absorption, never gain
    np.testing.assert_allclose(lnmag, -dz * g[m], rtol=1e-8)
    # phase = w0 tau_c x the absorption exponent; it wraps in the wings (tau_c
    # is long at low fields), so compare it through the residual phase factor
    ph_exp = -dz * P.w0 * tau_p[m] * g[m]
    resid = np.angle(r * np.exp(-1j * ph_exp))
    assert np.all(np.abs(resid) <= 1e-9 * (1 + np.abs(ph_exp)))
    # Q21 (Im part (1 + w0 tau_c) instead of w0 tau_c) leaves a residual of
    # -dz rho_nt F / 2, far outside
    Pq = O.params_from(inp, O.Q["Q21"])
    Pq.gamma2, Pq.sst_on = 0.0, 0
    Bq = A.copy()
    O.oracle().oracle_nonlinear(O.vp(Bq), C.byref(Pq), dz, None)
    rq = Bq[m] / A[m]
    resid_q = np.angle(rq * np.exp(-1j * ph_exp))
    assert np.abs(resid_q).max() > 0.1


def test_nonlinear_assembly_with_every_term():
    """Kerr + SST + plasma together: A' = A exp(dz A_nl) with
    A_nl = i gamma |A|^2 - S - (1 + i w0 tau_c) rho_nt F / 2, S from the pinned
    spectral derivative and tau_c, F from the physics numbers"""
    inp, A = _intense(8, 128, sst_on=1, plasm=1)
    P = inp.params
    dz = 1e-3
    S = np.zeros_like(A)
    O.oracle().oracle_sst(O.vp(A), O.vp(S), O.ip(O.dims_arr(A.shape)), O.dp(inp.w), P.gamma2, P.w0, 1)
    tau_p, _, _, F_p = _physics_ioniz(A, P)
    Anl = 1j * P.gamma2 * np.abs(A) ** 2 - S - (1 + 1j * P.w0 * tau_p) * P.rho_nt * F_p / 2
    ref = A * np.exp(dz * Anl)
    # every term moves the field by far more than the tolerance
    for term in (1j * P.gamma2 * np.abs(A) ** 2, S, (1 + 1j * P.w0 * tau_p) * P.rho_nt * F_p / 2):
        assert np.abs(dz * term).max() > 1e-6
    got = _nonlinear(inp, A, dz)
    scale = np.abs(A).max()
    assert np.abs(got - ref).max() < 1e-12 * scale
    # planted mistakes: +S, the plasma term's Q21 imaginary part, Q11 units
    wrong_s = A * np.exp(dz * (Anl + 2 * S))
    assert np.abs(got - wrong_s).max() > 1e-9 * scale
    for bit in ("Q21", "Q11"):
        Pq = O.params_from(inp, O.Q[bit])
        Bq = A.copy()
        O.oracle().oracle_nonlinear(O.vp(Bq), C.byref(Pq), dz, None)
        assert np.abs(Bq - ref).max() > 1e-9 * scale, bit


def test_mirror_aperture_and_loss_energy():
    """the end-of-pass mirror: |M|^2 = Ab_M inside the aperture radius and 0
    outside, so the energy after it is Ab_M x the energy inside the radius,
    whatever the curvature; with Q16 (no mirror loss) the factor is 1"""
    inp = make_inputs(configs.small(32, 64, mirror_R=0.3))
    A = inp.calibrated()
    r = inp.r
    ap = 1.2 * 0.2e-3                                   # 1.2 x the waist: clips ~6% of the energy
    inside = (r[None, :] ** 2 + r[:, None] ** 2) <= ap ** 2
    e_in = (np.abs(A) ** 2 * inside[None]).sum()
    assert 0 < e_in < (np.abs(A) ** 2).sum() * (1 - 1e-2)   # the aperture clips something
    for q, loss in ((O.OQ_PHYSICAL, 0.9), (O.Q["Q16"], 1.0)):
        P = O.params_from(inp, q)
        P.aperture, P.Ab_M = ap, 0.9
        B = A.copy()
        O.oracle().oracle_mirror(O.vp(B), C.byref(P))
        e_out = (np.abs(B) ** 2).sum()
        assert abs(e_out - loss * e_in) <= 1e-13 * e_in
        assert not B[:, ~inside].any()
        # inside, only the curvature phase: |B| = sqrt(loss) |A|
        np.testing.assert_allclose(np.abs(B[:, inside]), np.sqrt(loss) * np.abs(A[:, inside]), rtol=1e-14)
