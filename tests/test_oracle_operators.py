"""Known-answer tests that pin the oracle's operator restatement
(propagation.c) with closed forms and independent numpy computations."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from mpc_amd import configs, make_inputs


def _grid_inputs(nx=64, nt=64, **kw):
    return make_inputs(configs.small(nx, nt, **kw))


def test_exp_D0_closed_form():
    nt = 64
    betas = np.linspace(-3, 5, nt) ** 2
    out = np.zeros(nt, dtype=np.complex128)
    dz, alpha = 0.01, 0.3
    O.oracle().oracle_exp_D0(O.vp(out), O.dp(betas), alpha, dz, nt)
    ref = np.fft.ifftshift(np.exp(dz / 2 * (1j * betas - alpha / 2)))   # propagation.c:152,156
    assert np.abs(out - ref).max() < 1e-15


@pytest.mark.parametrize("z,expect", [(0.0, 1), (0.49, 1), (0.5, -1), (0.99, -1), (1.0, 1), (1.7, -1), (2.2, 1)])
def test_pass_sign_table(z, expect):
    assert O.oracle().oracle_pass_sign(z, 0.5) == expect


def test_calcexpR_literal_and_physical():
    n = 32
    wr = (np.arange(n) - n // 2) * 2 * np.pi / (n * 1e-4)
    dims = np.array([n, n, 16], dtype=np.int32)
    k0, dzh = 6.1e6, 1e-3
    L = O.oracle()
    # LITERAL: real exponential of the centred grid (propagation.c:386-388)
    out = np.zeros(n, dtype=np.complex128)
    sm = C.c_int(1)
    L.oracle_calcexpR(O.vp(out), C.byref(sm), 0.7, 0.5, k0, dzh, O.dp(wr), O.ip(dims), O.OQ_LITERAL)
    wc = (wr[n // 2 - 1] + wr[n // 2]) / 2
    ref = np.exp(-1 * dzh / (2 * k0) * (wr - wc) ** 2)      # s = -1 in the second pass
    assert sm.value == -1
    assert np.abs(out - ref).max() < 1e-15 * np.abs(ref).max()
    # PHYSICAL: phase with FFT-ordered kx (the intent at propagation.c:389)
    sm = C.c_int(1)
    L.oracle_calcexpR(O.vp(out), C.byref(sm), 0.2, 0.5, k0, dzh, O.dp(wr), O.ip(dims), O.OQ_PHYSICAL)
    kx = np.fft.ifftshift(wr)
    assert sm.value == 1
    assert np.abs(out - np.exp(1j * dzh / (2 * k0) * kx ** 2)).max() < 1e-14


def test_linearstep_equals_3d_fft_propagator():
    """linearstep == IFFT3(FFT3(A) * exp_D0[k] conj(expR[j]) conj(expR[i])) (exact composition)"""
    rng = np.random.default_rng(5)
    nt, n = 32, 16
    A = rng.standard_normal((nt, n, n)) + 1j * rng.standard_normal((nt, n, n))
    eD0 = np.exp(1j * rng.uniform(-3, 3, nt))
    eR = np.exp(1j * rng.uniform(-3, 3, n))
    dims = O.dims_arr(A.shape)
    got = A.copy()
    O.oracle().oracle_linearstep(O.vp(got), O.vp(eD0), O.vp(eR), O.ip(dims))
    D = eD0[:, None, None] * np.conj(eR)[None, :, None] * np.conj(eR)[None, None, :]
    ref = np.fft.ifftn(np.fft.fftn(A) * D)
    assert np.abs(got - ref).max() < 1e-13 * np.abs(ref).max()


def _linear_only(inp, dz, steps=1, z=0.1):
    """propagate with the oracle's linearstep only (no nonlinearity)"""
    P = inp.params
    dims = O.dims_arr(inp.A.shape)
    eD0 = np.zeros(dims[2], dtype=np.complex128)
    eR = np.zeros(dims[0], dtype=np.complex128)
    L = O.oracle()
    L.oracle_exp_D0(O.vp(eD0), O.dp(inp.betas), P.alpha, dz, int(dims[2]))
    sm = C.c_int(1)
    L.oracle_calcexpR(O.vp(eR), C.byref(sm), z, P.Cav, P.k0, dz / 2, O.dp(inp.wr), O.ip(dims), O.OQ_PHYSICAL)
    return eD0, eR


def test_gaussian_beam_width_and_gouy_phase():
    """diffraction of exp(-r^2/w0^2): |A(0,z)| = w0/w(z), phase -atan(z/zR)"""
    n, nt = 128, 32
    wl = configs.small(n, nt, waist=0.15e-3)
    wl.window_xy = 4e-3
    wl.noise_rel = 0.0
    wl.beta2_fs2_per_m = 0.0
    inp = make_inputs(wl)
    A = inp.A.copy()
    P = inp.params
    zR = P.k0 * wl.waist ** 2 / 2
    z_total = 0.6 * zR
    dz = z_total
    eD0, eR = _linear_only(inp, dz)
    dims = O.dims_arr(A.shape)
    O.oracle().oracle_linearstep(O.vp(A), O.vp(eD0), O.vp(eR), O.ip(dims))   # half step of dz/2 ...
    O.oracle().oracle_linearstep(O.vp(A), O.vp(eD0), O.vp(eR), O.ip(dims))   # ... twice = dz
    c = A[nt // 2, n // 2, n // 2] / inp.A[nt // 2, n // 2, n // 2]
    assert abs(abs(c) - 1 / np.sqrt(1 + (z_total / zR) ** 2)) < 1e-6
    assert abs(np.angle(c) + np.arctan(z_total / zR)) < 1e-6
    # second moment: <x^2> = w(z)^2 / 4
    I = np.abs(A[nt // 2]) ** 2
    x = inp.r
    wz = wl.waist * np.sqrt(1 + (z_total / zR) ** 2)
    mx2 = (I.sum(axis=0) * x ** 2).sum() / I.sum()
    assert abs(mx2 - wz ** 2 / 4) < 1e-6 * wz ** 2


def test_dispersion_chirp_broadening():
    """exp(-t^2/(2 T0^2)) under beta2: peak amplitude (1 + (z/LD)^2)^(-1/4)"""
    n, nt = 16, 1024
    wl = configs.small(n, nt)
    wl.window_t = 4e-12
    wl.noise_rel = 0.0
    wl.beta2_fs2_per_m = 2e5          # strong dispersion so LD is short
    inp = make_inputs(wl)
    t = (np.arange(nt) - nt // 2) * inp.params.deltat
    T0 = 60e-15
    A = np.broadcast_to(np.exp(-t ** 2 / (2 * T0 ** 2))[:, None, None], (nt, n, n)).astype(np.complex128).copy()
    beta2 = wl.beta2_fs2_per_m * 1e-30
    LD = T0 ** 2 / beta2
    z = 0.8 * LD
    eD0, eR = _linear_only(inp, z)
    eR[:] = 1.0                        # no diffraction
    dims = O.dims_arr(A.shape)
    O.oracle().oracle_linearstep(O.vp(A), O.vp(eD0), O.vp(eR), O.ip(dims))
    O.oracle().oracle_linearstep(O.vp(A), O.vp(eD0), O.vp(eR), O.ip(dims))
    peak = np.abs(A[:, 0, 0]).max()
    assert abs(peak - (1 + (z / LD) ** 2) ** -0.25) < 1e-9


def test_kerr_only_spm_and_literal_update():
    inp = _grid_inputs(16, 64, plasm=0, sst_on=0)
    P = O.params_from(inp)
    A = inp.calibrated()
    B = A.copy()
    dz = 0.01
    O.oracle().oracle_nonlinear(O.vp(B), C.byref(P), dz, None)
    ref = A * np.exp(1j * inp.params.gamma2 * np.abs(A) ** 2 * dz)
    assert np.abs(B - ref).max() < 1e-13 * np.abs(A).max()
    assert np.abs(np.abs(B) - np.abs(A)).max() < 1e-13 * np.abs(A).max()
    # LITERAL (propagation.c:325-326): for a real field x, x' = x cos(phi), y' = x' cos(phi) * 1
    P.quirks = O.Q["Q1"]
    R = np.full_like(A, 2.0 + 0j)
    phi = inp.params.gamma2 * 4.0 * dz
    O.oracle().oracle_nonlinear(O.vp(R), C.byref(P), dz, None)
    assert np.allclose(R.real, 2.0 * np.cos(phi), rtol=1e-14)
    assert np.allclose(R.imag, 2.0 * np.cos(phi) * np.cos(phi), rtol=1e-14)


def test_sst_analytic_derivative():
    """S = gamma2/w0 (2 dA A* + A dA*) with dA = -dA/dt (spectral) for a real Gaussian"""
    n, nt = 4, 512
    inp = _grid_inputs(n, nt)
    t = (np.arange(nt) - nt // 2) * inp.params.deltat
    T = 40e-15
    a = np.exp(-t ** 2 / (2 * T ** 2))
    A = np.broadcast_to(a[:, None, None], (nt, n, n)).astype(np.complex128).copy()
    S = np.zeros_like(A)
    g, w0 = 2.0, 3.0
    O.oracle().oracle_sst(O.vp(A), O.vp(S), O.ip(O.dims_arr(A.shape)), O.dp(inp.w), g, w0, 1)
    dA = (t / T ** 2) * a                       # -d/dt of the Gaussian
    ref = g / w0 * 3 * dA * a
    assert np.abs(S[:, 0, 0] - ref).max() < 1e-9 * np.abs(ref).max()
    S0 = np.ones_like(S)
    O.oracle().oracle_sst(O.vp(A), O.vp(S0), O.ip(O.dims_arr(A.shape)), O.dp(inp.w), g, w0, 0)
    assert not S0.any()


def test_energy_conserved_without_sst_and_plasma():
    inp = _grid_inputs(32, 128, plasm=0, sst_on=0, mirror_R=0.3, steps_per_pass=2, passes=3)
    A0 = inp.calibrated()
    A1, _ = O.run_fixed(inp, A0, 0, 5)
    e0, e1 = (np.abs(A0) ** 2).sum(), (np.abs(A1) ** 2).sum()
    assert abs(e1 / e0 - 1) < 1e-12
    assert O.rel_l2(A1, A0) > 1e-3          # the field did evolve


def test_parseval_temp_spec_and_beam_profiles():
    rng = np.random.default_rng(2)
    A = rng.standard_normal((64, 8, 8)) + 1j * rng.standard_normal((64, 8, 8))
    dims = O.dims_arr(A.shape)
    temp, spec = np.zeros(64), np.zeros(64)
    O.oracle().oracle_temp_spec(O.vp(A), O.ip(dims), O.dp(temp), O.dp(spec))
    assert np.allclose(temp, (np.abs(A) ** 2).sum(axis=(1, 2)), rtol=1e-14)
    assert np.allclose(spec, (np.abs(np.fft.fft(A, axis=0)) ** 2).sum(axis=(1, 2)), rtol=1e-12)
    assert abs(spec.sum() - 64 * temp.sum()) < 1e-11 * spec.sum()
    bp1, bp2 = np.zeros(8), np.zeros(8)
    O.oracle().oracle_beam_profiles(O.vp(A), O.ip(dims), O.dp(bp1), O.dp(bp2))
    assert np.allclose(bp1, (np.abs(A) ** 2).sum(axis=(0, 1)), rtol=1e-14)
    assert np.allclose(bp2, (np.abs(A) ** 2).sum(axis=(0, 2)), rtol=1e-14)


def _calc1oes(BP, r, q):
    dims = np.array([len(BP), len(BP), 1], dtype=np.int32)
    v, y = C.c_double(), np.zeros(4)
    O.oracle().oracle_calc1oes(O.dp(BP), O.ip(dims), O.dp(r), C.byref(v), O.dp(y), q)
    return v.value, y


def test_calc1oes_brute_force_and_tie_rule():
    n = 64
    r = (np.arange(n) - n // 2) * 1e-5
    BP = np.exp(-2 * r ** 2 / (1.7e-4) ** 2)
    v, y = _calc1oes(BP, r, O.OQ_PHYSICAL)
    nb = BP / BP.max()
    h = np.abs(nb - np.exp(-2))
    p1 = int(np.argmin(h[: n // 2]))
    p2 = n // 2 + int(np.argmin(h[n // 2:]))
    assert y[0] == r[p1] and y[1] == r[p2] and v == r[p2] - r[p1]
    assert y[2] == BP[p1] and y[3] == BP[p2]
    # tie: equal distances at k = 3 and 5 -> strict '<' keeps the first
    BP = np.full(n, 1.0)
    BP[3] = BP[5] = np.exp(-2) + 0.01
    v, y = _calc1oes(BP, r, O.OQ_PHYSICAL)
    assert y[0] == r[3]
    # Q15: argmin exactly at Nx/2 leaves posi2 at 0 in the reference
    BP = np.full(n, 1.0)
    BP[n // 2] = np.exp(-2)
    _, yl = _calc1oes(BP, r, O.Q["Q15"])
    _, yp = _calc1oes(BP, r, O.OQ_PHYSICAL)
    assert yl[1] == r[0] and yp[1] == r[n // 2]


@pytest.mark.parametrize("err,dz_out,bad", [
    (3e-3, 0.5, 1),                 # > 2 tol: halve and redo
    (2e-3, 1.0 * 2 ** (1 / 3), 0),  # == 2 tol exactly: falls through to "increase" (propagation.c:637-675)
    (1.5e-3, 2 ** (-1 / 3), 0),     # (tol, 2 tol): shrink, keep
    (1e-3, 1.0, 0),                 # (tol/2, tol]: keep
    (0.7e-3, 1.0, 0),
    (1e-4, 2 ** (1 / 3), 0),        # small: grow
])
def test_setstep_decision_table(err, dz_out, bad):
    dz, b = C.c_double(1.0), C.c_int(0)
    y = np.array([0.0, 1.0, 0, 0])
    O.oracle().oracle_setstep_decide(C.byref(dz), err, 1e-3, 1e-6, O.dp(y), O.dp(y), C.byref(b))
    assert b.value == bad and abs(dz.value - dz_out) < 1e-15


def test_setstep_limits():
    y = np.array([0.0, 1e-3, 0, 0])
    dz, b = C.c_double(1.0), C.c_int(0)
    O.oracle().oracle_setstep_decide(C.byref(dz), 1e-6, 1e-3, 1e-6, O.dp(y), O.dp(y), C.byref(b))
    assert dz.value == 1.0                     # growth capped at 100 * beam size
    dz = C.c_double(1e-6)
    O.oracle().oracle_setstep_decide(C.byref(dz), 1.0, 1e-3, 1e-6, O.dp(y), O.dp(y), C.byref(b))
    assert b.value == 0 and dz.value == 1e-6   # too small to halve: accepted as is


def test_richardson_and_calibration():
    rng = np.random.default_rng(0)
    c1 = rng.standard_normal(100) + 1j * rng.standard_normal(100)
    c2 = rng.standard_normal(100) + 1j * rng.standard_normal(100)
    a = np.zeros_like(c1)
    O.oracle().oracle_richardson(O.vp(a), O.vp(c1), O.vp(c2), 100)
    assert np.allclose(a, 4 / 3 * c2 - 1 / 3 * c1, rtol=1e-15)
    b = c1.copy()
    O.oracle().oracle_calibrate(O.vp(b), 100, 4.0, 1.0, O.Q["Q9"])
    assert np.array_equal(b.real, 2 * c1.real) and np.array_equal(b.imag, c1.imag)
    b = c1.copy()
    O.oracle().oracle_calibrate(O.vp(b), 100, 4.0, 1.0, O.OQ_PHYSICAL)
    assert np.array_equal(b, 2 * c1)


def test_mirror_thin_lens_q_parameter():
    """collimated Gaussian x mirror phase exp(-i k0 r^2/R), then free space:
    the width follows the thin-lens q-parameter law with f = R/2."""
    n, nt = 128, 16
    wl = configs.small(n, nt, waist=0.5e-3, mirror_R=1.0, plasm=0, sst_on=0)
    wl.window_xy = 6e-3
    wl.noise_rel = 0.0
    wl.beta2_fs2_per_m = 0.0
    inp = make_inputs(wl)
    P = O.params_from(inp)
    A = inp.A.copy()
    O.oracle().oracle_mirror(O.vp(A), C.byref(P))
    k0, w0, f = inp.params.k0, wl.waist, wl.mirror_R / 2
    zR = k0 * w0 ** 2 / 2
    z = 0.3
    eD0, eR = _linear_only(inp, z)
    dims = O.dims_arr(A.shape)
    O.oracle().oracle_linearstep(O.vp(A), O.vp(eD0), O.vp(eR), O.ip(dims))
    O.oracle().oracle_linearstep(O.vp(A), O.vp(eD0), O.vp(eR), O.ip(dims))
    # field ~ exp(+i k r^2 / (2 q)), q -> q + z under this propagator, waist
    # at the mirror q0 = -i zR, converging thin lens 1/q1 = 1/q0 - 1/f
    q0 = -1j * zR
    q1 = 1 / (1 / q0 - 1 / f)
    q = q1 + z
    wz2 = 2 / (k0 * np.imag(1 / q))
    I = np.abs(A[nt // 2]) ** 2     # the pulse centre (the window edges hold only roundoff)
    mx2 = (I.sum(axis=0) * inp.r ** 2).sum() / I.sum()
    assert abs(mx2 - wz2 / 4) < 1e-5 * wz2
