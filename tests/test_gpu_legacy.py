"""GPU parity of the legacy C entry points (propagation.h / plasma.h), called
through the C ABI exactly as the reference driver calls them, against the
FP64 oracle's restatement of the same functions."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from mpc_amd import configs, lib, make_inputs

pytestmark = pytest.mark.gpu


def _plans(dims):
    L = lib()
    hs = [C.c_void_p() for _ in range(6)]
    L.createPlans(0, O.ip(dims), None, *[C.byref(h) for h in hs])
    return hs


def _free(hs):
    for h in hs:
        lib().mpc_destroy_plan(h)


def test_linearstep_matches_oracle(gpu_lib):
    print("case legacy linearstep 32x32x512", flush=True)
    inp = make_inputs(configs.small(32, 512))
    A = inp.calibrated()
    dims = O.dims_arr(A.shape)
    o = O.oracle()
    eD0 = np.zeros(dims[2], dtype=np.complex128)
    eR = np.zeros(dims[0], dtype=np.complex128)
    o.oracle_exp_D0(O.vp(eD0), O.dp(inp.betas), 0.0, inp.params.dzz, int(dims[2]))
    sm = C.c_int(1)
    o.oracle_calcexpR(O.vp(eR), C.byref(sm), 0.1, inp.params.Cav, inp.params.k0, inp.params.dzz / 2,
                      O.dp(inp.wr), O.ip(dims), O.OQ_PHYSICAL)
    ref = A.copy()
    o.oracle_linearstep(O.vp(ref), O.vp(eD0), O.vp(eR), O.ip(dims))
    hs = _plans(dims)
    got = A.copy()
    lib().linearstep(O.vp(got), O.vp(eD0), O.vp(eR), hs[4], hs[2], hs[0], O.ip(dims))
    _free(hs)
    assert lib().mpc_last_error() == 0
    err = O.rel_l2(got, ref)
    print(f"  rel_l2={err:.3e}", flush=True)
    assert err < 1e-13


@pytest.mark.parametrize("change", [0, 1])
def test_ssfmprop_matches_oracle(gpu_lib, change):
    print(f"case legacy ssfmprop change={change}", flush=True)
    inp = make_inputs(configs.small(32, 512, mirror_R=0.0))
    A = inp.calibrated()
    P = inp.params
    dims = O.dims_arr(A.shape)
    z = 1.3 * P.Cav   # second pass: diffraction sign -1
    o = O.oracle()
    ref = A.copy()
    po = O.params_from(inp)
    sm_o, pe_o = C.c_int(1), C.c_double(P.puls_e)
    bo = [np.zeros(32), np.zeros(32), np.zeros(4), np.zeros(4)]
    b1o, b2o = C.c_double(0), C.c_double(0)
    o.oracle_ssfmprop(O.vp(ref), C.byref(po), P.dzz, C.byref(sm_o), z, C.byref(pe_o), *[O.dp(b) for b in bo],
                      C.byref(b1o), C.byref(b2o), change)
    hs = _plans(dims)
    got = A.copy()
    sm, pe = C.c_int(1), C.c_double(P.puls_e)
    bg = [np.zeros(32), np.zeros(32), np.zeros(4), np.zeros(4)]
    b1, b2 = C.c_double(0), C.c_double(0)
    lib().ssfmprop(O.vp(got), O.ip(dims), P.sst_on, P.dzz, O.dp(inp.betas), P.alpha, C.byref(sm), z, P.Cav, P.k0,
                   O.dp(inp.wr), *hs, O.dp(inp.w), P.gamma2, P.w0, P.plasm, P.deltat, P.gas, P.rho_c, P.rho_nt, P.n0,
                   C.byref(pe), O.dp(inp.r), *[O.dp(b) for b in bg], C.byref(b1), C.byref(b2), P.Ab_M, change, 0)
    _free(hs)
    assert lib().mpc_last_error() == 0
    err = O.rel_l2(got, ref)
    print(f"  rel_l2={err:.3e} signmem={sm.value}/{sm_o.value}", flush=True)
    assert err < 1e-10
    assert sm.value == sm_o.value == -1
    assert pe.value == pe_o.value
    if change:
        for g, r in zip(bg, bo):
            np.testing.assert_allclose(g, r, rtol=1e-10, atol=0)
        assert b1.value == b1o.value and b2.value == b2o.value


def test_ioniz_matches_oracle(gpu_lib):
    print("case legacy ioniz 16x16x1024", flush=True)
    inp = make_inputs(configs.small(16, 1024, energy=3e-3))
    A = inp.calibrated()
    P = inp.params
    dims = O.dims_arr(A.shape)
    N = A.size
    tau_o, F_o, rf = np.zeros(N), np.zeros(N), np.zeros(N)
    O.oracle().oracle_ioniz(O.vp(A), O.ip(dims), P.w0, P.deltat, P.gas, P.rho_c, O.dp(inp.w), P.rho_nt, P.n0,
                            O.dp(tau_o), O.dp(F_o), O.dp(rf), O.OQ_PHYSICAL)
    assert F_o.max() > 0
    tau, F = np.zeros(N), np.zeros(N)
    lib().ioniz(O.vp(A), O.ip(dims), P.w0, P.deltat, P.gas, P.rho_c, O.dp(inp.w), P.rho_nt, P.n0, O.dp(tau), O.dp(F))
    assert lib().mpc_last_error() == 0
    np.testing.assert_allclose(tau, tau_o, rtol=1e-13)
    err = np.abs(F - F_o).max() / np.abs(F_o).max()
    print(f"  F max rel err={err:.3e} peak rho frac={rf.max():.3e}", flush=True)
    assert err < 1e-12


def test_adk_matches_oracle(gpu_lib):
    print("case legacy ADK", flush=True)
    kon = np.zeros(6)
    lib().mpc_adk_constants(b"Argon", O.dp(kon))
    kon_o = np.zeros(6)
    O.oracle().oracle_adk_consts(b"Argon", O.OQ_PHYSICAL, O.dp(kon_o))
    np.testing.assert_allclose(kon, kon_o, rtol=1e-15)
    E = np.geomspace(1e-4, 0.2, 4096)
    dims = np.array([16, 16, 16], dtype=np.int32)
    W1, W2 = np.zeros(E.size), np.zeros(E.size)
    # reconstruct the reference's ADK inputs from the constants
    I1, I2 = C.c_double(), C.c_double()
    lib().ionizpot(b"Argon", C.byref(I1), C.byref(I2))
    Ei = np.array([I1.value / 27.21, I2.value / 27.21])
    nq = np.array([(2 * Ei[0]) ** -0.5, 2 * (2 * Ei[1]) ** -0.5])
    from math import gamma, sqrt
    cnl = np.array([sqrt(2 ** (2 * n) / (n * gamma(n + 1) * gamma(n))) for n in nq])
    lib().ADK(O.dp(W1), O.dp(W2), O.ip(dims), O.dp(cnl), O.dp(np.ones(2)), O.dp(nq), O.dp(np.zeros(2)), O.dp(Ei), O.dp(E))
    W1o, W2o = np.zeros(E.size), np.zeros(E.size)
    O.oracle().oracle_adk(O.dp(W1o), O.dp(W2o), E.size, O.dp(kon_o), O.dp(E))
    np.testing.assert_allclose(W1, W1o, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(W2, W2o, rtol=1e-12, atol=1e-300)


def test_setstep_matches_oracle(gpu_lib):
    print("case legacy setStep", flush=True)
    rng = np.random.default_rng(3)
    shape = (64, 16, 16)
    C2 = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    C1 = C2 + 1e-3 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    dims = O.dims_arr(shape)
    y1, y2 = np.array([0, 1.0, 0, 0]), np.array([0, 1.0, 0, 0])
    for de in (1e-4, 1e-3, 2e-3, 1e-2):
        dz, bad, err = C.c_double(1e-3), C.c_int(0), C.c_double(0)
        lib().setStep(O.ip(dims), C.byref(dz), O.vp(C1), O.vp(C2), de, 1e-6, O.dp(y1), O.dp(y2), C.byref(bad), C.byref(err))
        dzo, bado, erro = C.c_double(1e-3), C.c_int(0), C.c_double(0)
        O.oracle().oracle_setstep(O.ip(dims), C.byref(dzo), O.vp(C1), O.vp(C2), de, 1e-6, O.dp(y1), O.dp(y2),
                                  C.byref(bado), C.byref(erro))
        assert abs(err.value - erro.value) < 1e-13 * erro.value
        assert bad.value == bado.value and dz.value == dzo.value


def test_create_plans_writes_and_reads_the_variant_cache(gpu_lib, tmp_path):
    """createPlans' cache file plays FFTW wisdom's role (propagation.c:540-611):
    with no file the first ssfmprop tunes the kernel variants and writes
    "<cpath><Nx>"; a later createPlans(readin=1) reads it and the step runs
    those variants without tuning again (bit-identical to a context set to
    them); an unreadable file is reported and replaced"""
    from mpc_amd import Propagator
    print("case legacy createPlans cache 32x32x512", flush=True)
    inp = make_inputs(configs.small(32, 512, mirror_R=0.0))
    A = inp.calibrated()
    P = inp.params
    dims = O.dims_arr(A.shape)
    z = 0.3 * P.Cav
    cpath = str(tmp_path / "wisdom_")
    path = cpath + "32"
    L = lib()

    def step_with_cache():
        hs = [C.c_void_p() for _ in range(6)]
        L.createPlans(1, O.ip(dims), cpath.encode(), *[C.byref(h) for h in hs])
        got = A.copy()
        sm, pe = C.c_int(1), C.c_double(P.puls_e)
        bg = [np.zeros(32), np.zeros(32), np.zeros(4), np.zeros(4)]
        b1, b2 = C.c_double(0), C.c_double(0)
        L.ssfmprop(O.vp(got), O.ip(dims), P.sst_on, P.dzz, O.dp(inp.betas), P.alpha, C.byref(sm), z, P.Cav, P.k0,
                   O.dp(inp.wr), *hs, O.dp(inp.w), P.gamma2, P.w0, P.plasm, P.deltat, P.gas, P.rho_c, P.rho_nt,
                   P.n0, C.byref(pe), O.dp(inp.r), *[O.dp(b) for b in bg], C.byref(b1), C.byref(b2), P.Ab_M, 0, 0)
        _free(hs)
        assert L.mpc_last_error() == 0
        return got

    import os
    import re
    assert not os.path.exists(path)
    first = step_with_cache()                         # tunes, then writes the cache
    text = open(path).read()
    m = re.match(r"mpc-kernel-variants tline=(\d+) axis=(\d+) yaxis=(\d+) nx=32 ny=32 nt=512", text)
    assert m, text
    tv, xv, yv = map(int, m.groups())
    assert 0 <= tv < 2 and 0 <= xv < 32 and 0 <= yv < 32
    # a cache with other variants is read and used as it is
    open(path, "w").write("mpc-kernel-variants tline=1 axis=2 yaxis=27 nx=32 ny=32 nt=512\n")
    cached = step_with_cache()
    assert open(path).read().startswith("mpc-kernel-variants tline=1 axis=2 yaxis=27")   # not re-tuned
    with Propagator(inp, "f64") as p:
        p.upload(A)
        p.set_variants(1, 2, 27)
        p.step(P.dzz, z, 1)
        same = p.download()
    assert np.array_equal(cached, same)
    ref = A.copy()
    po = O.params_from(inp)
    sm_o, pe_o = C.c_int(1), C.c_double(P.puls_e)
    bo = [np.zeros(32), np.zeros(32), np.zeros(4), np.zeros(4)]
    O.oracle().oracle_ssfmprop(O.vp(ref), C.byref(po), P.dzz, C.byref(sm_o), z, C.byref(pe_o),
                               *[O.dp(b) for b in bo], C.byref(C.c_double()), C.byref(C.c_double()), 0)
    assert O.rel_l2(first, ref) < 1e-10 and O.rel_l2(cached, ref) < 1e-10
    # an unreadable cache is replaced by a tuned one
    open(path, "w").write("not a cache\n")
    step_with_cache()
    assert open(path).read().startswith("mpc-kernel-variants tline=")
