"""CPU checks of the C-ABI library (no GPU compute): it loads, exports every
function include/*.h declares, and its host logic (input generator, param.bin
format, .bak checkpoint helpers, 1-D table helpers) behaves as specified."""
import ctypes as C
import os
import re
import struct
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from mpc_amd import _lib, configs, lib, make_inputs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for h in os.listdir(INCLUDE):
        if not h.endswith(".h"):
            continue
        src = open(os.path.join(INCLUDE, h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        src = re.sub(r"#[^\n]*", "", src)
        src = re.sub(r"typedef struct \{.*?\}\s*\w+;", "", src, flags=re.S)
        src = re.sub(r"\btypedef\b[^;]*;", "", src)
        for m in re.finditer(r"([A-Za-z_][A-Za-z0-9_]*)\s*\(([^;{}()]|\([^()]*\))*\)\s*;", src):
            names.add(m.group(1))
    return names


def exported_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True, check=True)
    return {line.split()[-1] for line in out.stdout.splitlines() if " T " in line}


def test_every_declared_function_is_exported():
    decl = declared_functions()
    # the reference entry points this library replaces (propagation.h:7-14, plasma.h:7-10, utils.h:7-10)
    legacy = {"calcexpR", "fftshift", "calc1oes", "ssfmprop", "linearstep", "createPlans", "setStep", "ionizpot",
              "ADK", "Natoms", "ioniz", "backup", "restore", "copyfile", "opener"}
    assert legacy <= decl
    missing = decl - exported_symbols()
    assert not missing, f"declared but not exported: {sorted(missing)}"


def test_struct_layouts_match_python():
    L = lib()
    assert L.mpc_sizeof_params() == C.sizeof(_lib.MpcParams)
    assert L.mpc_sizeof_input_spec() == C.sizeof(_lib.MpcInputSpec)
    assert C.sizeof(_lib.MpcParams) == O.oracle().oracle_sizeof_params()


def test_generator_deterministic_and_seeded():
    a = make_inputs(configs.small(16, 64, seed=11))
    b = make_inputs(configs.small(16, 64, seed=11))
    c = make_inputs(configs.small(16, 64, seed=12))
    assert np.array_equal(a.A, b.A)
    assert not np.array_equal(a.A, c.A)
    n = a.A - np.abs(a.A).max() * 0  # noise statistics: imaginary part is pure noise
    assert abs(np.std(n.imag) / 1e-6 - 1) < 0.05
    P = a.params
    dx = P.dr
    Esig = (np.abs(a.A) ** 2).sum() * 1.054560652926899e-34 * P.w0 * dx * dx * P.deltat
    assert abs(P.Esig / Esig - 1) < 1e-12
    # grids (ascending centred, FFT conventions)
    nt, ny = 64, 16
    assert np.allclose(a.w, (np.arange(nt) - nt // 2) * 2 * np.pi / (nt * P.deltat))
    assert np.allclose(a.wr, (np.arange(ny) - ny // 2) * 2 * np.pi / (ny * dx))
    assert np.allclose(a.betas, 40e-30 / 2 * a.w ** 2)


def test_cfg_parameters_match_description():
    inp = make_inputs(configs.cfg1_cpu_64())
    P = inp.params
    assert tuple(P.dims) == (64, 64, 512)
    assert abs(P.dr - 62.5e-6) < 1e-12 and abs(P.deltat - 3.90625e-15) < 1e-25
    assert abs(P.gamma2 / 1.1765e-35 - 1) < 1e-3          # n2 k0 hbar w0
    assert abs(P.rho_c / 1.0509e27 - 1) < 1e-3
    assert abs(P.dzz - 0.5 / 200) < 1e-15
    I = np.abs(inp.calibrated()) ** 2 * 1.054560652926899e-34 * P.w0
    assert 1.1e14 < I.max() < 1.3e14                        # ~1.2e10 W/cm^2


def test_param_bin_roundtrip(tmp_path):
    inp = make_inputs(configs.small(16, 64))
    path = str(tmp_path / "param.bin").encode()
    assert lib().mpc_write_param_bin(path, C.byref(inp.params)) == 0
    nt, ny = 64, 16
    size = 12 + 8 * 2 + 8 * nt + 8 * 3 + 8 * ny + 4 + 8 * nt + 8 * 2 + 4 + 8 * 6 + 8 * ny + 8 * 4 + 9
    raw = open(path, "rb").read()
    assert len(raw) == size                                  # MPC.c:77,180-203, packed
    assert struct.unpack_from("3i", raw, 0) == (16, 16, 64)
    q = _lib.MpcParams()
    assert lib().mpc_read_param_bin(path, C.byref(q)) == 0
    for f in ("dist", "dzz", "alpha", "Cav", "k0", "gamma2", "w0", "deltat", "rho_c", "rho_nt", "n0", "puls_e",
              "Ab_M", "dr", "def_err", "dzmin", "Esig", "sst_on", "plasm"):
        assert getattr(q, f) == getattr(inp.params, f), f
    assert q.gas == b"Argon"
    assert np.array_equal(np.ctypeslib.as_array(q.betas, (nt,)), inp.betas)
    assert np.array_equal(np.ctypeslib.as_array(q.r, (ny,)), inp.r)
    lib().mpc_free_params(C.byref(q))


def test_backup_restore_semantics(tmp_path):
    """utils.c:4-69: copyfile(path, 0) writes <path>.bak, restore copies it back"""
    f = tmp_path / "bp1.bin"
    f.write_bytes(b"abc")
    lib().copyfile(str(f).encode(), 0)
    assert (tmp_path / "bp1.bin.bak").read_bytes() == b"abc"
    f.write_bytes(b"abcdef")
    lib().restore(str(f).encode())
    assert f.read_bytes() == b"abc"


def test_host_table_helpers_match_oracle():
    n = 32
    wr = (np.arange(n) - n // 2) * 2 * np.pi / (n * 1e-4)
    dims = np.array([n, n, 16], dtype=np.int32)
    for z in (0.1, 0.6):
        a, b = np.zeros(n, complex), np.zeros(n, complex)
        sa, sb = C.c_int(1), C.c_int(1)
        lib().calcexpR(O.vp(a), C.byref(sa), z, 0.5, 6.1e6, 1e-3, O.dp(wr), O.ip(dims))
        O.oracle().oracle_calcexpR(O.vp(b), C.byref(sb), z, 0.5, 6.1e6, 1e-3, O.dp(wr), O.ip(dims), O.OQ_PHYSICAL)
        assert sa.value == sb.value and np.abs(a - b).max() < 1e-15
    r = (np.arange(n) - n // 2) * 1e-5
    rng = np.random.default_rng(0)
    for _ in range(5):
        BP = rng.random(n)
        v1, y1 = C.c_double(), np.zeros(4)
        lib().calc1oes(O.dp(BP), O.ip(dims), O.dp(r), C.byref(v1), O.dp(y1))
        v2, y2 = C.c_double(), np.zeros(4)
        O.oracle().oracle_calc1oes(O.dp(BP), O.ip(dims), O.dp(r), C.byref(v2), O.dp(y2), O.OQ_PHYSICAL)
        assert v1.value == v2.value and np.array_equal(y1, y2)


def test_fftshift_axes():
    rng = np.random.default_rng(1)
    A = rng.standard_normal((8, 4, 6)) + 1j * rng.standard_normal((8, 4, 6))
    dims = O.dims_arr(A.shape)
    buf = np.zeros_like(A)
    for axis, npax in ((1, 2), (2, 1), (3, 0)):
        B = A.copy()
        lib().fftshift(O.vp(B), O.ip(dims), O.vp(buf), axis)
        assert np.array_equal(B, np.fft.fftshift(A, axes=npax))
    t = A[:, 0, 0].copy()
    lib().fftshift(O.vp(t), O.ip(dims), O.vp(buf), 0)
    assert np.array_equal(t, np.fft.fftshift(A[:, 0, 0]))


def test_unknown_gas_and_unsupported_sizes_are_rejected():
    kon = np.zeros(6)
    assert lib().mpc_adk_constants(b"Unobtainium", O.dp(kon)) != 0
    assert lib().mpc_adk_constants(b"Argon", O.dp(kon)) == 0


def test_guard_verdict_on_injected_non_finite_values():
    """the solver guard's verdict (host logic): NaN / Inf energies or widths,
    energy gain beyond its tolerance, two-sided energy and width checks"""
    L = lib()
    nan, inf = float("nan"), float("inf")
    ENONFINITE, EGUARD = -7, -8
    v = L.mpc_guard_verdict
    assert v(1.0, 1.0, 1.0, 1.0, 1e-3, 1e-3, 1e-3) == 0
    assert v(1.0, nan, 1.0, 1.0, 1e-3, -1, -1) == ENONFINITE
    assert v(1.0, inf, 1.0, 1.0, -1, -1, -1) == ENONFINITE       # checked even with every tolerance off
    assert v(1.0, -inf, 1.0, 1.0, -1, -1, -1) == ENONFINITE
    assert v(1.0, 1.0, 1.0, nan, -1, -1, 0.05) == ENONFINITE
    assert v(1.0, 1.0, 1.0, nan, -1, -1, -1) == 0                 # width not checked
    # one-sided gain check (the solver): losses pass, gains beyond the tolerance fail
    assert v(1.0, 0.5, 0, 0, 1e-2, -1, -1) == 0
    assert v(1.0, 1.009, 0, 0, 1e-2, -1, -1) == 0
    assert v(1.0, 1.011, 0, 0, 1e-2, -1, -1) == EGUARD
    # two-sided checks (the surrogate's fallback)
    assert v(1.0, 0.97, 1.0, 1.0, -1, 0.02, 0.05) == EGUARD
    assert v(1.0, 1.01, 1.0, 1.06, -1, 0.02, 0.05) == EGUARD
    assert v(1.0, 1.01, 1.0, 1.04, -1, 0.02, 0.05) == 0
    assert v(0.0, 0.0, 1.0, 1.0, -1, 0.02, -1) == EGUARD            # 0/0 is not within any tolerance
    assert lib().mpc_strerror(ENONFINITE) and lib().mpc_strerror(EGUARD)
