"""ctypes loader for the FP64 CPU oracle (oracle/liboracle.so).

Test infrastructure: used by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker / CPU baseline only.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")

OQ_PHYSICAL = 0
OQ_LITERAL = 0x1FFFF
Q = {
    "Q1": 1 << 0, "Q2": 1 << 1, "Q3": 1 << 2, "Q4": 1 << 3, "Q5": 1 << 4, "Q6": 1 << 5, "Q7": 1 << 6,
    "Q8": 1 << 7, "Q9": 1 << 8, "Q10": 1 << 9, "Q11": 1 << 10, "Q12": 1 << 11, "Q15": 1 << 12,
    "Q16": 1 << 13, "Q17": 1 << 14, "Q21": 1 << 15, "Q22": 1 << 16,
}

dptr = C.POINTER(C.c_double)
iptr = C.POINTER(C.c_int)


class OracleParams(C.Structure):
    _fields_ = [
        ("dims", C.c_int * 3),
        ("sst_on", C.c_int), ("plasm", C.c_int),
        ("steps_per_pass", C.c_int), ("npasses", C.c_int),
        ("dist", C.c_double), ("dzz", C.c_double), ("alpha", C.c_double), ("Cav", C.c_double),
        ("k0", C.c_double), ("gamma2", C.c_double), ("w0", C.c_double), ("deltat", C.c_double),
        ("rho_c", C.c_double), ("rho_nt", C.c_double), ("n0", C.c_double), ("puls_e", C.c_double),
        ("Ab_M", C.c_double), ("dr", C.c_double), ("def_err", C.c_double), ("dzmin", C.c_double),
        ("Esig", C.c_double),
        ("mirror_R", C.c_double), ("aperture", C.c_double),
        ("betas", dptr), ("wr", dptr), ("w", dptr), ("r", dptr),
        ("gas", C.c_char * 16),
        ("quirks", C.c_uint), ("pad_", C.c_int),
    ]


_o = None


def oracle() -> C.CDLL:
    global _o
    if _o is not None:
        return _o
    if not os.path.exists(ORACLE_SO):
        raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
    L = C.CDLL(ORACLE_SO)
    vp = C.c_void_p
    P = C.POINTER(OracleParams)
    sig = {
        "oracle_sizeof_params": (C.c_int, []),
        "oracle_fft_lines": (None, [vp, C.c_int, C.c_long, C.c_long, C.c_long, C.c_long, C.c_long, C.c_int]),
        "oracle_fft_axis": (None, [vp, iptr, C.c_int, C.c_int]),
        "oracle_exp_D0": (None, [vp, dptr, C.c_double, C.c_double, C.c_int]),
        "oracle_calcexpR": (None, [vp, iptr, C.c_double, C.c_double, C.c_double, C.c_double, dptr, iptr, C.c_uint]),
        "oracle_pass_sign": (C.c_int, [C.c_double, C.c_double]),
        "oracle_linearstep": (None, [vp, vp, vp, iptr]),
        "oracle_sst": (None, [vp, vp, iptr, dptr, C.c_double, C.c_double, C.c_int]),
        "oracle_ionizpot": (None, [C.c_char_p, dptr, dptr]),
        "oracle_natoms": (None, [C.c_char_p, dptr, dptr]),
        "oracle_adk_consts": (None, [C.c_char_p, C.c_uint, dptr]),
        "oracle_adk": (None, [dptr, dptr, C.c_long, dptr, dptr]),
        "oracle_ioniz": (None, [vp, iptr, C.c_double, C.c_double, C.c_char_p, C.c_double, dptr, C.c_double,
                                C.c_double, dptr, dptr, dptr, C.c_uint]),
        "oracle_density_line": (None, [dptr, dptr, dptr, dptr, C.c_int, dptr, dptr, C.c_uint]),
        "oracle_ssfmprop": (None, [vp, P, C.c_double, iptr, C.c_double, dptr, dptr, dptr, dptr, dptr, dptr, dptr,
                                   C.c_int]),
        "oracle_nonlinear": (None, [vp, P, C.c_double, dptr]),
        "oracle_mirror": (None, [vp, P]),
        "oracle_calc1oes": (None, [dptr, iptr, dptr, dptr, dptr, C.c_uint]),
        "oracle_beam_profiles": (None, [vp, iptr, dptr, dptr]),
        "oracle_temp_spec": (None, [vp, iptr, dptr, dptr]),
        "oracle_setstep": (None, [iptr, dptr, vp, vp, C.c_double, C.c_double, dptr, dptr, iptr, dptr]),
        "oracle_setstep_decide": (None, [dptr, C.c_double, C.c_double, C.c_double, dptr, dptr, iptr]),
        "oracle_richardson": (None, [vp, vp, vp, C.c_long]),
        "oracle_calibrate": (None, [vp, C.c_long, C.c_double, C.c_double, C.c_uint]),
        "oracle_run_fixed": (None, [vp, P, C.c_int, C.c_int, dptr]),
        "oracle_run_adaptive": (C.c_int, [vp, P, dptr, dptr, iptr, C.c_int, dptr, dptr]),
        "oracle_set_threads": (None, [C.c_int]),
        "oracle_last_rho_peak": (C.c_double, []),
        "oracle_get_threads": (C.c_int, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    assert L.oracle_sizeof_params() == C.sizeof(OracleParams)
    _o = L
    return L


def vp(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def dp(a: np.ndarray):
    return a.ctypes.data_as(dptr)


def dims_arr(shape) -> np.ndarray:
    """(Nt, Ny, Nx) numpy shape -> int32 dims {Nx, Ny, Nt}."""
    nt, ny, nx = shape
    return np.array([nx, ny, nt], dtype=np.int32)


def ip(a: np.ndarray):
    return a.ctypes.data_as(iptr)


def params_from(inp, quirks: int = OQ_PHYSICAL) -> OracleParams:
    """Oracle parameters from an mpc_amd Inputs (same field layout)."""
    src = inp.params
    p = OracleParams()
    for name, _t in OracleParams._fields_:
        if name in ("betas", "wr", "w", "r"):
            continue
        setattr(p, name, getattr(src, name))
    p.betas, p.wr, p.w, p.r = dp(inp.betas), dp(inp.wr), dp(inp.w), dp(inp.r)
    p.quirks = quirks
    return p


def run_fixed(inp, A: np.ndarray, step0: int, nsteps: int, quirks: int = OQ_PHYSICAL):
    """Oracle fixed-step run on a copy of A (reference layout); returns (A, rho_peak)."""
    L = oracle()
    B = np.ascontiguousarray(A, dtype=np.complex128).copy()
    p = params_from(inp, quirks)
    rho = C.c_double(0.0)
    L.oracle_run_fixed(vp(B), C.byref(p), step0, nsteps, C.byref(rho))
    return B, rho.value


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def observables(inp, A: np.ndarray) -> dict:
    """energy (sum of temp), rms spectral width over the Omega grid, temp and
    spec of a reference-layout field, from the oracle's temp_spec"""
    nt = A.shape[0]
    temp, spec = np.zeros(nt), np.zeros(nt)
    oracle().oracle_temp_spec(vp(A), ip(dims_arr(A.shape)), dp(temp), dp(spec))
    om = np.fft.ifftshift(inp.w)                 # the FFT-order Omega of each spectral bin
    s0, s1, s2 = spec.sum(), (om * spec).sum(), (om * om * spec).sum()
    width = float(np.sqrt(max(0.0, s2 / s0 - (s1 / s0) ** 2)))
    return {"energy": float(temp.sum()), "width": width, "temp": temp, "spec": spec}
