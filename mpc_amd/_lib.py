"""ctypes binding of libmpc_hip.so (the C ABI declared in include/*.h).

Loading fails loudly if the library has not been built: there is no CPU
fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmpc_hip.so")

MPC_F64 = 0
MPC_F32 = 1

dptr = C.POINTER(C.c_double)
iptr = C.POINTER(C.c_int)


class MpcParams(C.Structure):
    """mpc_params (include/mpc.h): param.bin content + runtime options."""
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


class MpcInputSpec(C.Structure):
    """mpc_input_spec (include/mpc.h): synthetic workload description."""
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nt", C.c_int),
        ("passes", C.c_int), ("steps_per_pass", C.c_int),
        ("sst_on", C.c_int), ("plasm", C.c_int), ("pad_", C.c_int),
        ("window_xy", C.c_double), ("window_t", C.c_double), ("lambda0", C.c_double),
        ("waist", C.c_double), ("tau_fwhm", C.c_double), ("energy", C.c_double),
        ("noise_rel", C.c_double), ("n2_per_bar", C.c_double), ("beta2", C.c_double),
        ("pressure", C.c_double), ("rho_nt_per_bar", C.c_double), ("cell_L", C.c_double),
        ("mirror_R", C.c_double), ("aperture", C.c_double), ("Ab_M", C.c_double),
        ("def_err", C.c_double), ("dzmin", C.c_double),
        ("seed", C.c_ulonglong),
        ("gas", C.c_char * 16),
    ]


class MpcDiag(C.Structure):
    _fields_ = [
        ("energy_sum", C.c_double), ("rho_peak", C.c_double), ("spec_rms_width", C.c_double),
        ("temp", dptr), ("spec", dptr), ("BP1", dptr), ("BP2", dptr),
    ]


_lib = None


def lib() -> C.CDLL:
    """Load libmpc_hip.so (raises if it is missing: no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run `python -m mpc_amd.build` first")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    sig = {
        "mpc_strerror": (C.c_char_p, [C.c_int]),
        "mpc_last_error": (C.c_int, []),
        "mpc_sizeof_params": (C.c_int, []),
        "mpc_sizeof_input_spec": (C.c_int, []),
        "mpc_device_count": (C.c_int, []),
        "mpc_create": (C.c_int, [C.POINTER(MpcParams), C.c_int, C.POINTER(vp)]),
        "mpc_destroy": (None, [vp]),
        "mpc_upload": (C.c_int, [vp, vp]),
        "mpc_download": (C.c_int, [vp, vp]),
        "mpc_snapshot_begin": (C.c_int, [vp]),
        "mpc_guard_verdict": (C.c_int, [C.c_double] * 7),
        "mpc_set_guard": (C.c_int, [vp, C.c_double]),
        "mpc_energy": (C.c_int, [vp, dptr]),
        "mpc_snapshot_poll": (C.c_int, [vp]),
        "mpc_snapshot_wait": (C.c_int, [vp, C.POINTER(C.c_void_p)]),
        "mpc_calibrate": (C.c_int, [vp]),
        "mpc_run": (C.c_int, [vp, C.c_int, C.c_int]),
        "mpc_step": (C.c_int, [vp, C.c_double, C.c_double, iptr]),
        "mpc_mirror": (C.c_int, [vp]),
        "mpc_linear": (C.c_int, [vp, C.c_double, C.c_int, C.c_int]),
        "mpc_diagnostics": (C.c_int, [vp, C.POINTER(MpcDiag)]),
        "mpc_run_adaptive": (C.c_int, [vp, dptr, dptr, iptr, C.c_int, iptr, dptr, dptr]),
        "mpc_setstep_noalloc": (C.c_int, [iptr, dptr, vp, vp, C.c_double, C.c_double, dptr, dptr, iptr, dptr]),
        "mpc_set_timing": (C.c_int, [vp, C.c_int]),
        "mpc_kernel_ms": (C.c_double, [vp, C.c_int]),
        "mpc_field_bytes": (C.c_longlong, [vp]),
        "mpc_autotune": (C.c_int, [vp]),
        "mpc_get_variants": (C.c_int, [vp, iptr, iptr]),
        "mpc_set_variants": (C.c_int, [vp, C.c_int, C.c_int]),
        "mpc_get_variants_xy": (C.c_int, [vp, iptr, iptr, iptr]),
        "mpc_set_variants_xy": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
        "mpc_variant_counts": (C.c_int, [iptr, iptr]),
        "mpc_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p)]),
        "mpc_batch_destroy": (None, [vp]),
        "mpc_batch_run": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
        "mpc_batch_graphs": (C.c_int, [vp]),
        "mpc_batch_autotune": (C.c_int, [vp]),
        "mpc_sizeof_gemm": (C.c_int, []),
        "mpc_gemm_bf16": (C.c_int, [vp, vp]),
        "mpc_fno_combine": (C.c_int, [vp, vp, C.c_long, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, vp]),
        "mpc_fno_bias_act": (C.c_int, [vp, vp, vp, C.c_long, C.c_int, C.c_long, C.c_int, vp]),
        "mpc_fno_act_bwd": (C.c_int, [vp, vp, vp, C.c_long, C.c_int, vp]),
        "mpc_sizeof_pw": (C.c_int, []),
        "mpc_fno_pointwise": (C.c_int, [vp, vp]),
        "mpc_fno_transpose": (C.c_int, [vp, vp, C.c_int, C.c_int, vp]),
        "mpc_fno_mix": (C.c_int, [vp, vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, vp]),
        "mpc_fno_mix_bwd": (C.c_int, [vp, vp, vp, vp, vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, vp]),
        "mpc_fno_bias_grad": (C.c_int, [vp, vp, C.c_long, C.c_int, C.c_long, C.c_int, vp]),
        "mpc_fno_adam": (C.c_int, [vp, vp, vp, vp, C.c_long, C.c_float, C.c_float, C.c_float, C.c_float,
                                   C.c_float, vp]),
        "mpc_fno_rel_l2": (C.c_int, [vp, vp, vp, vp, C.c_int, C.c_long, vp]),
        "mpc_dev_alloc": (C.c_int, [C.c_long, C.POINTER(C.c_void_p)]),
        "mpc_dev_free": (None, [vp]),
        "mpc_dev_upload": (C.c_int, [vp, vp, C.c_long]),
        "mpc_dev_download": (C.c_int, [vp, vp, C.c_long]),
        "mpc_dev_copy": (C.c_int, [vp, vp, C.c_long, vp]),
        "mpc_dev_fill": (C.c_int, [vp, C.c_long, C.c_float, vp]),
        "mpc_dev_sync": (C.c_int, []),
        "mpc_export_planar": (C.c_int, [vp, vp, C.c_double]),
        "mpc_import_planar": (C.c_int, [vp, vp, C.c_double]),
        "mpc_mkinput": (C.c_int, [C.POINTER(MpcInputSpec), C.POINTER(MpcParams), dptr, dptr, dptr, dptr, vp]),
        "mpc_write_param_bin": (C.c_int, [C.c_char_p, C.POINTER(MpcParams)]),
        "mpc_read_param_bin": (C.c_int, [C.c_char_p, C.POINTER(MpcParams)]),
        "mpc_free_params": (None, [C.POINTER(MpcParams)]),
    }
    # legacy entry points (propagation.h, plasma.h, utils.h)
    plan = C.c_void_p
    sig.update({
        "createPlans": (None, [C.c_int, iptr, C.c_char_p] + [C.POINTER(plan)] * 6),
        "mpc_destroy_plan": (None, [plan]),
        "ssfmprop": (None, [vp, iptr, C.c_int, C.c_double, dptr, C.c_double, iptr, C.c_double, C.c_double,
                            C.c_double, dptr, plan, plan, plan, plan, plan, plan, dptr, C.c_double, C.c_double,
                            C.c_int, C.c_double, C.c_char_p, C.c_double, C.c_double, C.c_double, dptr, dptr,
                            dptr, dptr, dptr, dptr, dptr, dptr, C.c_double, C.c_int, C.c_int]),
        "linearstep": (None, [vp, vp, vp, plan, plan, plan, iptr]),
        "setStep": (None, [iptr, dptr, vp, vp, C.c_double, C.c_double, dptr, dptr, iptr, dptr]),
        "ioniz": (None, [vp, iptr, C.c_double, C.c_double, C.c_char_p, C.c_double, dptr, C.c_double, C.c_double,
                         dptr, dptr]),
        "ADK": (None, [dptr, dptr, iptr, dptr, dptr, dptr, dptr, dptr, dptr]),
        "calcexpR": (None, [vp, iptr, C.c_double, C.c_double, C.c_double, C.c_double, dptr, iptr]),
        "fftshift": (None, [vp, iptr, vp, C.c_int]),
        "calc1oes": (None, [dptr, iptr, dptr, dptr, dptr]),
        "ionizpot": (None, [C.c_char_p, dptr, dptr]),
        "Natoms": (None, [C.c_char_p, dptr, dptr]),
        "copyfile": (None, [C.c_char_p, C.c_int]),
        "restore": (None, [C.c_char_p]),
        "mpc_adk_constants": (C.c_int, [C.c_char_p, dptr]),
    })
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    if L.mpc_sizeof_params() != C.sizeof(MpcParams):
        raise RuntimeError("mpc_params layout mismatch between C and Python")
    if L.mpc_sizeof_input_spec() != C.sizeof(MpcInputSpec):
        raise RuntimeError("mpc_input_spec layout mismatch between C and Python")
    _lib = L
    return L


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().mpc_strerror(rc).decode()
        raise RuntimeError(f"{what}: {msg} ({rc})")
