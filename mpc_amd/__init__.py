"""mpc_amd -- MI355X-native split-step Fourier propagator (drop-in for MPC's
propagation.h / plasma.h / utils.h entry points).

The product is the C-ABI library libmpc_hip.so (include/*.h); this package
is a thin Python front end over that ABI for tests and benchmarks:

    from mpc_amd import make_inputs, Propagator, configs
    inp = make_inputs(configs.cfg1_cpu_64())
    with Propagator(inp, "f64") as p:
        p.upload(inp.A); p.calibrate(); p.run(0, 200)
        A = p.download(); d = p.diagnostics()
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import configs
from ._lib import MPC_F32, MPC_F64, MpcDiag, MpcInputSpec, MpcParams, check, dptr, lib

__all__ = ["Inputs", "make_inputs", "Propagator", "configs", "MPC_F32", "MPC_F64", "lib"]


def _dp(a: np.ndarray):
    return a.ctypes.data_as(dptr)


@dataclass
class Inputs:
    """A workload's param.bin content and its unscaled input field."""
    params: MpcParams
    betas: np.ndarray
    wr: np.ndarray
    w: np.ndarray
    r: np.ndarray
    A: np.ndarray          # complex128, shape (Nt, Ny, Nx): reference layout A[t][y][x]

    @property
    def dims(self):
        return tuple(self.params.dims)

    def calibrated(self) -> np.ndarray:
        """A * sqrt(puls_e / Esig), both parts (MPC.c:206-216 with the real-only scaling fixed)."""
        k = np.sqrt(self.params.puls_e / self.params.Esig)
        return self.A * k


def make_inputs(wl: "configs.Workload") -> Inputs:
    s = MpcInputSpec()
    s.nx = s.ny = wl.nx
    s.nt = wl.nt
    s.passes = wl.passes
    s.steps_per_pass = wl.steps_per_pass
    s.sst_on = wl.sst_on
    s.plasm = wl.plasm
    s.window_xy = wl.window_xy
    s.window_t = wl.window_t
    s.lambda0 = wl.lambda0
    s.waist = wl.waist
    s.tau_fwhm = wl.tau_fwhm
    s.energy = wl.energy
    s.noise_rel = wl.noise_rel
    s.n2_per_bar = wl.n2_per_bar
    s.beta2 = wl.beta2_fs2_per_m * 1e-30
    s.pressure = wl.pressure
    s.rho_nt_per_bar = wl.rho_nt_per_bar
    s.cell_L = wl.cell_L
    s.mirror_R = wl.mirror_R
    s.aperture = wl.aperture
    s.Ab_M = wl.Ab_M
    s.def_err = wl.def_err
    s.dzmin = wl.dzmin
    s.seed = wl.seed
    s.gas = wl.gas.encode()
    nx, ny, nt = wl.nx, wl.nx, wl.nt
    betas, w = np.zeros(nt), np.zeros(nt)
    wr, r = np.zeros(ny), np.zeros(ny)
    A = np.empty((nt, ny, nx), dtype=np.complex128)
    p = MpcParams()
    check(lib().mpc_mkinput(C.byref(s), C.byref(p), _dp(betas), _dp(wr), _dp(w), _dp(r),
                            A.ctypes.data_as(C.c_void_p)), "mpc_mkinput")
    inp = Inputs(p, betas, wr, w, r, A)
    _bind_arrays(inp)
    return inp


def _bind_arrays(inp: Inputs) -> None:
    p = inp.params
    p.betas, p.wr, p.w, p.r = _dp(inp.betas), _dp(inp.wr), _dp(inp.w), _dp(inp.r)


def variant_counts() -> tuple:
    """(number of t-line kernel variants, number of x/y-line kernel variants)"""
    tv, av = C.c_int(), C.c_int()
    lib().mpc_variant_counts(C.byref(tv), C.byref(av))
    return tv.value, av.value


class Propagator:
    """A device-resident field and its run context (mpc_ctx)."""

    def __init__(self, inp: Inputs, precision: str = "f64"):
        self.inp = inp
        _bind_arrays(inp)
        self.prec = MPC_F32 if precision in ("f32", "fp32", MPC_F32) else MPC_F64
        self.ctx = C.c_void_p()
        check(lib().mpc_create(C.byref(inp.params), self.prec, C.byref(self.ctx)), "mpc_create")
        self.nt, self.ny, self.nx = inp.A.shape

    def close(self):
        if self.ctx:
            lib().mpc_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, A: np.ndarray) -> None:
        A = np.ascontiguousarray(A, dtype=np.complex128)
        assert A.shape == (self.nt, self.ny, self.nx)
        check(lib().mpc_upload(self.ctx, A.ctypes.data_as(C.c_void_p)), "mpc_upload")

    def download(self) -> np.ndarray:
        A = np.empty((self.nt, self.ny, self.nx), dtype=np.complex128)
        check(lib().mpc_download(self.ctx, A.ctypes.data_as(C.c_void_p)), "mpc_download")
        return A

    def set_guard(self, gain_tol: float) -> None:
        """mpc_run checks the energy stays finite and grows by at most gain_tol (< 0: off)"""
        check(lib().mpc_set_guard(self.ctx, gain_tol), "mpc_set_guard")

    def energy(self) -> float:
        e = C.c_double()
        check(lib().mpc_energy(self.ctx, C.byref(e)), "mpc_energy")
        return e.value

    def snapshot_begin(self) -> None:
        """queue an asynchronous copy of the field to the context's pinned host buffer"""
        check(lib().mpc_snapshot_begin(self.ctx), "mpc_snapshot_begin")

    def snapshot_ready(self) -> bool:
        rc = lib().mpc_snapshot_poll(self.ctx)
        if rc < 0:
            check(rc, "mpc_snapshot_poll")
        return rc == 1

    def snapshot_wait(self) -> np.ndarray:
        """the snapshot (a copy of the pinned buffer), reference layout"""
        p = C.c_void_p()
        check(lib().mpc_snapshot_wait(self.ctx, C.byref(p)), "mpc_snapshot_wait")
        n = self.nt * self.ny * self.nx
        buf = (C.c_double * (2 * n)).from_address(p.value)
        return np.frombuffer(buf, dtype=np.complex128).reshape(self.nt, self.ny, self.nx).copy()

    def calibrate(self) -> None:
        check(lib().mpc_calibrate(self.ctx), "mpc_calibrate")

    def run(self, step0: int, nsteps: int) -> None:
        check(lib().mpc_run(self.ctx, step0, nsteps), "mpc_run")

    def step(self, dzz: float, z: float, signmem: int = 1) -> int:
        sm = C.c_int(signmem)
        check(lib().mpc_step(self.ctx, dzz, z, C.byref(sm)), "mpc_step")
        return sm.value

    def mirror(self) -> None:
        check(lib().mpc_mirror(self.ctx), "mpc_mirror")

    def linear(self, length: float, sign: int = 1, mirror: bool = False) -> None:
        """linear propagation only over `length` (then the mirror if asked)"""
        check(lib().mpc_linear(self.ctx, length, sign, 1 if mirror else 0), "mpc_linear")

    def diagnostics(self) -> dict:
        d = MpcDiag()
        temp, spec = np.zeros(self.nt), np.zeros(self.nt)
        bp1, bp2 = np.zeros(self.nx), np.zeros(self.ny)
        d.temp, d.spec, d.BP1, d.BP2 = _dp(temp), _dp(spec), _dp(bp1), _dp(bp2)
        check(lib().mpc_diagnostics(self.ctx, C.byref(d)), "mpc_diagnostics")
        return {"energy_sum": d.energy_sum, "rho_peak": d.rho_peak, "spec_rms_width": d.spec_rms_width,
                "temp": temp, "spec": spec, "BP1": bp1, "BP2": bp2}

    def run_adaptive(self, z: float, dzz: float, signmem: int, max_steps: int):
        zc, dc, sm, acc = C.c_double(z), C.c_double(dzz), C.c_int(signmem), C.c_int(0)
        errs, dzs = np.zeros(max_steps), np.zeros(max_steps)
        check(lib().mpc_run_adaptive(self.ctx, C.byref(zc), C.byref(dc), C.byref(sm), max_steps, C.byref(acc),
                                     _dp(errs), _dp(dzs)), "mpc_run_adaptive")
        n = acc.value
        return {"z": zc.value, "dzz": dc.value, "signmem": sm.value, "accepted": n,
                "errs": errs[:n], "dzs": dzs[:n]}

    def autotune(self) -> tuple:
        check(lib().mpc_autotune(self.ctx), "mpc_autotune")
        return self.variants()

    def variants(self) -> tuple:
        """(t-line, x-line, y-line) kernel variants in use"""
        tv, xv, yv = C.c_int(), C.c_int(), C.c_int()
        check(lib().mpc_get_variants_xy(self.ctx, C.byref(tv), C.byref(xv), C.byref(yv)), "mpc_get_variants_xy")
        return tv.value, xv.value, yv.value

    def set_variants(self, tline: int, x: int, y: int | None = None) -> None:
        check(lib().mpc_set_variants_xy(self.ctx, tline, x, x if y is None else y), "mpc_set_variants_xy")

    def set_timing(self, on: bool) -> None:
        check(lib().mpc_set_timing(self.ctx, 1 if on else 0), "mpc_set_timing")

    def kernel_ms(self, which: int) -> float:
        return lib().mpc_kernel_ms(self.ctx, which)


class Batch:
    """Pulses advancing together (mpc_batch): every sweep of a step is one
    launch per group of same-size pulses (at most 256 MB of fields per
    group), and with graph=True each pass's launches are a HIP graph
    captured once and replayed.  The propagators stay owned by the
    caller; close the batch first."""

    def __init__(self, props: list):
        self.props = list(props)
        arr = (C.c_void_p * len(self.props))(*[p.ctx.value for p in self.props])
        self.b = C.c_void_p()
        check(lib().mpc_batch_create(arr, len(self.props), C.byref(self.b)), "mpc_batch_create")

    def run(self, step0: int, nsteps: int, graph: bool = True) -> None:
        check(lib().mpc_batch_run(self.b, step0, nsteps, 1 if graph else 0), "mpc_batch_run")

    def graphs(self) -> int:
        return lib().mpc_batch_graphs(self.b)

    def autotune(self) -> None:
        """pick each launch class's t-line form by timing the batch's own
        launches (after the pulses' own autotune; fields are left as they were)"""
        check(lib().mpc_batch_autotune(self.b), "mpc_batch_autotune")

    def close(self):
        if self.b:
            lib().mpc_batch_destroy(self.b)
            self.b = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
