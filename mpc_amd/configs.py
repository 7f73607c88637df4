"""Synthetic workloads: the five benchmark configurations plus small grids.

Physical parameters follow the configuration descriptions (wavelength
1030 nm, argon, Gaussian pulse with seeded complex noise at 1e-6 of the peak
amplitude); param.bin quantities are derived by mpc_mkinput (host.c).
"""
from __future__ import annotations

from dataclasses import dataclass, field

FS = 1e-15


@dataclass
class Workload:
    name: str
    nx: int
    nt: int
    window_xy: float
    window_t: float
    waist: float
    tau_fwhm: float
    energy: float
    pressure: float
    beta2_fs2_per_m: float
    passes: int
    steps_per_pass: int
    cell_L: float = 0.5
    mirror_R: float = 0.3
    aperture: float = 0.0
    Ab_M: float = 1.0
    sst_on: int = 1
    plasm: int = 1
    seed: int = 20261202
    noise_rel: float = 1e-6
    lambda0: float = 1030e-9
    n2_per_bar: float = 1e-23
    rho_nt_per_bar: float = 2.5e25
    def_err: float = 1e-3
    dzmin: float = 1e-6
    gas: str = "Argon"
    extra: dict = field(default_factory=dict)


def cfg1_cpu_64() -> Workload:
    """64x64x512, 4 mm x 2 ps, 10 uJ, 200 fs, w0 0.5 mm, Ar 1 bar, one pass of 200 steps."""
    return Workload("64x64x512 Ar 1 bar, 10 uJ 200 fs, 1 pass x 200 steps", 64, 512, 4e-3, 2e-12,
                    0.5e-3, 200 * FS, 10e-6, 1.0, 20.0, 1, 200, seed=20261202 + 0)


def cfg2_production_256() -> Workload:
    """256x256x2048, 6 mm x 4 ps, 1 mJ, 300 fs, w0 0.8 mm, Ar 2 bar, 20 passes x 500 steps."""
    return Workload("256x256x2048 Ar 2 bar, 1 mJ 300 fs, 20 passes x 500 steps", 256, 2048, 6e-3, 4e-12,
                    0.8e-3, 300 * FS, 1e-3, 2.0, 40.0, 20, 500, seed=20261202 + 1)


def cfg3_precision_64() -> Workload:
    """config-1 workload, run in the FP32 form."""
    w = cfg1_cpu_64()
    w.name = "64x64x512 FP32 precision form of the 64x64x512 workload"
    w.seed = 20261202 + 2
    return w


def cfg4_batch(npulses: int = 48, seed: int = 20261202 + 3) -> list:
    """48 pulses with grids drawn by seed from {128^2 x 512, 192^2 x 1024,
    256^2 x 2048}, energies log-uniform 0.1-2 mJ, durations uniform 100-500
    fs, argon 0.5-3 bar, w0 0.5-1.0 mm; 10 passes x 300 steps.  Every grid
    spans the 6 mm x 4 ps window of the 256^2 x 2048 workload."""
    import numpy as np
    rng = np.random.default_rng(seed)
    grids = [(128, 512), (192, 1024), (256, 2048)]
    out = []
    for i in range(npulses):
        nx, nt = grids[int(rng.integers(len(grids)))]
        energy = float(10 ** rng.uniform(np.log10(0.1e-3), np.log10(2e-3)))
        tau = float(rng.uniform(100, 500)) * FS
        pressure = float(rng.uniform(0.5, 3.0))
        waist = float(rng.uniform(0.5e-3, 1.0e-3))
        out.append(Workload(f"batch pulse {i}: {nx}x{nx}x{nt}", nx, nt, 6e-3, 4e-12, waist, tau, energy, pressure,
                            40.0, 10, 300, seed=seed * 100 + i))
    return out


def small(nx: int = 16, nt: int = 64, plasm: int = 1, sst_on: int = 1, mirror_R: float = 0.3,
          energy: float = 1e-3, steps_per_pass: int = 4, passes: int = 2, seed: int = 7,
          pressure: float = 2.0, waist: float = 0.2e-3) -> Workload:
    """Tiny grids for parity tests: intense enough that plasma terms are non-zero."""
    return Workload(f"{nx}x{nx}x{nt} test", nx, nt, 2e-3, 1e-12, waist, 100 * FS, energy, pressure, 40.0,
                    passes, steps_per_pass, cell_L=0.05, mirror_R=mirror_R, sst_on=sst_on, plasm=plasm,
                    seed=seed)


CONFIGS = {
    "cfg1": cfg1_cpu_64,
    "cfg2": cfg2_production_256,
    "cfg3": cfg3_precision_64,
}
