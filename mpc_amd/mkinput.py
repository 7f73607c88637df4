"""Write a synthetic workload as the reference driver's input files.

    python -m mpc_amd.mkinput --config cfg1 --dir /path/data [--run 0]

writes <dir>/datar.bin, <dir>/datai.bin (N doubles each, A[t][y][x]) and
<dir>%04d/param.bin (MPC.c:76-77,159-168,180-203 formats) through the
library's C generator (mpc_mkinput) and param.bin writer.
"""
from __future__ import annotations

import argparse
import ctypes as C
import os

import numpy as np

from . import configs, lib, make_inputs


def write_inputs(inp, base: str, run: int = 0) -> str:
    os.makedirs(base, exist_ok=True)
    rundir = f"{base}{run:04d}"
    os.makedirs(rundir, exist_ok=True)
    inp.A.real.astype(np.float64).tofile(os.path.join(base, "datar.bin"))
    inp.A.imag.astype(np.float64).tofile(os.path.join(base, "datai.bin"))
    rc = lib().mpc_write_param_bin(os.path.join(rundir, "param.bin").encode(), C.byref(inp.params))
    if rc != 0:
        raise RuntimeError("mpc_write_param_bin failed")
    return rundir


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--config", default="cfg1", help="cfg1 | cfg2 | cfg3 | small")
    ap.add_argument("--dir", required=True)
    ap.add_argument("--run", type=int, default=0)
    a = ap.parse_args(argv)
    wl = configs.small() if a.config == "small" else configs.CONFIGS[a.config]()
    print(write_inputs(make_inputs(wl), a.dir, a.run))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
