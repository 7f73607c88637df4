"""Build libmpc_hip.so (gfx950) in-tree, and the oracle's liboracle.so.

    python -m mpc_amd.build            # library + driver tools
    python -m mpc_amd.build --oracle   # also oracle/liboracle.so (test infrastructure)

The library is compiled every time this runs (no staleness check): hipcc
for the HIP/C++ translation units, gcc for the plain-C host code, linked
with hipcc.  Objects go to mpc_amd/_build/ (git-ignored).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmpc_hip.so")
ARCH = os.environ.get("MPC_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
# packed FP32 vector ops are kept out of the generated code on purpose
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
C_FLAGS = ["-O2", "-fPIC", "-std=c11", "-fopenmp", "-Wall", "-Wno-unused-result"]

HIP_SOURCES = ["tline.hip", "axis.hip", "kernels.hip", "fno.hip", "noalloc.hip", "api.cpp"]
# per-file flags: the persistent sweep kernels keep machine-level LICM from
# hoisting constants and twiddle products out of their loops into registers
EXTRA_FLAGS = {"tline.hip": ["-mllvm", "-disable-machine-licm"],
               "axis.hip": ["-mllvm", "-disable-machine-licm"]}
C_SOURCES = ["host.c"]


def _run(cmd: list[str]) -> None:
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + p.stdout + p.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")


def build_library(verbose: bool = False) -> str:
    os.makedirs(OUT, exist_ok=True)
    jobs = []
    objs = []
    for s in HIP_SOURCES:
        o = os.path.join(OUT, os.path.splitext(s)[0] + ".o")
        src = os.path.join(CSRC, s)
        lang = ["-x", "hip"] if s.endswith(".cpp") else []
        jobs.append([HIPCC, *HIP_FLAGS, *EXTRA_FLAGS.get(s, []), *lang, "-c", src, "-o", o])
        objs.append(o)
    for s in C_SOURCES:
        o = os.path.join(OUT, os.path.splitext(s)[0] + ".o")
        jobs.append(["gcc", *C_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    with cf.ThreadPoolExecutor(max_workers=min(4, len(jobs))) as ex:
        list(ex.map(_run, jobs))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-fopenmp", "-lm"])
    os.replace(tmp, LIB)
    # the run driver (MPC.c's replacement), linked against the library
    _run(["gcc", *C_FLAGS, "-pthread", "-o", os.path.join(HERE, "mpc"), os.path.join(CSRC, "mpc_main.c"),
          f"-L{HERE}", "-lmpc_hip", "-Wl,-rpath,$ORIGIN", "-lm"])
    if verbose:
        print("built", LIB, "and", os.path.join(HERE, "mpc"))
    return LIB


def build_oracle(verbose: bool = False) -> str:
    d = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-C", d, "-B", "liboracle.so"])
    if verbose:
        print("built", os.path.join(d, "liboracle.so"))
    return os.path.join(d, "liboracle.so")


def main(argv: list[str]) -> int:
    build_library(verbose=True)
    if "--oracle" in argv:
        build_oracle(verbose=True)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
