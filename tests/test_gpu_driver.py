"""GPU test of the run driver (mpc_amd/mpc, the MPC.c replacement): files,
dump formats and kill-after-checkpoint / resume (-r) against an
uninterrupted run, which must agree bit for bit (same device arithmetic)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from mpc_amd import configs, make_inputs
from mpc_amd.mkinput import write_inputs

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MPC = os.path.join(ROOT, "mpc_amd", "mpc")


def _run(args, cwd):
    print("driver:", " ".join(args), flush=True)
    p = subprocess.run([MPC, *args], cwd=cwd, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    return p.stdout


def test_driver_fixed_checkpoint_resume(gpu_lib, tmp_path):
    wl = configs.small(16, 512, steps_per_pass=10, passes=2, mirror_R=0.3)
    inp = make_inputs(wl)
    base = str(tmp_path / "data")
    rundir = write_inputs(inp, base, 0)
    alt = str(tmp_path / "alt")
    shutil.copytree(tmp_path, tmp_path.parent / (tmp_path.name + "_copy"))
    copy = str(tmp_path.parent / (tmp_path.name + "_copy") / "data")
    args = ["-d", base, "-n", "0", "-F", "10", "-R", "0.3", "-k", "1"]
    _run(args, str(tmp_path))
    full = np.fromfile(os.path.join(rundir, "out.bin"))
    # interrupted run: stop after 15 dumps (last checkpoint at step 10), resume
    argsc = ["-d", copy, "-n", "0", "-F", "10", "-R", "0.3", "-k", "1"]
    _run(argsc + ["-m", "15"], os.path.dirname(copy))
    _run(argsc + ["-r"], os.path.dirname(copy))
    crun = copy + "0000"
    resumed = np.fromfile(os.path.join(crun, "out.bin"))
    assert np.array_equal(full, resumed)
    for f in ("bp1.bin", "bp2.bin", "temp.bin", "spec.bin", "other.bin"):
        a = open(os.path.join(rundir, f), "rb").read()
        b = open(os.path.join(crun, f), "rb").read()
        assert a == b, f
    # formats: 20 dumps of Nx / Nt doubles; out.bin = 2N doubles
    assert os.path.getsize(os.path.join(rundir, "bp1.bin")) == 20 * 16 * 8
    assert os.path.getsize(os.path.join(rundir, "temp.bin")) == 20 * 512 * 8
    assert full.size == 2 * 16 * 16 * 512
    assert os.path.getsize(os.path.join(rundir, "resparam.bin")) == 32
    # and the field agrees with the oracle's fixed-step run
    ref, _ = O.run_fixed(inp, inp.calibrated(), 0, 20)
    got = full.view(np.complex128).reshape(ref.shape)
    assert O.rel_l2(got, ref) < 1e-10


def test_driver_adaptive_runs(gpu_lib, tmp_path):
    wl = configs.small(16, 512, steps_per_pass=10, passes=1, mirror_R=0.0)
    wl.def_err = 1e-4
    inp = make_inputs(wl)
    base = str(tmp_path / "data")
    rundir = write_inputs(inp, base, 0)
    out = _run(["-d", base, "-n", "0", "-m", "4"], str(tmp_path))
    assert "We are here" in out
    lines = open(os.path.join(rundir, "other.bin")).read().strip().splitlines()
    assert len(lines) == 4 and all("\t" in l for l in lines)
    # the driver's 4 accepted steps against the oracle's adaptive run: the
    # field, and err / dz as other.bin records them ("%f\t%f")
    import ctypes as C
    ref = inp.calibrated()
    po = O.params_from(inp)
    z, dz, sm = C.c_double(0.0), C.c_double(inp.params.dzz), C.c_int(1)
    errs, dzs = np.zeros(4), np.zeros(4)
    n = O.oracle().oracle_run_adaptive(O.vp(ref), C.byref(po), C.byref(z), C.byref(dz), C.byref(sm), 4,
                                       O.dp(errs), O.dp(dzs))
    assert n == 4
    got = np.fromfile(os.path.join(rundir, "out.bin")).view(np.complex128).reshape(ref.shape)
    err = O.rel_l2(got, ref)
    print(f"  driver adaptive vs oracle: rel_l2={err:.3e}", flush=True)
    assert err < 1e-10
    for line, e, d in zip(lines, errs, dzs):
        le, ld = (float(v) for v in line.split("\t"))
        assert le == float(f"{e:f}") and ld == float(f"{d:f}")


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_snapshot_is_the_field_at_its_begin(gpu_lib, prec):
    """mpc_snapshot_begin queues the copy and returns; steps run afterwards
    (across a mirror) and a download in between do not change what the
    snapshot holds; a second snapshot replaces it"""
    from mpc_amd import Propagator
    print(f"case snapshot prec={prec}", flush=True)
    inp = make_inputs(configs.small(32, 512, steps_per_pass=3, passes=3))
    A0 = inp.calibrated()
    with Propagator(inp, prec) as p:
        p.upload(A0)
        p.run(0, 2)
        at2 = p.download()
        p.snapshot_begin()
        p.run(2, 4)
        at6 = p.download()
        snap = p.snapshot_wait()
        assert np.array_equal(snap, at2)
        assert p.snapshot_ready()
        p.snapshot_begin()
        p.run(6, 1)
        assert np.array_equal(p.snapshot_wait(), at6)
        assert not np.array_equal(at6, at2)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_guard_stops_a_run_on_a_non_finite_field(gpu_lib, prec):
    """with the guard on, a field holding a NaN makes mpc_run fail with
    MPC_ENONFINITE; a normal run passes the guard with the same bits as one
    without it"""
    from mpc_amd import Propagator
    print(f"case guard prec={prec}", flush=True)
    inp = make_inputs(configs.small(32, 512, steps_per_pass=3, passes=2))
    A0 = inp.calibrated()
    with Propagator(inp, prec) as p:
        p.upload(A0)
        p.run(0, 4)
        ref = p.download()
        p.upload(A0)
        p.set_guard(1e-2)
        e0 = p.energy()
        assert abs(e0 / (np.abs(A0) ** 2).sum() - 1) < 1e-6
        p.run(0, 4)
        assert np.array_equal(p.download(), ref)
        bad = A0.copy()
        bad[100, 5, 7] = np.nan
        p.upload(bad)
        with pytest.raises(RuntimeError, match="non-finite"):
            p.run(0, 1)
        p.set_guard(-1.0)
        p.upload(bad)
        p.run(0, 1)                                # unguarded: runs, the NaN spreads
        assert not np.isfinite(p.download()).all()
