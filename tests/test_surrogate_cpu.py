"""CPU: the pass surrogate's host logic (no device calls).

* the pulse workloads are drawn inside their stated ranges, reproducibly
  per seed, with a grid the network's configuration matches;
* the constant input channels are O(1) over those ranges.
"""
import math

from mpc_amd import surrogate
from mpc_amd.fno import FnoConfig


def test_pulse_workloads_ranges_and_reproducibility():
    a = surrogate.pulse_workloads(64, seed=3, passes=5)
    b = surrogate.pulse_workloads(64, seed=3, passes=5)
    c = surrogate.pulse_workloads(64, seed=4, passes=5)
    assert [(w.energy, w.pressure, w.waist, w.tau_fwhm, w.seed) for w in a] == \
           [(w.energy, w.pressure, w.waist, w.tau_fwhm, w.seed) for w in b]
    assert [w.energy for w in a] != [w.energy for w in c]
    cfg = FnoConfig()
    for w in a:
        assert 5e-6 <= w.energy <= 20e-6
        assert 0.5 <= w.pressure <= 2.0
        assert 0.4e-3 <= w.waist <= 0.6e-3
        assert 150e-15 <= w.tau_fwhm <= 250e-15
        assert (w.nx, w.nt) == (cfg.X, cfg.T) and cfg.X == cfg.Y
        assert w.passes == 5 and w.steps_per_pass == 200
        assert math.isclose(w.cell_L / w.steps_per_pass * w.steps_per_pass, w.cell_L)
    assert len({w.seed for w in a}) == len(a)


def test_scalar_channels_are_order_one():
    for w in surrogate.pulse_workloads(64, seed=5, passes=2):
        le, p = surrogate.scalars_of(w)
        assert abs(le) <= math.log(2.0) + 1e-12
        assert -0.5 <= p <= 1.0
