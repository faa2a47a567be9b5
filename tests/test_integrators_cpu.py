"""The ring driver's integrators on the CPU oracle: backward Euler is first
order and the one-step integrator's trapezoid rule (N_TIA_OneStep.C:272-335,
463-500) second order in the step.  The node voltages of a 5-stage ring at
0.2 ns are compared with a trapezoid run at h/16 of the finest step; halving
h divides the error by about 2 (backward Euler) or 4 (trapezoid).  A dropped
1/2 on F or on the previous step's F leaves the rule first order or
inconsistent, which fails the order check."""
import numpy as np
import pytest

from ring_oracle import OracleLoader
from xyce_amd.ring import Ring, transient

T = 0.2e-9


def _final(ring, h, method):
    _, V, _ = transient(OracleLoader(ring), ring, t_stop=T, h=h, method=method)
    return V[-1, :ring.N]


def test_convergence_orders():
    ring = Ring(stages=5)
    ref = _final(ring, 1e-12 / 16, "trap")
    err = {m: [float(np.max(np.abs(_final(ring, h, m) - ref))) for h in (2e-12, 1e-12, 0.5e-12)]
           for m in ("be", "trap")}
    for m, lo, hi in (("be", 1.6, 3.0), ("trap", 3.2, 6.5)):
        e = err[m]
        assert lo < e[0] / e[1] < hi and lo < e[1] / e[2] < hi, (m, e)
    assert err["trap"][1] < 0.05 * err["be"][1]


def test_unknown_method_rejected():
    ring = Ring(stages=3)
    with pytest.raises(ValueError):
        transient(OracleLoader(ring), ring, t_stop=1e-12, method="gear")
