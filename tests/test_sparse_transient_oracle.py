"""The oracle's netlist transient (tests/sparse_transient_oracle.py), pinned
by the closed-form recurrences of its integrator on a linear RC driven by a
voltage step: backward Euler for the first two steps,
v1 = (v0 + h/tau) / (1 + h/tau), then the trapezoid rule,
v1 = (v0 (1 - h/(2 tau)) + h/tau) / (1 + h/(2 tau)), tau = RC
(N_TIA_OneStep.C:272-335, 463-500; order 1 for two steps,
N_TIA_OneStep.C:2130,2211).  A planted error -- trapezoid from the first
step -- fails it."""
import numpy as np

from sparse_transient_oracle import OracleTransient
from xyce_amd.netlist_reader import Netlist

TEXT = """rc step
V1 in 0 PULSE(0 1 0 1e-16 1e-16 1 2)
R1 in out 1000
C1 out 0 1e-12
"""


def _expected(steps, h, tau, be_steps=2):
    v = [0.0]
    for k in range(steps):
        a = h / tau
        if k < be_steps:
            v.append((v[-1] + a) / (1 + a))
        else:
            v.append((v[-1] * (1 - a / 2) + a) / (1 + a / 2))
    return np.array(v)


def test_rc_step_follows_the_integrator_recurrences():
    nl = Netlist(TEXT)
    ot = OracleTransient(nl, vtol=1e-13)
    out = nl.node("out")
    t, V, nw = ot.run(np.zeros(ot.n), 40e-12, 1e-12, record=[out])
    want = _expected(40, 1e-12, 1e-9)
    np.testing.assert_allclose(V[:, 0], want, rtol=0, atol=1e-13)
    # the trapezoid rule from the first step would differ by O(h/tau)^2
    assert np.max(np.abs(_expected(40, 1e-12, 1e-9, be_steps=0) - want)) > 1e-8
