"""Ring-oscillator driver on the CPU oracle: the DC operating point solves
KCL at every free node, the initial condition holds, and the 51-stage ring
oscillates."""
import numpy as np

from ring_oracle import OracleLoader
from xyce_amd.ring import Ring, crossings, transient


def test_dc_operating_point_solves_kcl():
    ring = Ring(stages=5)
    loader = OracleLoader(ring)
    t, V, stats = transient(loader, ring, t_stop=0.0)
    x = V[0]
    assert x[ring.ic_node] == ring.ic_value and x[ring.supply] == ring.vdd
    out = loader.load(x, newtonIter=5, dcopFlag=1)
    free = [k for k in range(ring.N) if k != ring.ic_node]
    # node currents balance to far below a device's drive current
    assert np.max(np.abs(out["f"][free])) < 1e-15
    # the inverter chain alternates away from the held node
    assert x[1] > 0.9 * ring.vdd and x[2] < 0.1 * ring.vdd


def test_ring51_oscillates():
    ring = Ring(stages=51)
    t, V, stats = transient(OracleLoader(ring), ring, t_stop=1.5e-9)
    assert stats["newton_iterations"] < 5 * stats["steps"]
    # every stage output swings through mid-rail at least once
    for k in range(ring.N):
        assert len(crossings(t, V[:, k], 0.5, rising=True)) + \
               len(crossings(t, V[:, k], 0.5, rising=False)) >= 1
    assert V[:, :ring.N].min() > -0.1 and V[:, :ring.N].max() < 1.1
