"""The multi-type circuit harness on the oracle alone (no GPU), on reduced
netlists: the load sequence runs, every entry is finite, and each device
type's terms land where its stamps say (a circuit of only wire resistors
gives the nodal conductance matrix and Ohm's-law currents)."""
import numpy as np

from circuit import CircuitPair, two_terminal
from netlist_circuits import hb_circuit, logic_circuit


def test_reduced_netlists_load():
    for build in (lambda: logic_circuit(4000), lambda: hb_circuit(4, 200)):
        parts, n_nodes, volts = build()
        cp = CircuitPair(parts, n_nodes, gpu=False)
        cp.set_inputs(nextSol=cp.unknowns(volts, 1e-4))
        ro, _ = cp.load(newtonIter=0, dcopFlag=1, initJctFlag=1)
        for k, v in ro.items():
            assert np.all(np.isfinite(v)), k
        assert np.any(ro["dFdx"] != 0) and np.any(ro["q"] != 0)


def test_resistor_network_is_nodal_analysis():
    rng = np.random.default_rng(5)
    n_nodes, m = 30, 80
    nodes = rng.integers(-1, n_nodes, size=(m, 2)).astype(np.int32)
    nodes = nodes[nodes[:, 0] != nodes[:, 1]]
    R = rng.uniform(10.0, 1000.0, len(nodes))
    cp = CircuitPair([two_terminal("resistor", nodes, R, "R")], n_nodes, gpu=False)
    x = rng.uniform(-1, 1, n_nodes)
    cp.set_inputs(nextSol=x)
    ro, _ = cp.load(newtonIter=1, dcopFlag=1)
    G = np.zeros((n_nodes, n_nodes))
    for (a, b), r in zip(nodes, R):
        for i, j, s in ((a, a, 1), (b, b, 1), (a, b, -1), (b, a, -1)):
            if i >= 0 and j >= 0:
                G[i, j] += s / r
    dense = np.zeros((n_nodes, n_nodes))
    rows = np.repeat(np.arange(n_nodes), np.diff(cp.row_ptr))
    dense[rows, cp.col_idx] = ro["dFdx"]
    np.testing.assert_allclose(dense, G, rtol=1e-13, atol=1e-16)
    np.testing.assert_allclose(ro["f"], G @ x, rtol=1e-12, atol=1e-14)
