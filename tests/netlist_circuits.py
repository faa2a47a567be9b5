"""The shared-node netlists as multi-type circuits for the parity tests:
the flat logic netlist (BSIM4 transistors, wire resistors and capacitors,
the supply and 20 PULSE inputs) and the harmonic-balance samples of the
differential amplifier (BSIM4 transistors, the output RC filter and the
supply, bias and input sources of every sample copy)."""
import numpy as np

from circuit import Part, two_terminal
from xyce_amd import netlists


def logic_circuit(n_transistors=200_000, seed=netlists.SEED):
    cards, mi, params, lids, n_nodes, volts = netlists.logic_dag(n_transistors, seed=seed)
    r_nodes, R, c_nodes, Cv, _ = netlists.logic_wires(n_transistors, seed=seed)
    parts = [Part("bsim4", cards, mi, params, lids),
             two_terminal("resistor", r_nodes, R, "R"),
             two_terminal("capacitor", c_nodes, Cv, "C"),
             _sources(*netlists.logic_sources())]
    return parts, n_nodes, volts


def _sources(nodes, params):
    n = len(nodes)
    return Part("vsource", [(1, {})], np.zeros(n, np.int32), params, nodes.reshape(-1))


def hb_circuit(samples=128, n_transistors=2000, seed=netlists.SEED):
    cards, mi, params, lids, _, _ = netlists.hb_diffamp(samples, n_transistors, seed)
    r_nodes, R, c_nodes, Cv, n_nodes, volts = netlists.hb_filter(samples, n_transistors, seed)
    parts = [Part("bsim4", cards, mi, params, lids),
             two_terminal("resistor", r_nodes, R, "R"),
             two_terminal("capacitor", c_nodes, Cv, "C"),
             _sources(*netlists.hb_sources(samples, n_transistors, seed))]
    return parts, n_nodes, volts
