"""The shared-node netlists as multi-type circuits for the parity tests:
the flat logic netlist (BSIM4 transistors, wire resistors and capacitors)
and the harmonic-balance samples of the differential amplifier (BSIM4
transistors and the output RC filter of every sample copy)."""
from circuit import Part, two_terminal
from xyce_amd import netlists


def logic_circuit(n_transistors=200_000, seed=netlists.SEED):
    cards, mi, params, lids, n_nodes, volts = netlists.logic_dag(n_transistors, seed=seed)
    r_nodes, R, c_nodes, Cv, _ = netlists.logic_wires(n_transistors, seed=seed)
    parts = [Part("bsim4", cards, mi, params, lids),
             two_terminal("resistor", r_nodes, R, "R"),
             two_terminal("capacitor", c_nodes, Cv, "C")]
    return parts, n_nodes, volts


def hb_circuit(samples=128, n_transistors=2000, seed=netlists.SEED):
    cards, mi, params, lids, _, _ = netlists.hb_diffamp(samples, n_transistors, seed)
    r_nodes, R, c_nodes, Cv, n_nodes, volts = netlists.hb_filter(samples, n_transistors, seed)
    parts = [Part("bsim4", cards, mi, params, lids),
             two_terminal("resistor", r_nodes, R, "R"),
             two_terminal("capacitor", c_nodes, Cv, "C")]
    return parts, n_nodes, volts
