"""The SPICE-netlist subset reader (xa_netlist_*): host code only, no GPU.

The reader follows the reference's lexical rules -- title line, `*` and `;`
comments, `+` continuations, case-insensitive names, SPICE scale suffixes
(N_UTL_Misc.C:197-236), node 0 as ground -- and maps element letters and
.MODEL types onto the device types the library loads (M NMOS/PMOS level 1 ->
MOSFET1, 14/54 -> BSIM4; D; Q NPN/PNP; R; C; V)."""
import numpy as np
import pytest

from xyce_amd import XaError
from xyce_amd.netlist_reader import Netlist

RING3 = """three-stage ring
* inverters of the default BSIM4 cards
.model nch nmos level=54
.model pch pmos level=54 toxe=1.8n
MN1 n1 n3 0 0 nch W=90n L=45n
MP1 n1 n3 vdd vdd pch W=180n
+ L=45n                         ; continuation
MN2 n2 n1 0 0 nch W=90n L=45n
MP2 n2 n1 vdd vdd pch W=180n L=45n
MN3 n3 n2 0 0 nch W=90n L=45n
MP3 n3 n2 vdd vdd pch W=180n L=45n
C1 n1 0 1f
C2 n2 0 1f
C3 n3 0 1f
VDD vdd 0 DC 1
.ic V(n1)=0
.tran 1p 2n
.print tran v(n1)
.end
"""


def test_ring_netlist_parts_nodes_and_controls():
    nl = Netlist(RING3)
    assert nl.node_names == ["N1", "N3", "VDD", "N2"]
    assert nl.n_parts == 3
    assert nl.tran == (1e-12, 2 * 1e-9, 0.0, 0.0)
    nodes, vals = nl.ic
    assert list(nodes) == [0] and list(vals) == [0.0]
    m, c, v = nl.parts()
    assert m["type"] == "bsim4" and c["type"] == "capacitor" and v["type"] == "vsource"
    # a scaled number is the mantissa times the scale, as N_UTL_Misc.C:197-236
    # computes it (1.8 * 1e-9, one rounding off the literal 1.8e-9)
    assert m["models"] == [(54, {"TYPE": 1.0}), (54, {"TYPE": -1.0, "TOXE": 1.8 * 1e-9})]
    assert list(m["model_index"]) == [0, 1, 0, 1, 0, 1]
    assert list(m["ext_nodes"][:8]) == [0, 1, -1, -1, 0, 1, 2, 2]
    np.testing.assert_array_equal(m["params"]["W"], [90 * 1e-9, 180 * 1e-9] * 3)
    np.testing.assert_array_equal(m["params"]["L"], [45 * 1e-9] * 6)
    assert c["models"] == [(1, {})]
    np.testing.assert_array_equal(c["params"]["C"], [1e-15] * 3)
    assert list(c["ext_nodes"]) == [0, -1, 3, -1, 1, -1]
    assert v["params"] == {"DCV0": np.array([1.0])}


@pytest.mark.parametrize("tok,val", [
    ("1", 1.0), ("2.5n", 2.5 * 1e-9), ("1k", 1e3), ("1kohm", 1e3), ("1MEG", 1e6),
    ("3meg", 3 * 1e6), ("1m", 1e-3), ("1mH", 1e-3), ("2mil", 2 * 25.4e-6), ("10pF", 10 * 1e-12),
    ("3f", 3 * 1e-15), ("4u", 4 * 1e-6), ("1T", 1e12), ("2g", 2 * 1e9), ("1e-3", 1e-3),
    ("-2.5e2k", -2.5e2 * 1e3), ("5v", 5.0),
])
def test_scale_suffixes(tok, val):
    nl = Netlist(f"t\nR1 a 0 {tok}\n")
    assert nl.part(0)["params"]["R"][0] == val


def test_sources_waveforms():
    nl = Netlist("""sources
V1 a 0 SIN(0 1 1G 10p 0 90)
V2 b 0 PULSE(0 1 0 50p 50p 450p 1n)
V3 c 0 EXP(0 1 0 10p 100p 20p)
V4 d 0 DC=0.5
""")
    assert nl.n_parts == 4          # four different parameter sets
    p = [nl.part(k)["params"] for k in range(4)]
    assert {k: float(v[0]) for k, v in p[0].items()} == {
        "V0": 0.0, "VA": 1.0, "FREQ": 1e9, "TD": 10 * 1e-12, "THETA": 0.0, "PHASE": 90.0,
        "TRANSIENTSOURCETYPE": 0.0}
    assert {k: float(v[0]) for k, v in p[1].items()} == {
        "V1": 0.0, "V2": 1.0, "TD": 0.0, "TR": 50 * 1e-12, "TF": 50 * 1e-12,
        "PW": 450 * 1e-12, "PER": 1e-9, "TRANSIENTSOURCETYPE": 2.0}
    assert float(p[2]["TRANSIENTSOURCETYPE"][0]) == 1.0 and float(p[2]["TAU2"][0]) == 20 * 1e-12
    assert {k: float(v[0]) for k, v in p[3].items()} == {"DCV0": 0.5}


def test_diode_bjt_mosfet1_and_models():
    nl = Netlist("""mixed
.MODEL dm D IS=1e-14 N=1.1
.MODEL qn NPN BF=100
.MODEL qp PNP
.MODEL m1 NMOS LEVEL=1 VTO=0.7
D1 a b dm 2
D2 b 0 dm
Q1 c b e qn
Q2 c b e sub qp 3
M1 d g 0 0 m1 W=2u L=1u
""")
    parts = nl.parts()
    assert [p["type"] for p in parts] == ["diode", "diode", "bjt", "bjt", "mosfet1"]
    assert parts[0]["params"] == {"AREA": np.array([2.0])}
    assert parts[1]["params"] == {}
    assert parts[0]["models"] == [(1, {"IS": 1e-14, "N": 1.1})]
    names = nl.node_names
    assert list(parts[2]["ext_nodes"]) == [names.index("C"), names.index("B"), names.index("E"), -1]
    assert parts[2]["models"] == [(1, {"TYPE": 1.0, "BF": 100.0})]
    assert list(parts[3]["ext_nodes"]) == [names.index(x) for x in ("C", "B", "E", "SUB")]
    assert parts[3]["models"] == [(1, {"TYPE": -1.0})]
    assert parts[3]["params"] == {"AREA": np.array([3.0])}
    assert parts[4]["models"] == [(1, {"TYPE": 1.0, "VTO": 0.7})]


def test_step_lin_list_and_device_parameters():
    text = """steps
.model nch nmos level=54
M1 d g 0 0 nch W=90n L=45n
R1 d vdd 1k
V1 vdd 0 1
V2 g 0 0.5
.STEP nch:TOXE 1n 2n 0.5n
.STEP R1:R LIST 1k 2k 3k
"""
    nl = Netlist(text)
    assert nl.n_steps == 3
    for s, (tox, r) in enumerate([(1e-9, 1e3), (1.5e-9, 2e3), (2e-9, 3e3)]):
        m, rr, _ = nl.parts(step=s)
        assert m["models"][0][1]["TOXE"] == pytest.approx(tox, rel=1e-15)
        assert rr["params"]["R"][0] == r
    # the netlist's own values outside the sweep; a stepped parameter not on
    # the card is added from the sweep only when a step is selected
    m, rr, _ = nl.parts()
    assert "TOXE" not in m["models"][0][1] and rr["params"]["R"][0] == 1e3


def test_step_adds_parameter_not_on_the_element():
    nl = Netlist("t\nR1 a 0 1k\nR2 a 0 2k\nC1 a 0 1f\n.STEP C1:TC1 LIST 0.1 0.2\n")
    c = nl.part(1)
    assert set(c["params"]) == {"C", "TC1"}
    assert nl.part(1, step=1)["params"]["TC1"][0] == 0.2


@pytest.mark.parametrize("text,msg", [
    ("t\nL1 a b 1n\n", "not a device-load element"),
    ("t\nX1 a b sub\n", "subcircuit SUB not defined"),
    ("t\n.subckt s a b\nR1 a b 1\n.ends\nX1 n1 s\n", "X1 connects 1 nodes, S has 2 ports"),
    ("t\n.ends\n", ".ENDS without .SUBCKT"),
    ("t\n.subckt s a\n.subckt u b\n.ends\n.ends\n", "nested .SUBCKT"),
    ("t\n.subckt s a b params: w=1\n.ends\n", "subcircuit parameters"),
    ("t\n.subckt s a\nR1 a 0 1\n", ".SUBCKT without .ENDS"),
    ("t\n.subckt s a\nX1 a s\n.ends\nX1 n s\n", "nested deeper than 32"),
    ("t\n.model m nmos level=8\nM1 a b c d m\n", "MOSFET level 8"),
    ("t\nM1 a b c d nomodel\n", "model NOMODEL not defined"),
    ("t\n.model dm d\nM1 a b c d dm\n", "does not fit"),
    ("t\nR1 a 0 1k\n.STEP R2:R LIST 1 2\n", "target R2 not found"),
    ("t\n.STEP R1:R 1 2 -1\nR1 a 0 1\n", "wrong sign"),
    ("t\n.include models.lib\n", "control line .INCLUDE is not read"),
    ("t\nV1 a 0 SIN(0 1\n", "unclosed SIN("),
    ("t\n.ic V(0)=1\n", ".IC on ground"),
])
def test_errors_name_the_problem(text, msg):
    with pytest.raises(XaError, match=msg.replace("(", r"\(")):
        Netlist(text)


def test_error_names_the_line():
    with pytest.raises(XaError, match="netlist line 4"):
        Netlist("t\n* comment\nR1 a 0 1\nR2 a 0 bogus-value\n")


def test_end_stops_reading_and_options_are_skipped():
    nl = Netlist("t\n.options abstol=1e-6\nR1 a 0 1\n.end\nthis line is never read\n")
    assert nl.n_parts == 1 and nl.node_names == ["A"]


def _logic_text(n_transistors=600):
    """The synthetic logic netlist written as SPICE text (nodes named by
    their generator ids, values printed exactly)."""
    from xyce_amd import netlists
    cards, mi, params, lids, n_nodes, _ = netlists.logic_dag(n_transistors)
    r_nodes, R, c_nodes, Cv, _ = netlists.logic_wires(n_transistors)
    src_nodes, src = netlists.logic_sources()
    name = lambda k: "0" if k < 0 else f"n{k}"
    lines = ["logic netlist"]
    for j, (lev, card) in enumerate(cards):
        kind = "nmos" if card["TYPE"] > 0 else "pmos"
        lines.append(f".model c{j} {kind} level={lev}")
    # sources first, supply before the pulses, so the parts keep the
    # generator's order of sources
    for k, (a, b) in enumerate(src_nodes):
        if k == 0:
            lines.append(f"VS{k} {name(a)} {name(b)} DC {float(src['DCV0'][0])!r}")
        else:
            args = " ".join(repr(float(src[p][k])) for p in ("V1", "V2", "TD", "TR", "TF", "PW", "PER"))
            lines.append(f"VS{k} {name(a)} {name(b)} PULSE({args})")
    lid = np.asarray(lids).reshape(-1, 4)
    for i in range(len(mi)):
        nodes = " ".join(name(x) for x in lid[i])
        extra = " ".join(f"{k}={float(v[i])!r}" for k, v in params.items())
        lines.append(f"M{i} {nodes} c{mi[i]} {extra}")
    for k, ((a, b), r) in enumerate(zip(np.asarray(r_nodes).reshape(-1, 2), R)):
        lines.append(f"R{k} {name(a)} {name(b)} {float(r)!r}")
    for k, ((a, b), c) in enumerate(zip(np.asarray(c_nodes).reshape(-1, 2), Cv)):
        lines.append(f"C{k} {name(a)} {name(b)} {float(c)!r}")
    lines.append(".tran 1p 1n")
    lines.append(".end")
    return "\n".join(lines) + "\n", (cards, mi, params, lid, R, Cv, r_nodes, c_nodes)


def test_logic_netlist_text_reads_back_exactly():
    text, (cards, mi, params, lid, R, Cv, r_nodes, c_nodes) = _logic_text()
    nl = Netlist(text)
    ids = np.array([-1 if n == "0" else int(n[1:]) for n in ["0"]] +
                   [int(n[1:]) for n in nl.node_names])
    to_gen = lambda a: np.where(np.asarray(a) < 0, -1, ids[1:][np.maximum(np.asarray(a), 0)])
    parts = nl.parts()
    assert [p["type"] for p in parts] == ["vsource", "vsource", "bsim4", "resistor", "capacitor"]
    m = parts[2]
    assert m["models"] == [(lev, dict(card)) for lev, card in cards]
    np.testing.assert_array_equal(m["model_index"], mi)
    for k, v in params.items():
        np.testing.assert_array_equal(m["params"][k], v)
    np.testing.assert_array_equal(to_gen(m["ext_nodes"]).reshape(-1, 4), lid)
    np.testing.assert_array_equal(parts[3]["params"]["R"], R)
    np.testing.assert_array_equal(to_gen(parts[3]["ext_nodes"]), np.asarray(r_nodes).reshape(-1))
    np.testing.assert_array_equal(parts[4]["params"]["C"], Cv)
    np.testing.assert_array_equal(to_gen(parts[4]["ext_nodes"]), np.asarray(c_nodes).reshape(-1))


RING17_SUBCKT = """17-stage ring oscillator through a subcircuit (the shape of the
* reference's ringosc_17stg.sp, with BSIM4 level-54 inverters)
.options abstol=1e-6 reltol=1e-6
.model nch nmos level=54
.model pch pmos level=54
vdd supply 0 dc=1.0
.subckt inv vin vout vdd gnd
Mp1 vout vin vdd vdd pch W=180n L=45n
Mn1 vout vin gnd gnd nch W=90n L=45n
.ends
""" + "".join(f"Xinv{k} {k} {k % 17 + 1} supply 0 inv\n" for k in range(1, 18)) + """Cl 1 0 1f
.ic 1=1
.tran 1p 1n
.print tran v(1)
.measure tran t1 when v(1)=0.5 cross=1
.end
"""


def test_subcircuit_ring_flattens():
    nl = Netlist(RING17_SUBCKT)
    names = nl.node_names
    assert names[0] == "SUPPLY" and set(names[1:]) == {str(k) for k in range(1, 18)}
    m, c, v = nl.parts()[1], nl.parts()[2], nl.parts()[0]
    assert v["type"] == "vsource" and m["type"] == "bsim4" and c["type"] == "capacitor"
    assert len(m["model_index"]) == 34
    # instance k's PMOS: drain = node k+1 (the output), gate = node k, source = bulk = supply
    ext = m["ext_nodes"].reshape(-1, 4)
    for k in range(17):
        inp, out = names.index(str(k + 1)), names.index(str((k + 1) % 17 + 1))
        assert list(ext[2 * k]) == [out, inp, 0, 0]            # Mp1
        assert list(ext[2 * k + 1]) == [out, inp, -1, -1]      # Mn1
    np.testing.assert_array_equal(m["model_index"], [1, 0] * 17)
    nodes, vals = nl.ic
    assert [names[i] for i in nodes] == ["1"] and list(vals) == [1.0]


def test_nested_subcircuits_and_internal_nodes():
    nl = Netlist("""nested
.subckt rdiv a b
R1 a mid 1k
R2 mid b 1k
.ends
.subckt two x y
Xa x m rdiv
Xb m y rdiv
.ends
X1 in 0 two
V1 in 0 1
""")
    # nodes in order of first appearance in the expanded lines
    assert nl.node_names == ["IN", "X1:XA:MID", "X1:M", "X1:XB:MID"]
    r = nl.part(0)
    assert r["type"] == "resistor"
    np.testing.assert_array_equal(r["ext_nodes"], [0, 1, 1, 2, 2, 3, 3, -1])
