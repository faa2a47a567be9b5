"""Seeded synthetic netlists for load-pass measurements on shared nodes.

* ``logic_dag``: a flat logic netlist -- a random DAG of INV / NAND2 / NOR2
  cells of BSIM4 transistors (two shared cards, two shared sizes), each net a
  driver node and a far-end node joined by a wire resistor, fanout 1-4, 20
  primary inputs.  ``logic_wires`` gives that netlist's wires: one resistor
  per net from the driver to the far end (R lognormal, median 50 ohm, sigma
  0.5) and one capacitor from the far end to ground (C uniform 0.5-5 fF).
* ``hb_diffamp``: the time samples of a harmonic-balance analysis of a
  differential amplifier: a 2k-transistor amplifier replicated once per time
  sample, every copy with its own nodes (the HB loader evaluates each
  sample's devices at that sample's voltages).  ``hb_filter`` gives each
  copy's output RC filter (a series resistor from each output of the last
  cell to a filter node, and a capacitor from that node to ground).

Each transistor netlist returns (models, model_index, params, node_lids,
n_nodes, voltages): instance-major external node LIDs (D, G, S, B; ground
-1) and one voltage per node, ready for :class:`xyce_amd.workloads.Workload`.
The wire functions return two-terminal device lists (node pairs, values);
the source functions return voltage-source lists (node pairs, instance
parameter columns for the "vsource" group).
"""
import numpy as np

SEED = 20261203
CARDS = [(54, {"TYPE": 1.0}), (54, {"TYPE": -1.0})]


def logic_dag(n_transistors=200_000, n_inputs=20, seed=SEED, **kw):
    return _logic(n_transistors, n_inputs, seed, **kw)[0]


def _logic(n_transistors, n_inputs, seed, vdd=1.0, window=2000, wn=90e-9, wp=180e-9, l=45e-9):
    """(transistor netlist, nets[k] = (driver, far end), primary-input nodes)"""
    rng = np.random.default_rng(np.random.SeedSequence(seed))
    VDD = 0
    nxt = 1
    pool = []           # (far-end node, remaining fanout)
    for _ in range(n_inputs):
        pool.append([nxt, int(rng.integers(1, 5))])
        nxt += 1
    mi, w, lids = [], [], []

    def fet(pmos, d, g, s):
        mi.append(1 if pmos else 0)
        w.append(wp if pmos else wn)
        lids.extend([d, g, s, VDD if pmos else -1])

    nets = []          # (driver node, far-end node) of every cell output
    count = 0
    while count < n_transistors:
        kind = int(rng.integers(0, 3))          # 0 INV, 1 NAND2, 2 NOR2
        k = 1 if kind == 0 else 2
        lo = max(0, len(pool) - window)
        ins = []
        for _ in range(k):
            j = int(rng.integers(lo, len(pool)))
            ins.append(pool[j][0])
            pool[j][1] -= 1
            if pool[j][1] == 0:
                pool[j] = pool[-1]
                pool.pop()
                lo = max(0, min(lo, len(pool) - 1))
        out, far = nxt, nxt + 1
        nxt += 2
        if kind == 0:
            fet(False, out, ins[0], -1)
            fet(True, out, ins[0], VDD)
        elif kind == 1:                          # NAND2: series NMOS, parallel PMOS
            x = nxt
            nxt += 1
            fet(False, out, ins[0], x)
            fet(False, x, ins[1], -1)
            fet(True, out, ins[0], VDD)
            fet(True, out, ins[1], VDD)
        else:                                    # NOR2: series PMOS, parallel NMOS
            y = nxt
            nxt += 1
            fet(True, y, ins[0], VDD)
            fet(True, out, ins[1], y)
            fet(False, out, ins[0], -1)
            fet(False, out, ins[1], -1)
        count += 2 if kind == 0 else 4
        pool.append([far, int(rng.integers(1, 5))])
        nets.append((out, far))
    n = len(mi)
    params = {"W": np.asarray(w), "L": np.full(n, l)}
    v = rng.uniform(0.0, vdd, nxt)
    v[VDD] = vdd
    fets = (CARDS, np.asarray(mi, np.int32), params, np.asarray(lids, np.int32), nxt, v)
    return fets, np.asarray(nets, np.int32).reshape(-1, 2), np.arange(1, 1 + n_inputs,
                                                                     dtype=np.int32)


def logic_wires(n_transistors=200_000, n_inputs=20, seed=SEED, **kw):
    """Wires of ``logic_dag`` with the same arguments: returns
    (r_nodes[n, 2], R[n], c_nodes[n, 2], C[n], input_nodes): resistor k from
    net k's driver to its far end, capacitor k from the far end to ground
    (-1), and the primary-input nodes.  The values come from their own seeded
    stream, so the transistors are those of ``logic_dag``."""
    _, nets, inputs = _logic(n_transistors, n_inputs, seed, **kw)
    rng = np.random.default_rng(np.random.SeedSequence([seed, 1]))
    R = 50.0 * np.exp(0.5 * rng.standard_normal(len(nets)))
    C = rng.uniform(0.5e-15, 5e-15, len(nets))
    c_nodes = np.stack([nets[:, 1], np.full(len(nets), -1, np.int32)], axis=1)
    return nets, R, c_nodes, C, inputs


def hb_diffamp(samples=128, n_transistors=2000, seed=SEED, **kw):
    return _diffamp(samples, n_transistors, seed, **kw)[0]


def _diffamp(samples, n_transistors, seed, vdd=1.0, l=45e-9):
    """(transistor netlist, outputs[s] = the last cell's two outputs of copy s).
    One copy per time sample of an amplifier of n_transistors/5 cells.
    Cell: differential pair (NMOS, gates on the two inputs), PMOS
    current-mirror load, NMOS tail source gated by a shared bias node; cell
    outputs feed the next cell's inputs.  Copy s sees the inputs of sample s
    of a two-tone (1.00/1.01 GHz, 10 mV) drive around mid-rail."""
    rng = np.random.default_rng(np.random.SeedSequence(seed))
    cells = n_transistors // 5
    mi, w, lids, volts, outs = [], [], [], [], []
    nxt = 0
    t = np.arange(samples) / samples * 100e-9        # 100 ns: one beat period
    drive = 0.01 * (np.sin(2 * np.pi * 1.00e9 * t) + np.sin(2 * np.pi * 1.01e9 * t))
    for s in range(samples):
        VDD, BIAS, INP, INN = nxt, nxt + 1, nxt + 2, nxt + 3
        nxt += 4
        node_v = [vdd, 0.45, 0.5 + drive[s], 0.5 - drive[s]]
        a, b = INP, INN
        for c in range(cells):
            tail, outp, outn = nxt, nxt + 1, nxt + 2
            nxt += 3
            node_v += [0.05 + 0.01 * rng.standard_normal(),
                       0.6 + 0.05 * rng.standard_normal(), 0.6 + 0.05 * rng.standard_normal()]
            for pmos, d, g, src, width in ((False, outn, a, tail, 180e-9),
                                           (False, outp, b, tail, 180e-9),
                                           (True, outn, outn, VDD, 360e-9),
                                           (True, outp, outn, VDD, 360e-9),
                                           (False, tail, BIAS, -1, 360e-9)):
                mi.append(1 if pmos else 0)
                w.append(width)
                lids.extend([d, g, src, VDD if pmos else -1])
            a, b = outp, outn
        volts.extend(node_v)
        outs.append((a, b))
    n = len(mi)
    params = {"W": np.asarray(w), "L": np.full(n, l)}
    fets = (CARDS, np.asarray(mi, np.int32), params, np.asarray(lids, np.int32), nxt,
            np.asarray(volts))
    return fets, np.asarray(outs, np.int32).reshape(-1, 2)


def hb_filter(samples=128, n_transistors=2000, seed=SEED, r=1000.0, c=50e-15, **kw):
    """Output RC filter of every copy of ``hb_diffamp`` with the same
    arguments: returns (r_nodes, R, c_nodes, C, n_nodes, voltages) -- per
    copy a resistor from each of the last cell's two outputs to a new filter
    node, and a capacitor from each filter node to ground; the node count and
    voltages cover the amplifier's nodes followed by the filter nodes."""
    (_, _, _, _, n0, v0), outs = _diffamp(samples, n_transistors, seed, **kw)
    k = 2 * len(outs)
    filt = n0 + np.arange(k, dtype=np.int32)
    r_nodes = np.stack([outs.reshape(-1), filt], axis=1).astype(np.int32)
    c_nodes = np.stack([filt, np.full(k, -1, np.int32)], axis=1)
    volts = np.concatenate([v0, v0[outs.reshape(-1)]])
    return r_nodes, np.full(k, r), c_nodes, np.full(k, c), n0 + k, volts


def logic_sources(n_inputs=20, vdd=1.0, period=1e-9, edge=50e-12):
    """Sources of ``logic_dag``: the supply (node 0, DC vdd) and one PULSE
    per primary input (nodes 1..n_inputs): 0 -> vdd at 1 GHz (PER = 1 ns),
    edges of 50 ps, pulse width half a period less an edge, staggered delays.
    Returns (nodes[n, 2], params)."""
    nodes = np.array([[0, -1]] + [[k, -1] for k in range(1, n_inputs + 1)], np.int32)
    n = len(nodes)
    kind = np.full(n, 2.0)
    kind[0] = np.nan                      # the supply: DC only
    params = {"DCV0": np.concatenate([[vdd], np.zeros(n_inputs)]),
              "TRANSIENTSOURCETYPE": kind,
              "V1": np.zeros(n), "V2": np.full(n, vdd),
              "TD": np.concatenate([[0.0], np.arange(n_inputs) * period / n_inputs]),
              "TR": np.full(n, edge), "TF": np.full(n, edge),
              "PW": np.full(n, period / 2 - edge), "PER": np.full(n, period)}
    return nodes, params


def hb_sources(samples=128, n_transistors=2000, seed=SEED, **kw):
    """Sources of every sample copy of ``hb_diffamp``: VDD, the bias and the
    two inputs, each a DC source at that sample's value (the HB loader hands
    each sample's devices the sources' values at that sample's time).
    Returns (nodes[n, 2], params)."""
    (_, _, _, _, _, volts), _ = _diffamp(samples, n_transistors, seed, **kw)
    cells = n_transistors // 5
    stride = 4 + 3 * cells                # nodes per copy
    base = np.arange(samples, dtype=np.int32) * stride
    nodes = np.stack([np.repeat(base, 4) + np.tile(np.arange(4, dtype=np.int32), samples),
                      np.full(4 * samples, -1, np.int32)], axis=1)
    return nodes, {"DCV0": volts[nodes[:, 0]]}


def diffamp_netlist(n_transistors=2000, seed=SEED, r=1000.0, c=50e-15, vdd=1.0, bias=0.45,
                    vcm=0.5, l=45e-9):
    """The differential amplifier of ``hb_diffamp`` (one copy) with its
    output RC filter and its sources, as SPICE text for the netlist reader.
    Nodes are named after the generator's ids (``n<id>``; the filter nodes
    follow the amplifier's), so a reader node maps back to the id by its
    name.  The sources are DC: VDD, VBIAS, and the inputs VINP and VINN at
    the common-mode level, ready for per-sample values (`.STEP` lines).
    Returns (text, initial voltages by generator id)."""
    (cards, mi, params, lids, n0, volts), outs = _diffamp(1, n_transistors, seed, vdd=vdd, l=l)
    lid = np.asarray(lids).reshape(-1, 4)
    name = lambda k: "0" if k < 0 else f"n{int(k)}"
    lines = [f"differential amplifier, {len(mi)} BSIM4 transistors, RC filter"]
    for j, (lev, card) in enumerate(cards):
        kind = "nmos" if card["TYPE"] > 0 else "pmos"
        lines.append(f".model c{j} {kind} level={lev}")
    for i in range(len(mi)):
        extra = " ".join(f"{k}={float(v[i])!r}" for k, v in params.items())
        lines.append(f"M{i} {' '.join(name(x) for x in lid[i])} c{mi[i]} {extra}")
    a, b = (int(x) for x in outs[0])
    lines += [f"RF1 {name(a)} {name(n0)} {float(r)!r}", f"RF2 {name(b)} {name(n0 + 1)} {float(r)!r}",
              f"CF1 {name(n0)} 0 {float(c)!r}", f"CF2 {name(n0 + 1)} 0 {float(c)!r}",
              f"VDD {name(0)} 0 DC {float(vdd)!r}", f"VBIAS {name(1)} 0 DC {float(bias)!r}",
              f"VINP {name(2)} 0 DC {float(vcm)!r}", f"VINN {name(3)} 0 DC {float(vcm)!r}"]
    v0 = np.concatenate([volts, volts[[a, b]]])
    v0[2] = v0[3] = vcm
    return "\n".join(lines) + "\n", v0


def logic_netlist(n_transistors=200_000, n_inputs=20, seed=SEED, t_step=1e-12, t_stop=1e-9,
                  **kw):
    """The flat logic netlist (``logic_dag`` with ``logic_wires`` and
    ``logic_sources``) as SPICE text for the netlist reader: nodes named
    after the generator's ids (``n<id>``), the supply and the PULSE inputs
    first, then the transistors, the wire resistors and capacitors, and the
    transient.  Values print exactly.  Returns (text, voltages by generator
    id)."""
    cards, mi, params, lids, n_nodes, volts = logic_dag(n_transistors, n_inputs, seed, **kw)
    r_nodes, R, c_nodes, Cv, _ = logic_wires(n_transistors, n_inputs, seed, **kw)
    src_nodes, src = logic_sources(n_inputs)
    name = lambda k: "0" if k < 0 else f"n{int(k)}"
    lines = [f"flat logic netlist, {len(mi)} BSIM4 transistors"]
    for j, (lev, card) in enumerate(cards):
        kind = "nmos" if card["TYPE"] > 0 else "pmos"
        lines.append(f".model c{j} {kind} level={lev}")
    for k, (a, b) in enumerate(src_nodes):
        if k == 0:
            lines.append(f"VS{k} {name(a)} {name(b)} DC {float(src['DCV0'][0])!r}")
        else:
            args = " ".join(repr(float(src[p][k]))
                            for p in ("V1", "V2", "TD", "TR", "TF", "PW", "PER"))
            lines.append(f"VS{k} {name(a)} {name(b)} PULSE({args})")
    lid = np.asarray(lids).reshape(-1, 4)
    keys = list(params.keys())
    for i in range(len(mi)):
        extra = " ".join(f"{k}={float(params[k][i])!r}" for k in keys)
        lines.append(f"M{i} {' '.join(name(x) for x in lid[i])} c{mi[i]} {extra}")
    for k, ((a, b), r) in enumerate(zip(np.asarray(r_nodes).reshape(-1, 2), R)):
        lines.append(f"R{k} {name(a)} {name(b)} {float(r)!r}")
    for k, ((a, b), c) in enumerate(zip(np.asarray(c_nodes).reshape(-1, 2), Cv)):
        lines.append(f"C{k} {name(a)} {name(b)} {float(c)!r}")
    lines.append(f".tran {float(t_step)!r} {float(t_stop)!r}")
    lines.append(".end")
    return "\n".join(lines) + "\n", np.asarray(volts)
