"""Circuits built by the netlist reader on the GPU (xa_circuit_build /
xa_circuit_load) against the harness circuits built by hand from the same
parsed parts, and through them against the oracle.

* The builder's LIDs, state/store slots and CSR are the harness's
  convention (netlist nodes, then each part's internal nodes instance by
  instance), so the CSR must be identical and every output of a circuit
  load bit-identical to the harness's GPU load of the same groups
  (DeviceMgr's order over device types, N_DEV_DeviceMgr.C:2053-2502).
* The harness's GPU load meets the parity bar against the oracle
  (assert_parity: 1e-12 of each entry's sum of |contributions| plus its
  libm-noise spread).
* A .STEP netlist built batched (every step's copy of the circuit in one
  set of groups) gives, copy by copy, exactly the outputs of the circuit
  built for that step alone.
"""
import numpy as np
import pytest

from circuit import CircuitPair, Part
from harness import OUT_VECS, STATE_VECS, assert_parity

pytestmark = pytest.mark.gpu

PDT = 1e12
TIMES = dict(startingTimeStep=1e-12, finalTime=1e-9, bpTol=1e-18)
ITERS = [dict(newtonIter=0, dcopFlag=1, initJctFlag=1, **TIMES),
         dict(newtonIter=0, transientFlag=1, initTranFlag=1, pdt=PDT, currTime=1e-12, **TIMES),
         dict(newtonIter=1, transientFlag=1, pdt=PDT, currTime=3.7e-10, **TIMES)]
KEYS = OUT_VECS + STATE_VECS + ("dFdx", "dQdx")


def _parts(nl, step=-1):
    return [Part(p["type"], p["models"], p["model_index"], p["params"], p["ext_nodes"])
            for p in nl.parts(step)]


class CircuitRun:
    """A reader-built circuit with its own device vectors."""

    def __init__(self, ctx, nl, step=-1):
        import torch
        import xyce_amd
        from xyce_amd.netlist_reader import Circuit
        self.torch, self.xa = torch, xyce_amd
        self.c = Circuit(ctx, nl, step=step)
        n, ns, nt, nz = self.c.n_unknowns, self.c.n_sta, self.c.n_sto, self.c.nnz
        size = dict(nextSol=n, currSol=n, flagSol=n, nextSta=ns, currSta=ns, lastSta=ns,
                    nextStaDeriv=ns, nextSto=nt, currSto=nt, f=n, q=n, b=n, dFdxdVp=n,
                    dQdxdVp=n, storeLeadF=nt, storeLeadQ=nt, dFdx=nz, dQdx=nz)
        self.size = size
        self.d = {k: torch.zeros(max(v, 1), dtype=torch.float64, device="cuda")
                  for k, v in size.items()}

    def load(self, inputs, flags):
        for k, a in inputs.items():
            self.d[k].zero_()
            if len(a):
                self.d[k][:len(a)] = self.torch.from_numpy(np.ascontiguousarray(a)).cuda()
        for k in OUT_VECS + ("dFdx", "dQdx", "b"):
            self.d[k].zero_()
        ss = self.xa.default_solver_state(gainScale=1.0, nltermScale=1.0, **flags)
        self.c.load(ss, **self.d)
        self.torch.cuda.synchronize()
        return {k: self.d[k][:self.size[k]].cpu().numpy().copy() for k in KEYS}


def _inputs(n, ns, nt, volts, rng, it, vdd=1.0):
    x = np.full(n, 0.0)
    x[:len(volts)] = volts if it == 0 else np.clip(volts + rng.normal(0, 0.05, len(volts)), 0, vdd)
    x[len(volts):] = rng.normal(0.0, 1e-4, n - len(volts))
    return dict(nextSol=x, currSol=x * 0.99,
                currSto=rng.uniform(0.0, vdd, nt), currSta=rng.uniform(-1e-15, 1e-15, ns),
                lastSta=rng.uniform(-1e-15, 1e-15, ns), nextStaDeriv=np.zeros(ns),
                nextSta=np.zeros(ns), nextSto=np.zeros(nt), b=np.zeros(n))


def _check_against_pair(nl, seed, parity=True):
    import xyce_amd
    cp = CircuitPair(_parts(nl), nl.n_nodes, noise_runs=2 if parity else 0, gpu=True)
    run = CircuitRun(cp.ctx, nl)
    assert run.c.n_unknowns == cp.n_nodes and run.c.n_sta == cp.n_sta and run.c.n_sto == cp.n_sto
    rp, ci = run.c.csr()
    np.testing.assert_array_equal(rp, cp.row_ptr)
    np.testing.assert_array_equal(ci, cp.col_idx)
    rng = np.random.default_rng(seed)
    volts = rng.uniform(0.0, 1.0, nl.n_nodes)
    for it, flags in enumerate(ITERS):
        inp = _inputs(cp.n_nodes, cp.n_sta, cp.n_sto, volts, rng, it)
        print(f"  iteration {it}: {flags}", flush=True)
        cp.set_inputs(**inp)
        rc = run.load(inp, flags)
        ro, rg = cp.load(**flags)
        for k in KEYS:
            np.testing.assert_array_equal(rc[k], rg[k], err_msg=f"iter {it} {k}: reader vs harness")
            if parity and k in ro:
                assert_parity(rg[k], ro[k], what=f"iter {it} {k}", scale=cp.mag.get(k),
                              noise=cp.noise.get(k))
    del xyce_amd
    return cp, run


def test_ring_netlist_circuit():
    from xyce_amd.netlist_reader import Netlist
    from xyce_amd.ring import Ring
    ring = Ring(stages=51)
    nl = Netlist(ring.netlist())
    assert nl.n_nodes == 52 and [p["type"] for p in nl.parts()] == ["bsim4", "capacitor",
                                                                    "vsource"]
    print("ring netlist: 102 BSIM4, 51 C, 1 V", flush=True)
    _check_against_pair(nl, 81)


def _mixed_text(rng, n_nodes=40, n_each=60):
    nodes = lambda k: " ".join(f"n{int(x)}" if x >= 0 else "0"
                               for x in rng.integers(-1, n_nodes, k))
    L = ["every device type",
         ".model nch nmos level=54", ".model pch pmos level=54 toxe=1.8e-9",
         ".model m1n nmos level=1 vto=0.7 kp=1e-4 cgso=1e-10 cgdo=1e-10 cbd=1e-15 cbs=1e-15",
         ".model m1p pmos level=1 vto=-0.7 kp=5e-5 lambda=0.02",
         ".model dm d is=1e-14 rs=10 cjo=1e-12 tt=1e-11",
         ".model qn npn bf=100 is=1e-16 rb=10 rc=5 re=1 cje=1e-12 cjc=5e-13 tf=1e-11 vaf=50",
         ".model qp pnp bf=50 is=1e-16",
         "VDD n0 0 DC 1",
         "VIN n1 0 PULSE(0 1 1e-11 5e-11 5e-11 4.5e-10 1e-9)",
         "VS n2 0 SIN(0.5 0.5 1e9 0 0 0)"]
    for i in range(n_each):
        L.append(f"MB{i} {nodes(4)} {'nch' if i % 2 == 0 else 'pch'} "
                 f"W={float(rng.uniform(90e-9, 1e-6))!r} L={float(rng.uniform(45e-9, 1e-7))!r}")
        L.append(f"MA{i} {nodes(4)} {'m1n' if i % 3 else 'm1p'} "
                 f"W={float(rng.uniform(1e-6, 1e-5))!r} L={float(rng.uniform(1e-6, 2e-6))!r}")
        L.append(f"D{i} {nodes(2)} dm {float(rng.uniform(0.5, 2))!r}")
        L.append(f"Q{i} {nodes(3)} {'qn' if i % 4 else 'qp'}")
        L.append(f"R{i} {nodes(2)} {float(rng.uniform(10, 1e4))!r}")
        L.append(f"C{i} {nodes(2)} {float(rng.uniform(1e-15, 1e-13))!r}")
    L.append(".tran 1e-12 1e-9")
    return "\n".join(L) + "\n"


def test_mixed_netlist_circuit():
    from xyce_amd.netlist_reader import Netlist
    nl = Netlist(_mixed_text(np.random.default_rng(82)))
    types = [p["type"] for p in nl.parts()]
    print(f"mixed netlist: parts {types}, {nl.n_nodes} nodes", flush=True)
    assert set(types) == {"vsource", "bsim4", "mosfet1", "diode", "bjt", "resistor", "capacitor"}
    _check_against_pair(nl, 83)


STEP_TEXT = """stepped inverter chain with a diode clamp
.model nch nmos level=54
.model pch pmos level=54
.model dm d is=1e-14 cjo=1e-12
MN1 b a 0 0 nch W=9e-08 L=4.5e-08
MP1 b a vdd vdd pch W=1.8e-07 L=4.5e-08
MN2 c b 0 0 nch W=9e-08 L=4.5e-08
MP2 c b vdd vdd pch W=1.8e-07 L=4.5e-08
D1 c vdd dm
R1 c out 1000
C1 out 0 1e-15
VDD vdd 0 DC 1
VIN a 0 PULSE(0 1 1e-11 5e-11 5e-11 4.5e-10 1e-9)
.STEP nch:TOXE LIST 1.6e-9 1.8e-9 2e-9 2.2e-9
.STEP R1:R LIST 500 1000 2000 4000
.tran 1e-12 1e-9
"""


def test_step_batched_equals_each_step():
    import xyce_amd
    from xyce_amd.netlist_reader import Netlist
    nl = Netlist(STEP_TEXT)
    ctx = xyce_amd.Context(0)
    batched = CircuitRun(ctx, nl, step=-2)
    singles = [CircuitRun(ctx, nl, step=s) for s in range(nl.n_steps)]
    assert batched.c.n_copies == nl.n_steps == 4
    U, S, T, Z = (singles[0].c.n_unknowns, singles[0].c.n_sta, singles[0].c.n_sto,
                  singles[0].c.nnz)
    assert (batched.c.n_unknowns, batched.c.n_sta, batched.c.n_sto, batched.c.nnz) == \
        (4 * U, 4 * S, 4 * T, 4 * Z)
    rp, ci = batched.c.csr()
    rp1, ci1 = singles[0].c.csr()
    for c in range(4):           # block diagonal: copy c's block is the single circuit's
        np.testing.assert_array_equal(rp[c * U:(c + 1) * U + 1] - rp[c * U], rp1)
        np.testing.assert_array_equal(ci[rp[c * U]:rp[(c + 1) * U]] - c * U, ci1)
    rng = np.random.default_rng(84)
    volts = [rng.uniform(0.0, 1.0, nl.n_nodes) for _ in range(4)]
    per = {"nextSol": U, "currSol": U, "b": U, "currSto": T, "nextSto": T, "currSta": S,
           "lastSta": S, "nextSta": S, "nextStaDeriv": S}
    for it, flags in enumerate(ITERS):
        inps = [_inputs(U, S, T, volts[c], rng, it) for c in range(4)]
        outs = [singles[c].load(inps[c], flags) for c in range(4)]
        big = {k: np.concatenate([inps[c][k] for c in range(4)]) for k in per}
        ob = batched.load(big, flags)
        size = {k: len(outs[0][k]) for k in KEYS}
        for k in KEYS:
            for c in range(4):
                np.testing.assert_array_equal(ob[k][c * size[k]:(c + 1) * size[k]], outs[c][k],
                                              err_msg=f"iter {it} {k} copy {c}")
        # the copies differ (the sweep reaches the loads)
        assert not np.array_equal(outs[0]["f"], outs[3]["f"])


def test_subcircuit_ring_circuit():
    """The 17-stage ring written with an inverter subcircuit (the shape of
    the reference's ringosc_17stg.sp): flattened, built and loaded like the
    flat netlists."""
    from test_netlist_reader import RING17_SUBCKT
    from xyce_amd.netlist_reader import Netlist
    nl = Netlist(RING17_SUBCKT)
    print(f"subcircuit ring: parts {[p['type'] for p in nl.parts()]}", flush=True)
    _check_against_pair(nl, 85)
