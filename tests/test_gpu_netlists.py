"""Full-size GPU parity of the shared-node netlists against the oracle:

* the flat logic netlist: 200k BSIM4 transistors in a random DAG of
  INV/NAND2/NOR2 cells, with a resistor and a capacitor per net (the wire
  RC), the supply and 20 PULSE inputs at 1 GHz; its VDD row collects over 5*10^5 dF/dx terms, so it runs through the
  chunked rail sum;
* the harmonic-balance load: 128 time samples of a 2000-transistor
  differential amplifier with its output RC filter and sources (256k BSIM4 instances
  in one launch, as the HB loader hands every sample's devices to the load,
  N_LOA_HBLoader.C:618-710).

Each is loaded as four device types on one set of vectors, in DeviceMgr's
order (N_DEV_DeviceMgr.C:2053-2502), over a DC operating-point iteration
with junction initialisation, the first transient iteration and a later one.
Bar: assert_parity (1e-12 of each entry's sum of |contributions|, plus its
libm-noise spread) on every output.
"""
import numpy as np
import pytest

from circuit import CircuitPair
from harness import assert_parity
from netlist_circuits import hb_circuit, logic_circuit

pytestmark = pytest.mark.gpu

PDT = 1e12   # 1 ps steps
TIMES = dict(startingTimeStep=1e-12, finalTime=1e-9, bpTol=1e-18)
ITERS = [dict(newtonIter=0, dcopFlag=1, initJctFlag=1, **TIMES),
         dict(newtonIter=0, transientFlag=1, initTranFlag=1, pdt=PDT, currTime=1e-12, **TIMES),
         dict(newtonIter=1, transientFlag=1, pdt=PDT, currTime=3.7e-10, **TIMES)]


def _drive(cp, volts, rng, vdd=1.0):
    cp.set_inputs(currSto=rng.uniform(0.0, vdd, cp.n_sto),
                  currSta=rng.uniform(-1e-15, 1e-15, cp.n_sta))
    base = cp.unknowns(volts)
    extra = np.arange(cp.n_nodes) >= len(volts)     # branch currents of the sources
    for it, flags in enumerate(ITERS):
        x = base.copy() if it == 0 else np.clip(base + rng.normal(0.0, 0.05, cp.n_nodes), 0.0, vdd)
        x[extra] = rng.normal(0.0, 1e-4, int(extra.sum()))
        cp.set_inputs(nextSol=x)
        print(f"  iteration {it}: {flags}", flush=True)
        ro, rg = cp.load(**flags)
        for k in ro:
            st = assert_parity(rg[k], ro[k], what=f"iter {it} {k}", scale=cp.mag.get(k),
                               noise=cp.noise.get(k))
            print(f"    {k:12s} n {st['n']:8d} bit-exact {st['bitexact']:.4f} "
                  f"outliers {st['outliers']} max rel {st['max_rel']:.2e}", flush=True)


def test_logic_netlist_full_size():
    parts, n_nodes, volts = logic_circuit()
    print(f"logic netlist: {parts[0].n} BSIM4, {parts[1].n} R, {parts[2].n} C, "
          f"{parts[3].n} V, {n_nodes} nodes", flush=True)
    cp = CircuitPair(parts, n_nodes, noise_runs=4)
    vdd_terms = np.diff(cp.row_ptr)[0]
    print(f"  VDD row: {vdd_terms} CSR slots", flush=True)
    _drive(cp, volts, np.random.default_rng(71))


def test_hb_samples_full_size():
    parts, n_nodes, volts = hb_circuit()
    print(f"HB samples: {parts[0].n} BSIM4, {parts[1].n} R, {parts[2].n} C, "
          f"{n_nodes} nodes", flush=True)
    cp = CircuitPair(parts, n_nodes, noise_runs=4)
    _drive(cp, volts, np.random.default_rng(72))
