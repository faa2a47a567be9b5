"""GPU parity of the resistor and capacitor loads against the oracle (through
the C ABI).

Bar: bit-identical.  The loads are products and sums of inputs and of the
per-instance G or C (computed at setup with the same operations as the
oracle), added into each entry in the reference's order, so every F/Q/lead/
state/CSR entry must equal the oracle's bits, across operating-point and
transient iterations, private, shared and grounded topologies, IC= branches
and lead currents; the fused one-call load equals the four calls.  Rows of
more than 1024 terms (supply rails) are the exception: they are summed in a
fixed chunked order and meet the 1e-12 bar instead (test_rails)."""
import numpy as np
import pytest

from gpu_common import ITERS, run_parity, twin_drive
from harness import Pair
from specs import capacitor_spec, resistor_spec

pytestmark = pytest.mark.gpu


def _all_exact(summary):
    for k, (n, exact, _, mx) in summary.items():
        assert exact == n, f"{k}: {n - exact} of {n} entries differ (max rel {mx:.2e})"


@pytest.mark.parametrize("topology,sheet", [("private", False), ("shared", False),
                                            ("shared", True)])
def test_resistor_parity(topology, sheet):
    n = 20000
    spec = resistor_spec(n, seed=501, temp_spread=True, sheet=sheet)
    lead = (np.arange(n) % 3 == 0).astype(np.int32)
    p = Pair("resistor", spec, topology=topology, seed=502, lead=lead, ground_frac=0.05)
    print(f"resistor {topology} sheet={sheet}", flush=True)
    _all_exact(run_parity(p, np.random.default_rng(503)))


@pytest.mark.parametrize("mode", ["plain", "ic", "semi", "age"])
def test_capacitor_parity(mode):
    n = 20000
    spec = capacitor_spec(n, seed=504, temp_spread=True, ic=mode == "ic", semi=mode == "semi",
                          age=mode == "age")
    lead = (np.arange(n) % 4 == 0).astype(np.int32)
    p = Pair("capacitor", spec, topology="shared", seed=505, lead=lead, ground_frac=0.05)
    print(f"capacitor {mode}", flush=True)
    _all_exact(run_parity(p, np.random.default_rng(506)))


@pytest.mark.parametrize("model", ["resistor", "capacitor-ic"])
def test_fused_equals_four_calls(model):
    n = 5000
    if model == "resistor":
        name, spec = "resistor", resistor_spec(n, seed=507)
    else:
        name, spec = "capacitor", capacitor_spec(n, seed=507, ic=True)
    lead = (np.arange(n) % 5 == 0).astype(np.int32)
    twin_drive(name, spec, 508, np.random.default_rng(509), True, topology="blocked", lead=lead,
               iters=ITERS)


def test_rails():
    """Rails: every resistor and capacitor between two of 3 nodes, so each row
    sums thousands of terms.  Rows of more than 1024 terms are summed as
    256-term chunks in order and then a fixed pairwise tree over the chunk
    sums: the same bits on every run, within the 1e-12 bar of the oracle's
    sequential sum (run_parity asserts it), but not its bits."""
    n = 6000
    for model, spec, seed in (("resistor", resistor_spec(n, seed=510), 511),
                              ("capacitor", capacitor_spec(n, seed=513), 514)):
        p = Pair(model, spec, topology="shared", seed=seed, n_shared_nodes=3)
        summary = run_parity(p, np.random.default_rng(seed + 1))
        assert max(v[3] for v in summary.values()) <= 1e-13
