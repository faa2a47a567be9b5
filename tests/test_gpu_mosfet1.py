"""GPU parity of the Level-1 MOSFET load against the oracle, through the C ABI.

Bar: harness.assert_parity (1e-12 of each entry's error scale, or twice its
libm noise; at most 0.05% outliers, none beyond 1e-6).  Covered: the parity
cards (internal drain/source nodes, NSUB-derived parameters, explicit and
area-scaled junction capacitances), lead currents, Newton iterations of DC and
transient analysis, private and shared topologies with ground nodes, limiter
on and off, and the Meyer charges carried from the previous step.  Fused one
call equals the four calls; two runs are bit-identical.
"""
import numpy as np
import pytest

from gpu_common import NOISE_RUNS, run_parity, twin_drive
from harness import Pair
from specs import MOSFET1_CARDS, mosfet1_spec

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("topology", ["private", "shared"])
def test_mosfet1_parity(topology):
    n = 20000
    spec = mosfet1_spec(n, seed=301, temp_spread=True)
    lead = (np.arange(n) % 4 == 0).astype(np.int32)
    print(f"mosfet1 parity {topology} n={n}", flush=True)
    p = Pair("mosfet1", spec, topology=topology, seed=302, lead=lead, ground_frac=0.05,
             noise_runs=NOISE_RUNS)
    run_parity(p, np.random.default_rng(303))


def test_mosfet1_parity_limiter_off_plain():
    import oracle
    n = 8000
    spec = mosfet1_spec(n, seed=304, geometry=False)
    print(f"mosfet1 limiter off n={n}", flush=True)
    p = Pair("mosfet1", spec, options=oracle.default_options(voltageLimiterFlag=0), seed=305,
             noise_runs=NOISE_RUNS)
    run_parity(p, np.random.default_rng(306))


@pytest.mark.parametrize("topology", ["private", "blocked"])
def test_mosfet1_fused_equals_split(topology):
    spec = mosfet1_spec(6000, seed=307)
    print(f"mosfet1 fused vs split {topology} n=6000", flush=True)
    lead = (np.arange(6000) % 3 == 0).astype(np.int32)
    twin_drive("mosfet1", spec, 308, np.random.default_rng(309), fused_b=True, topology=topology, lead=lead)


def test_mosfet1_run_to_run_bit_identical():
    spec = mosfet1_spec(6000, seed=310)
    print("mosfet1 determinism n=6000", flush=True)
    twin_drive("mosfet1", spec, 311, np.random.default_rng(312), fused_b=False, topology="shared")
