"""GPU parity of the BSIM4 option groups (the option kernel and its stamp
lists) against the oracle, through the C ABI.

Each option card -- gate tunnelling (igcMod 1/2, igbMod), bias-dependent
source/drain resistance (rdsMod), the capMod 0 and 1 charge models under
each charge partition, the gate resistance networks (rgateMod 1-3), the body
resistance networks (rbodyMod 1/2), NQS (trnqsMod, acnqsMod) -- is loaded
beside a default card in one group (so the group takes the option kernel for
both) over DC, transient-OP and transient Newton iterations, private and
shared topologies with ground.  Bars: the production library against the
oracle by harness.assert_parity (as the other models), and the parity builds
(shared portable exp/log/pow) bit for bit on every output.  Initial
conditions (IC1-IC3) are loaded in DC and transient calls.
"""
import numpy as np
import pytest

from gpu_common import NOISE_RUNS, run_parity, twin_drive
from harness import Pair
from specs import BSIM4_CARDS, BSIM4_OPT_CARDS, bsim4_spec
from test_gpu_bitexact import run_exact

pytestmark = pytest.mark.gpu
NAMES = list(BSIM4_OPT_CARDS)


def _spec(name, n, seed, **kw):
    return bsim4_spec(n, seed=seed, cards=[BSIM4_OPT_CARDS[name], BSIM4_CARDS[0]], **kw)


@pytest.mark.parametrize("name", NAMES)
def test_option_parity(name):
    n = 1500
    spec = _spec(name, n, 301 + NAMES.index(name), geometry=True, nf_mix=True, temp_spread=True)
    print(f"bsim4 option {name} parity n={n}", flush=True)
    p = Pair("bsim4", spec, topology="shared", seed=302, ground_frac=0.05, noise_runs=NOISE_RUNS)
    run_parity(p, np.random.default_rng(303))


@pytest.mark.parametrize("name", NAMES)
def test_option_bitexact(name):
    n = 1500
    spec = _spec(name, n, 311 + NAMES.index(name), geometry=True, nf_mix=True, temp_spread=True)
    print(f"bsim4 option {name} bit-exact n={n}", flush=True)
    p = Pair("bsim4", spec, variant="pm", topology="private", seed=312, ground_frac=0.05)
    run_exact(p, np.random.default_rng(313))


def test_all_options_one_group_bitexact():
    """Every option card in one group, plain instances, limiter off."""
    import oracle
    spec = bsim4_spec(3000, seed=321, cards=list(BSIM4_OPT_CARDS.values()) + BSIM4_CARDS[:2],
                      geometry=False)
    print("bsim4 all option cards n=3000", flush=True)
    p = Pair("bsim4", spec, variant="pm", topology="shared", seed=322,
             options=oracle.default_options(voltageLimiterFlag=0))
    run_exact(p, np.random.default_rng(323))


def _ic_spec(n, seed):
    spec = bsim4_spec(n, seed=seed, cards=[BSIM4_OPT_CARDS["igc1"], BSIM4_CARDS[1]], geometry=False)
    rng = np.random.default_rng(seed + 1)
    spec["inst_params"].update({"IC1": rng.uniform(0.0, 1.0, n), "IC2": rng.uniform(0.0, 1.0, n),
                                "IC3": rng.uniform(-0.3, 0.0, n)})
    return spec


def test_initial_conditions_bitexact():
    """IC1/IC2/IC3: the branch-current unknowns and their equations in DC,
    the unit branch diagonal otherwise."""
    spec = _ic_spec(1200, 331)
    print("bsim4 initial conditions n=1200", flush=True)
    p = Pair("bsim4", spec, variant="pm", topology="private", seed=332)
    run_exact(p, np.random.default_rng(333))


def test_initial_conditions_parity():
    spec = _ic_spec(1200, 341)
    print("bsim4 initial conditions parity n=1200", flush=True)
    p = Pair("bsim4", spec, topology="private", seed=342, noise_runs=NOISE_RUNS)
    run_parity(p, np.random.default_rng(343))


def test_option_group_run_to_run_bit_identical():
    """Two groups of the same option circuit driven alike give the same bits."""
    spec = bsim4_spec(2000, seed=351, cards=list(BSIM4_OPT_CARDS.values()), geometry=True)
    print("bsim4 option group twin n=2000", flush=True)
    twin_drive("bsim4", spec, 352, np.random.default_rng(353), fused_b=False, topology="shared")
