"""GPU parity of the BSIM4 (level 54) load against the oracle, through the C ABI.

Bar (harness.assert_parity): every F/Q/limiter/store/state/CSR entry within
1e-12 of its error scale (the larger of the two values and the sum of
|contributions| assembled into it) or within twice its libm noise (its spread
when the oracle's exp/log/pow results move by up to 2 ulps); at most 0.05% of
entries outside that, none beyond 1e-6 of its scale.  Covered: the parity
cards, Newton iterations of DC and transient analysis (initJct, initTran),
private and shared topologies with ground nodes.  Assembly is bit-identical
from run to run and the fused one-call load equals the four calls.
"""
import numpy as np
import pytest

from gpu_common import NOISE_RUNS, run_parity, twin_drive
from harness import Pair
from specs import BSIM4_CARDS, BSIM4_VTL_CARD, bsim4_spec

pytestmark = pytest.mark.gpu
ALL_CARDS = BSIM4_CARDS + [BSIM4_VTL_CARD]


@pytest.mark.parametrize("topology", ["private", "shared"])
def test_bsim4_parity(topology):
    n = 6000
    spec = bsim4_spec(n, seed=201, cards=ALL_CARDS, geometry=True, nf_mix=True, temp_spread=True)
    print(f"bsim4 parity {topology} n={n}", flush=True)
    p = Pair("bsim4", spec, topology=topology, seed=202, ground_frac=0.05, noise_runs=NOISE_RUNS)
    run_parity(p, np.random.default_rng(203))


def test_bsim4_parity_plain_instances():
    """Only W and L given: default areas/perimeters from the geometry model."""
    n = 6000
    spec = bsim4_spec(n, seed=204, cards=ALL_CARDS, geometry=False)
    print(f"bsim4 plain n={n}", flush=True)
    p = Pair("bsim4", spec, topology="private", seed=205, noise_runs=NOISE_RUNS)
    run_parity(p, np.random.default_rng(206))


def test_bsim4_shared_size_records():
    """Few distinct (W, L): instances share size-dependent records created
    at the first instance's temperature."""
    n = 3000
    spec = bsim4_spec(n, seed=207, cards=BSIM4_CARDS, geometry=False, temp_spread=True, sizes=17)
    print(f"bsim4 shared sizes n={n}", flush=True)
    p = Pair("bsim4", spec, topology="shared", seed=208, noise_runs=NOISE_RUNS)
    run_parity(p, np.random.default_rng(209))


def test_bsim4_parity_rails():
    """24 shared nodes for 3000 instances: every target collects hundreds of
    contributions (the one-wave long-segment scatter; the rows past its
    in-order limit take the chunked sum)."""
    n = 3000
    spec = bsim4_spec(n, seed=219, cards=BSIM4_CARDS, geometry=False)
    print(f"bsim4 rails n={n}", flush=True)
    p = Pair("bsim4", spec, topology="shared", seed=220, n_shared_nodes=24, noise_runs=NOISE_RUNS)
    run_parity(p, np.random.default_rng(221))


def test_bsim4_parity_limiter_off():
    import oracle
    n = 3000
    spec = bsim4_spec(n, seed=210, cards=ALL_CARDS)
    print(f"bsim4 limiter off n={n}", flush=True)
    p = Pair("bsim4", spec, options=oracle.default_options(voltageLimiterFlag=0), seed=211,
             noise_runs=NOISE_RUNS)
    run_parity(p, np.random.default_rng(212))


@pytest.mark.parametrize("topology", ["private", "blocked"])
def test_bsim4_fused_equals_split(topology):
    spec = bsim4_spec(4000, seed=213, cards=BSIM4_CARDS[:2], geometry=False)
    print(f"bsim4 fused vs split {topology} n=4000", flush=True)
    twin_drive("bsim4", spec, 214, np.random.default_rng(215), fused_b=True, topology=topology)


def test_bsim4_run_to_run_bit_identical():
    spec = bsim4_spec(4000, seed=216, cards=ALL_CARDS)
    print("bsim4 determinism n=4000", flush=True)
    twin_drive("bsim4", spec, 217, np.random.default_rng(218), fused_b=False, topology="shared")


@pytest.mark.parametrize("variant", ["plain", "delvto", "temp"])
def test_bsim4_card_modes(variant):
    """Production parity in each card mode of the update kernel (parameter
    columns served from the model cards where card-uniform)."""
    from test_gpu_bitexact import CARD_LAYOUTS, _card_mode_spec
    p = Pair("bsim4", _card_mode_spec(variant, 3000), topology="shared", seed=715,
             ground_frac=0.05, noise_runs=NOISE_RUNS)
    if variant in CARD_LAYOUTS:
        assert p.gpu.param_layout() == CARD_LAYOUTS[variant]
    run_parity(p, np.random.default_rng(716))


@pytest.mark.parametrize("blocks", [7, -1])
def test_bsim4_persistent_update_equals_tiles(blocks):
    """The persistent update (a resident grid looping over the tiles, here 7
    workgroups or the GPU's capacity) gives the bits of one workgroup per
    tile."""
    spec = bsim4_spec(20000, seed=231, cards=ALL_CARDS, geometry=True, temp_spread=True)
    print(f"bsim4 persistent update blocks={blocks} n=20000", flush=True)
    twin_drive("bsim4", spec, 232, np.random.default_rng(233), fused_b=False,
               prepare_b=lambda p: p.gpu.set_persistent(blocks))
