"""GPU parity of the BSIM4 (level 54) load against the oracle, through the C ABI.

Bar: every F/Q/limiter/store/state/CSR entry within 1e-12 of the oracle
relative to the entry's error scale (the larger of the two values and the sum
of |contributions| assembled into it), over the parity cards, Newton
iteration kinds of DC and transient analysis, private and shared topologies
with ground nodes; assembly bit-identical from run to run; the fused one-call
load equal to the four calls.
"""
import numpy as np
import pytest

from harness import Pair, assert_close, max_rel_err
from specs import BSIM4_CARDS, BSIM4_VTL_CARD, bsim4_spec

pytestmark = pytest.mark.gpu
RTOL = 1e-12
PDT = 2.0e9   # 1/dt of a 0.5 ns step
ITERS = [dict(newtonIter=0, dcopFlag=1, initJctFlag=1),
         dict(newtonIter=1, dcopFlag=1),
         dict(newtonIter=0, transientFlag=1, initTranFlag=1, pdt=PDT),
         dict(newtonIter=1, transientFlag=1, pdt=PDT),
         dict(newtonIter=2, transientFlag=1, pdt=PDT)]
ALL_CARDS = BSIM4_CARDS + [BSIM4_VTL_CARD]


def _run(p, rng, iters=ITERS, fused=False):
    p.set_inputs(currSto=rng.uniform(-0.5, 1.5, p.topo.n_sto))
    worst = {}
    for it, flags in enumerate(iters):
        print(f"  iteration {it}: {flags}", flush=True)
        p.set_inputs(nextSol=rng.uniform(-0.5, 1.5, p.topo.n_nodes))
        ro, rg = p.load(fused=fused, **flags)
        for k in ro:
            worst[k] = max(worst.get(k, 0.0), max_rel_err(rg[k], ro[k]))
            assert_close(rg[k], ro[k], RTOL, 1e-300, what=f"iter {it} {flags} {k}",
                         scale=p.mag.get(k))
    return worst


@pytest.mark.parametrize("topology", ["private", "shared"])
def test_bsim4_parity(topology):
    n = 6000
    spec = bsim4_spec(n, seed=201, cards=ALL_CARDS, geometry=True, nf_mix=True, temp_spread=True)
    print(f"bsim4 parity {topology} n={n}", flush=True)
    p = Pair("bsim4", spec, topology=topology, seed=202, ground_frac=0.05)
    print("worst rel err", _run(p, np.random.default_rng(203)))


def test_bsim4_parity_plain_instances():
    """Only W and L given: default areas/perimeters from the geometry model."""
    n = 6000
    spec = bsim4_spec(n, seed=204, cards=ALL_CARDS, geometry=False)
    print(f"bsim4 plain n={n}", flush=True)
    p = Pair("bsim4", spec, topology="private", seed=205)
    print("worst rel err", _run(p, np.random.default_rng(206)))


def test_bsim4_shared_size_records():
    """Few distinct (W, L): instances share size-dependent records created
    at the first instance's temperature."""
    n = 3000
    spec = bsim4_spec(n, seed=207, cards=BSIM4_CARDS, geometry=False, temp_spread=True, sizes=17)
    print(f"bsim4 shared sizes n={n}", flush=True)
    p = Pair("bsim4", spec, topology="shared", seed=208)
    _run(p, np.random.default_rng(209))


def test_bsim4_parity_limiter_off():
    import oracle
    n = 3000
    spec = bsim4_spec(n, seed=210, cards=ALL_CARDS)
    print(f"bsim4 limiter off n={n}", flush=True)
    p = Pair("bsim4", spec, options=oracle.default_options(voltageLimiterFlag=0), seed=211)
    _run(p, np.random.default_rng(212))


def test_bsim4_fused_equals_split():
    n = 4000
    spec = bsim4_spec(n, seed=213, cards=BSIM4_CARDS[:2], geometry=False)
    print(f"bsim4 fused vs split n={n}", flush=True)
    p = Pair("bsim4", spec, topology="private", seed=214)
    rng = np.random.default_rng(215)
    p.set_inputs(currSto=rng.uniform(-0.5, 1.5, p.topo.n_sto))
    for flags in ITERS:
        p.set_inputs(nextSol=rng.uniform(-0.5, 1.5, p.topo.n_nodes))
        _, split = p.load(fused=False, **flags)
        _, fused = p.load(fused=True, **flags)
        for k in split:
            np.testing.assert_array_equal(fused[k], split[k], err_msg=f"{flags} {k}")


def test_bsim4_run_to_run_bit_identical():
    n = 4000
    spec = bsim4_spec(n, seed=216, cards=ALL_CARDS)
    print(f"bsim4 determinism n={n}", flush=True)
    p = Pair("bsim4", spec, topology="shared", seed=217, ground_frac=0.05)
    rng = np.random.default_rng(218)
    p.set_inputs(currSto=rng.uniform(-0.5, 1.5, p.topo.n_sto),
                 nextSol=rng.uniform(-0.5, 1.5, p.topo.n_nodes))
    flags = ITERS[3]
    _, a = p.load(**flags)
    _, b = p.load(**flags)
    for k in a:
        np.testing.assert_array_equal(a[k].view(np.int64), b[k].view(np.int64), err_msg=k)
