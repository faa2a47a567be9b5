"""GPU parity of the diode load against the oracle (through the C ABI).

Bar: every F/Q/limiter/lead/store/CSR entry within 1e-12 relative of the
oracle (FP64 everywhere; the difference budget is the last-ulp difference of
the GPU's exp/log/pow against glibc's), across Newton iteration kinds, private
and shared topologies with ground nodes; assembly bit-identical from run to run;
the fused one-call load equal to the four calls.
"""
import numpy as np
import pytest

from harness import Pair, assert_close, max_rel_err
from specs import diode_spec

pytestmark = pytest.mark.gpu
RTOL = 1e-12
ITERS = [dict(newtonIter=0, dcopFlag=1, initJctFlag=1),
         dict(newtonIter=1, dcopFlag=1),
         dict(newtonIter=0, transientFlag=1),
         dict(newtonIter=1, transientFlag=1),
         dict(newtonIter=2, transientFlag=1)]


def _run(p, rng, iters=ITERS, fused=False):
    p.set_inputs(currSto=rng.uniform(-0.5, 1.5, p.topo.n_sto))
    worst = 0.0
    for it, flags in enumerate(iters):
        p.set_inputs(nextSol=rng.uniform(-0.5, 1.5, p.topo.n_nodes))
        ro, rg = p.load(fused=fused, **flags)
        for k in ro:
            worst = max(worst, max_rel_err(rg[k], ro[k]))
            assert_close(rg[k], ro[k], RTOL, 1e-300, what=f"iter {it} {flags} {k}",
                         scale=p.mag.get(k))
    return worst


@pytest.mark.parametrize("topology", ["private", "shared"])
def test_diode_parity(topology):
    n = 20000
    spec = diode_spec(n, seed=101, temp_spread=True)
    lead = (np.arange(n) % 3 == 0).astype(np.int32)
    p = Pair("diode", spec, topology=topology, seed=102, lead=lead, ground_frac=0.05)
    print(f"diode {topology}: worst rel err", _run(p, np.random.default_rng(103)))


@pytest.mark.parametrize("lambertw", [1, 2])
def test_diode_parity_limiter_off_lambertw(lambertw):
    """Lambert-W current forms.  One LAMBERTW value per group: an RS = 0 card
    shared by LAMBERTW=1 and =0 instances is rewritten to RS = 1e-12 by the
    first Lambert-W instance (N_DEV_Diode.C:444-445), after which the
    reference's registerLIDs reads an internal LID that was never allocated."""
    import oracle
    n = 4000
    spec = diode_spec(n, seed=104)
    spec["inst_params"]["LAMBERTW"] = np.full(n, float(lambertw))
    opts = oracle.default_options(voltageLimiterFlag=0)
    p = Pair("diode", spec, options=opts, topology="private", seed=105)
    _run(p, np.random.default_rng(106))


def test_diode_parity_rails():
    """128 shared nodes for 20000 diodes: segments of a few hundred
    contributions, added in instance order by one wavefront each."""
    n = 20000
    spec = diode_spec(n, seed=116)
    p = Pair("diode", spec, topology="shared", seed=117, n_shared_nodes=128)
    print("diode rails: worst rel err", _run(p, np.random.default_rng(118)))


def test_diode_parity_supply_rail():
    """Two shared nodes for 40000 diodes: each row collects more contributions
    than the in-order limit, so it is summed in fixed 256-term chunks whose
    sums combine in a fixed tree; the result stays within 1e-12 of the
    contributions' magnitude sum of the one-by-one order."""
    n = 40000
    spec = diode_spec(n, seed=119)
    p = Pair("diode", spec, topology="shared", seed=120, n_shared_nodes=2)
    print("diode supply rail: worst rel err", _run(p, np.random.default_rng(121)))


def test_diode_fused_equals_split():
    n = 10000
    spec = diode_spec(n, seed=107)
    p = Pair("diode", spec, topology="private", seed=108)
    _run(p, np.random.default_rng(109), fused=True)


def test_diode_fused_equals_split_blocked():
    """Nodes shared within blocks of instances: the fused load writes them
    from LDS in the scatter's order."""
    from gpu_common import twin_drive
    spec = diode_spec(10000, seed=113)
    lead = (np.arange(10000) % 3 == 0).astype(np.int32)
    twin_drive("diode", spec, 114, np.random.default_rng(115), fused_b=True, topology="blocked",
               lead=lead, iters=ITERS)


def test_diode_run_to_run_bit_identical():
    n = 30000
    spec = diode_spec(n, seed=110)
    outs = []
    for rep in range(2):
        p = Pair("diode", spec, topology="shared", seed=111, ground_frac=0.02)
        rng = np.random.default_rng(112)
        p.set_inputs(currSto=rng.uniform(-0.5, 1.5, p.topo.n_sto),
                     nextSol=rng.uniform(-0.5, 1.5, p.topo.n_nodes))
        _, rg = p.load(newtonIter=1, transientFlag=1)
        outs.append(rg)
    for k in outs[0]:
        assert np.array_equal(outs[0][k].view(np.int64), outs[1][k].view(np.int64)), k
