"""GPU parity of the independent voltage source against the oracle, through
the C ABI: DC values at the operating point, the PULSE/SIN/EXP/DC waveforms
at a series of transient times (with parameters left to the solver-state
defaults), shared and grounded nodes, the fused one-call load.

Bar: every entry within 1e-12 of its error scale (the sum of |terms|); the
DC and PULSE values are sums of IEEE operations on the same inputs and match
bit for bit, SIN and EXP differ by the last bits of the GPU's sin/exp
(bit-exact in the parity build, tests/test_gpu_bitexact.py)."""
import numpy as np
import pytest

from harness import Pair, assert_parity
from specs import vsource_spec

pytestmark = pytest.mark.gpu

TIMES = dict(startingTimeStep=1e-12, finalTime=2e-9, bpTol=1e-18)


def _iters():
    its = [dict(newtonIter=0, dcopFlag=1, initJctFlag=1, **TIMES),
           dict(newtonIter=1, dcopFlag=1, tranopFlag=1, currTime=0.0, **TIMES)]
    for t in np.linspace(1e-11, 2.5e-9, 12):
        its.append(dict(newtonIter=1, transientFlag=1, currTime=float(t), **TIMES))
    return its


@pytest.mark.parametrize("fused", [False, True])
def test_vsource_parity(fused):
    n = 5000
    p = Pair("vsource", vsource_spec(n, seed=601), topology="shared", seed=602, ground_frac=0.2)
    rng = np.random.default_rng(603)
    exact = {}
    pulse_dc = np.isin(p.spec["inst_params"]["TRANSIENTSOURCETYPE"], (2.0, 6.0)) | \
        np.isnan(p.spec["inst_params"]["TRANSIENTSOURCETYPE"])
    for flags in _iters():
        p.set_inputs(nextSol=rng.uniform(-1.0, 1.0, p.topo.n_nodes))
        ro, rg = p.load(fused=fused, **flags)
        for k in ro:
            st = assert_parity(rg[k], ro[k], what=f"{flags} {k}", scale=p.mag.get(k))
            exact[k] = min(exact.get(k, 1.0), st["bitexact"])
        # the branch rows of DC and PULSE sources carry no transcendental
        br = p.topo.node_lids.reshape(n, 3)[:, 2]
        np.testing.assert_array_equal(rg["f"][br[pulse_dc]].view(np.int64),
                                      ro["f"][br[pulse_dc]].view(np.int64))
    print("  bit-exact fraction per output:", {k: round(v, 4) for k, v in exact.items()})
    assert exact["dFdx"] == 1.0
