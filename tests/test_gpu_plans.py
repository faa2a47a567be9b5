"""Scatter plan shapes.  A private-node group of one node-collapse pattern
(the microbenchmark layout) takes the pattern plan for every output kind it
stamps (targets at fixed offsets from a per-instance base), and a group whose
instances share nodes takes the segmented plan.  The pattern plan's bits are
checked elsewhere: the private-topology parity tests run it against the
oracle, and the fused-equals-split tests against the fused load's block
plans, which are built from the segmented plan."""
import pytest

from harness import Pair
from specs import diode_spec
from xyce_amd.workloads import Workload

pytestmark = pytest.mark.gpu
KINDS = ["F", "Q", "dFdxdVp", "dQdxdVp", "dFdx", "dQdx"]


@pytest.mark.parametrize("model", ["diode", "mosfet1", "bjt", "bsim4"])
def test_private_groups_take_pattern_plans(model):
    wl = Workload(model, 5000)
    info = {k: wl.master.plan_info(i) for i, k in enumerate(KINDS)}
    print(f"{model}: {info}", flush=True)
    for k, (nt, ns, slots) in info.items():
        if nt:
            assert slots > 0 and nt == slots * 5000, (k, info[k])


def test_shared_nodes_take_segmented_plans():
    p = Pair("diode", diode_spec(4000, seed=601), topology="shared", seed=602, n_shared_nodes=64)
    for i, k in enumerate(KINDS):
        nt, ns, slots = p.gpu.plan_info(i)
        if nt:
            assert slots == 0, k
