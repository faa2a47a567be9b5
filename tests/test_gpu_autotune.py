"""The block-shape autotuner (xa_group_autotune): after tuning on a group's
own vectors, xa_load_all (the tuned path and block) gives the four calls'
bits on a twin group driven with the same inputs (after the caller's vectors
are reset, as the tuning loads write them), so the tuning left the
instances' own state as it was; a topology whose nodes are shared across
blocks has no fused load and tunes to the four calls."""
import numpy as np
import pytest

from gpu_common import ITERS, twin_drive
from harness import Pair
from specs import BSIM4_CARDS, bjt_spec, bsim4_spec, diode_spec, mosfet1_spec

pytestmark = pytest.mark.gpu


def _tune(p, want=None):
    ss = p._ss("gpu", newtonIter=1, transientFlag=1, pdt=1e10)
    route, block = p.gpu.autotune(ss, p._vecs("gpu"), reps=2)
    print(f"  tuned: {route}, block {block}", flush=True)
    # the tuning loads wrote the caller's vectors (outputs, store, state, as a
    # load does); the twin starts from the same zeroed vectors as the other
    for t in p.d.values():
        t.zero_()
    assert route in ("split", "fused") and block in (64, 128, 256)
    if want is not None:
        assert route == want
    return route, block


@pytest.mark.parametrize("model,topology", [("diode", "private"), ("bsim4", "private"),
                                            ("mosfet1", "blocked"), ("bjt", "private")])
def test_autotuned_load_all_equals_four_calls(model, topology):
    spec = {"diode": lambda: diode_spec(20000, seed=501),
            "bsim4": lambda: bsim4_spec(3000, seed=502, cards=BSIM4_CARDS),
            "mosfet1": lambda: mosfet1_spec(6000, seed=503),
            "bjt": lambda: bjt_spec(6000, seed=504)}[model]()
    print(f"{model} {topology}", flush=True)
    twin_drive(model, spec, 505, np.random.default_rng(506), fused_b=True, topology=topology,
               prepare_b=_tune)


def test_shared_nodes_tune_to_four_calls():
    p = Pair("diode", diode_spec(8000, seed=507), topology="shared", seed=508,
             n_shared_nodes=64)
    _tune(p, want="split")


@pytest.mark.parametrize("path,block", [("pattern", 64), ("pattern", 128), ("pattern", 256),
                                        ("fused", 64), ("split", 128)])
def test_pinned_paths_equal_four_calls(path, block):
    """Each pinned xa_load_all path and block size gives the four calls' bits
    (the pattern-fused load: every stamp added into the instance's staged
    output span in stamp order).  Private nodes without ground and one
    node-collapse pattern, so every kind has a pattern plan."""
    # cards without RSH: no instance has internal nodes, so every instance's
    # stamps follow one pattern
    spec = bsim4_spec(3000, seed=509, cards=[BSIM4_CARDS[i] for i in (0, 1, 3)])
    print(f"bsim4 {path} {block}", flush=True)
    twin_drive("bsim4", spec, 510, np.random.default_rng(511), fused_b=True, topology="private",
               prepare_b=lambda p: p.gpu.set_path(path, block), ground_frac=0.0)


def test_pattern_path_needs_pattern_plans():
    p = Pair("bsim4", bsim4_spec(2000, seed=512, cards=BSIM4_CARDS), topology="shared", seed=513,
             n_shared_nodes=64)
    with pytest.raises(RuntimeError):
        p.gpu.set_path("pattern", 64)
