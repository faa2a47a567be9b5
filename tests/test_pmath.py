"""The portable exp/log/pow/tan of the parity builds (xa_pmath.h, compiled
into the oracle's parity build) against the C library, and the parity
build's loads against the production oracle's.

Accuracy bars (ulps of the C library's result, which is within 1 ulp of
the true value): exp and log 1, tan 3 for |x| <= 20, pow 2 + 2|y log x|.
Special values follow C99 F.9 (checked by value).  A coefficient or split
constant off in its last digits moves results by many ulps and fails these.
"""
import math

import numpy as np
import pytest

import oracle
from harness import Pair
from specs import BSIM4_CARDS, bjt_spec, bsim4_spec, diode_spec, mosfet1_spec

PM = None


def _pm(which, a, b=0.0):
    global PM
    if PM is None:
        PM = oracle.lib("pm")
    return PM.or_math(which, float(a), float(b))


def _ulps(got, want):
    if got == want:
        return 0.0
    return abs(got - want) / (math.ulp(want) if want != 0 else math.ulp(0.0))


def test_exp_log_within_one_ulp():
    rng = np.random.default_rng(1)
    for x in np.concatenate([rng.uniform(-745, 709.7, 20000), rng.uniform(-1, 1, 20000),
                             rng.uniform(-40, 40, 20000)]):
        assert _ulps(_pm(0, x), math.exp(x)) <= 1.0, x
    for x in np.concatenate([np.exp(rng.uniform(-700, 700, 20000)),
                             rng.uniform(0.5, 2.0, 20000), [5e-324, 1e-310, 2.2e-308]]):
        assert _ulps(_pm(1, x), math.log(x)) <= 1.0, x


def test_tan_within_three_ulps():
    rng = np.random.default_rng(2)
    for x in np.concatenate([rng.uniform(-20, 20, 40000), rng.uniform(-1.5, 1.5, 20000)]):
        assert _ulps(_pm(3, x), math.tan(x)) <= 3.0, x


def test_sin_within_three_ulps():
    rng = np.random.default_rng(4)
    for x in np.concatenate([rng.uniform(-20, 20, 40000), rng.uniform(-1.5, 1.5, 20000)]):
        want = math.sin(x)
        got = _pm(4, x)
        assert _ulps(got, want) <= 3.0 or abs(got - want) <= 1e-16, x
    for k in range(-12, 13):          # near the zeros (k pi rounded): absolute error
        x = k * math.pi
        assert abs(_pm(4, x) - math.sin(x)) <= 1e-17 + 3 * math.ulp(math.sin(x)), k


def test_pow_within_bound():
    rng = np.random.default_rng(3)
    for _ in range(40000):
        x, y = rng.uniform(1e-3, 10.0), rng.uniform(-5.0, 5.0)
        bound = 2.0 + 2.0 * abs(y * math.log(x))
        assert _ulps(_pm(2, x, y), math.pow(x, y)) <= bound, (x, y)


def test_special_values():
    inf, nan = math.inf, math.nan
    assert _pm(0, 0.0) == 1.0 and _pm(0, 710.0) == inf and _pm(0, -746.0) == 0.0
    assert _pm(0, -inf) == 0.0 and math.isnan(_pm(0, nan))
    assert _pm(1, 1.0) == 0.0 and _pm(1, 0.0) == -inf and _pm(1, inf) == inf
    assert math.isnan(_pm(1, -1.0))
    assert _pm(2, 2.0, 10.0) == 1024.0
    assert _pm(2, -2.0, 3.0) == pytest.approx(-8.0, rel=1e-15)   # odd power keeps the sign
    assert _pm(2, -2.0, 2.0) == pytest.approx(4.0, rel=1e-15)
    assert _pm(2, 0.0, 2.0) == 0.0 and _pm(2, 0.0, -1.0) == inf and _pm(2, -0.0, -1.0) == -inf
    assert _pm(2, 5.0, 0.0) == 1.0 and _pm(2, 1.0, nan) == 1.0
    assert math.isnan(_pm(2, -2.0, 0.5))
    assert _pm(3, 0.0) == 0.0 and math.isnan(_pm(3, inf))
    # log of a power of two is k ln 2 to the last bit or one ulp
    for k in range(-1074, 1024, 7):
        assert _ulps(_pm(1, 2.0 ** k), math.log(2.0 ** k)) <= 1.0, k


@pytest.mark.parametrize("model", ["diode", "mosfet1", "bjt", "bsim4"])
def test_parity_build_loads_match_production_oracle(model):
    """The parity build changes only the last bits of exp/log/pow/tan: its
    loads meet the production oracle's within the GPU parity bar's scale."""
    from gpu_common import ITERS
    from harness import assert_parity
    spec = {"diode": lambda: diode_spec(400, seed=9),
            "mosfet1": lambda: mosfet1_spec(400, seed=9),
            "bjt": lambda: bjt_spec(400, seed=9),
            "bsim4": lambda: bsim4_spec(400, seed=9, cards=BSIM4_CARDS)}[model]()
    a = Pair(model, spec, topology="shared", seed=10, gpu=False, noise_runs=4)
    b = Pair(model, spec, topology="shared", seed=10, gpu=False, variant="pm")
    rng = np.random.default_rng(12)
    cs = rng.uniform(-0.5, 1.5, a.topo.n_sto)
    a.set_inputs(currSto=cs)
    b.set_inputs(currSto=cs)
    for flags in ITERS:
        x = rng.uniform(-0.5, 1.5, a.topo.n_nodes)
        a.set_inputs(nextSol=x)
        b.set_inputs(nextSol=x)
        ra, _ = a.load(**flags)
        rb, _ = b.load(**flags)
        for k in ra:
            assert_parity(rb[k], ra[k], what=f"{model} {flags} {k}", scale=a.mag.get(k),
                          noise=a.noise.get(k))
