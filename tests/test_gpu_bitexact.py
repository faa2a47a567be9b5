"""Bit-exact GPU = oracle parity under the parity builds.

The production GPU library calls the GPU math library's exp/log/pow/tan and
the oracle the C library's, so their loads agree to the conditioning of each
model (tests/test_gpu_*.py bound that).  In the parity builds both sides
take the same portable functions (xyce_amd/csrc/xa_pmath.h); every other
operation is IEEE-754 (+ - * /, sqrt, conversions) in the same order, with
contraction off, and the stamps are added in the reference's order.  So
every output -- F, Q, limiter vectors, lead currents, store, state and both
Jacobians -- must be equal bit for bit, over the cases of the production
parity tests: the parity cards of each model, DC / transient-OP / transient
Newton iterations, private and shared topologies with ground, limiter on
and off, lead currents, Lambert-W forms, IC and OFF instances, inputOP
flags.  (Rows of more than 1024 terms are summed in a fixed chunked order,
not the reference's one-by-one order, so no test here has such rows.)
"""
import numpy as np
import pytest

from gpu_common import ITERS
from harness import Pair
from specs import (BSIM4_CARDS, BSIM4_VTL_CARD, bjt_spec, bsim4_spec, diode_spec,
                   mosfet1_spec, vsource_spec)
from test_gpu_bjt import BJT_ITERS

pytestmark = pytest.mark.gpu


def run_exact(p, rng, iters=ITERS, fused=False, flag_sol=None):
    p.set_inputs(currSto=rng.uniform(-0.5, 1.5, p.topo.n_sto),
                 currSta=rng.uniform(-1e-15, 1e-15, p.topo.n_sta),
                 lastSta=rng.uniform(-1e-15, 1e-15, p.topo.n_sta))
    if flag_sol is not None:
        p.set_inputs(flagSol=flag_sol)
    bad = []
    total = 0
    for it, flags in enumerate(iters):
        p.set_inputs(nextSol=rng.uniform(-0.5, 1.5, p.topo.n_nodes))
        ro, rg = p.load(fused=fused, **flags)
        for k in ro:
            a, b = rg[k], ro[k]
            total += a.size
            # NaN payloads differ between the host and the GPU: both NaN is equal
            diff = (a.view(np.int64) != b.view(np.int64)) & ~(np.isnan(a) & np.isnan(b))
            if diff.any():
                idx = np.nonzero(diff)[0]
                rel = np.abs(a[idx] - b[idx]) / np.maximum(np.maximum(np.abs(a[idx]),
                                                                      np.abs(b[idx])), 1e-300)
                bad.append(f"iter {it} {k}: {idx.size} of {a.size} differ, first {idx[:4].tolist()}"
                           f" gpu {a[idx[:2]].tolist()} oracle {b[idx[:2]].tolist()}"
                           f" max rel {rel.max():.2e}")
    print(f"  {total} entries compared, {len(bad)} outputs with differences", flush=True)
    assert not bad, "\n".join(bad)


def _pair(model, spec, **kw):
    return Pair(model, spec, variant="pm", **kw)


@pytest.mark.parametrize("topology", ["private", "shared"])
def test_diode_bitexact(topology):
    n = 8000
    spec = diode_spec(n, seed=101, temp_spread=True)
    lead = (np.arange(n) % 3 == 0).astype(np.int32)
    run_exact(_pair("diode", spec, topology=topology, seed=102, lead=lead, ground_frac=0.05),
              np.random.default_rng(103))


@pytest.mark.parametrize("lambertw", [1, 2])
def test_diode_bitexact_lambertw_limiter_off(lambertw):
    import oracle
    n = 4000
    spec = diode_spec(n, seed=104)
    spec["inst_params"]["LAMBERTW"] = np.full(n, float(lambertw))
    p = _pair("diode", spec, options=oracle.default_options(voltageLimiterFlag=0), seed=105)
    run_exact(p, np.random.default_rng(106))


@pytest.mark.parametrize("topology", ["private", "shared"])
def test_mosfet1_bitexact(topology):
    n = 6000
    spec = mosfet1_spec(n, seed=301, temp_spread=True)
    lead = (np.arange(n) % 5 == 0).astype(np.int32)
    run_exact(_pair("mosfet1", spec, topology=topology, seed=302, lead=lead, ground_frac=0.05),
              np.random.default_rng(303))


def test_mosfet1_bitexact_limiter_off_plain():
    import oracle
    spec = mosfet1_spec(4000, seed=304, geometry=False)
    p = _pair("mosfet1", spec, options=oracle.default_options(voltageLimiterFlag=0), seed=305)
    run_exact(p, np.random.default_rng(306))


@pytest.mark.parametrize("topology", ["private", "shared"])
def test_bjt_bitexact(topology):
    n = 8000
    spec = bjt_spec(n, seed=401, temp_spread=True, lambertw_frac=0.25)
    lead = (np.arange(n) % 4 == 0).astype(np.int32)
    run_exact(_pair("bjt", spec, topology=topology, seed=402, lead=lead, ground_frac=0.05),
              np.random.default_rng(403), iters=BJT_ITERS)


def test_bjt_bitexact_ic_off_inputop():
    n = 6000
    p = _pair("bjt", bjt_spec(n, seed=404, ic=True, off_frac=0.25), seed=405)
    run_exact(p, np.random.default_rng(406),
              iters=[dict(newtonIter=0, dcopFlag=1, initJctFlag=1),
                     dict(newtonIter=1, dcopFlag=1, initFixFlag=1),
                     dict(newtonIter=2, dcopFlag=1)])
    q = _pair("bjt", bjt_spec(n, seed=407), topology="shared", seed=408, ground_frac=0.05)
    rng = np.random.default_rng(409)
    run_exact(q, rng, iters=[dict(newtonIter=0, dcopFlag=1, initJctFlag=1, inputOPFlag=1)],
              flag_sol=(rng.random(q.topo.n_nodes) < 0.9).astype(np.float64))


def test_bjt_bitexact_new_excess_phase():
    import oracle
    spec = bjt_spec(6000, seed=426, temp_spread=True)
    p = _pair("bjt", spec, options=oracle.default_options(newExcessPhase=1), seed=427,
              topology="shared", ground_frac=0.05)
    run_exact(p, np.random.default_rng(428), iters=BJT_ITERS)


def test_bjt_bitexact_limiter_off():
    import oracle
    spec = bjt_spec(4000, seed=410, lambertw_frac=0.25)
    p = _pair("bjt", spec, options=oracle.default_options(voltageLimiterFlag=0), seed=411)
    run_exact(p, np.random.default_rng(412), iters=BJT_ITERS)


@pytest.mark.parametrize("topology", ["private", "shared"])
def test_bsim4_bitexact(topology):
    n = 4000
    spec = bsim4_spec(n, seed=201, cards=BSIM4_CARDS + [BSIM4_VTL_CARD], geometry=True,
                      nf_mix=True, temp_spread=True)
    run_exact(_pair("bsim4", spec, topology=topology, seed=202, ground_frac=0.05),
              np.random.default_rng(203))


def test_bsim4_bitexact_plain_limiter_off_fused():
    import oracle
    spec = bsim4_spec(3000, seed=204, cards=BSIM4_CARDS, geometry=False)
    p = _pair("bsim4", spec, options=oracle.default_options(voltageLimiterFlag=0), seed=205)
    run_exact(p, np.random.default_rng(206))
    q = _pair("bsim4", bsim4_spec(2000, seed=207, cards=BSIM4_CARDS[:2], geometry=False),
              seed=208)
    q.gpu.set_path("fused", 128)
    run_exact(q, np.random.default_rng(209), fused=True)


def test_vsource_bitexact():
    """SIN and EXP waveforms through the shared portable sin/exp."""
    p = _pair("vsource", vsource_spec(4000, seed=701), topology="shared", seed=702,
              ground_frac=0.2)
    times = dict(startingTimeStep=1e-12, finalTime=2e-9, bpTol=1e-18)
    iters = [dict(newtonIter=1, transientFlag=1, currTime=float(t), **times)
             for t in np.linspace(0.0, 2.5e-9, 9)]
    run_exact(p, np.random.default_rng(703), iters=iters)


def _card_mode_spec(variant, n):
    spec = bsim4_spec(n, seed=711, cards=BSIM4_CARDS[:2], geometry=False)
    rng = np.random.default_rng(712)
    if variant == "delvto":
        spec["inst_params"]["DELVTO"] = rng.uniform(-0.02, 0.02, n)
    elif variant == "temp":
        spec["inst_params"]["TEMP"] = rng.uniform(-20.0, 120.0, n)
    return spec


# (per-instance columns, card columns) of each variant: binless cards with
# W/L only serve all but the geometry-derived columns from the cards; a
# per-instance DELVTO keeps the instance values per instance
CARD_LAYOUTS = {"plain": (14, 123), "delvto": (51, 86)}


@pytest.mark.parametrize("variant", ["plain", "delvto", "temp"])
def test_bsim4_card_modes_bitexact(variant):
    """The update kernel's card modes (parameter columns served from the
    model cards when card-uniform) read the same bits as the columns."""
    p = _pair("bsim4", _card_mode_spec(variant, 3000), topology="shared", seed=713,
              ground_frac=0.05)
    layout = p.gpu.param_layout()
    print(f"  {variant}: per-instance columns {layout[0]}, card columns {layout[1]}")
    if variant in CARD_LAYOUTS:
        assert layout == CARD_LAYOUTS[variant]
    run_exact(p, np.random.default_rng(714))
