"""GPU parity of the Gummel-Poon BJT load against the oracle, through the C ABI.

Bar: harness.assert_parity (1e-12 of each entry's error scale, or twice its
libm noise; at most 0.05% outliers, none beyond 1e-6).  Covered: the parity
cards (internal collector/base/emitter nodes, high injection with roll-off,
Early effect both ways, leakage currents, base-current crowding, substrate
and transit-time charges with the VTF/ITF bias dependence, old-style excess
phase, and the new-DAE excess phase of DeviceOptions newExcessPhase with its
two unknowns per instance), Lambert-W junction currents, lead currents,
Newton iterations of DC,
transient-operating-point and transient analysis (first step with
beginIntegration, later steps with unequal step sizes), initial conditions,
OFF instances under initJct/initFix, inputOP flags, limiter on and off,
private and shared topologies with ground nodes.  Fused one call equals the
four calls; two runs are bit-identical.
"""
import numpy as np
import pytest

from gpu_common import NOISE_RUNS, PDT, run_parity, twin_drive
from harness import Pair
from specs import bjt_spec

pytestmark = pytest.mark.gpu

STEP = dict(transientFlag=1, pdt=PDT, currTimeStep=5e-10)
BJT_ITERS = [dict(newtonIter=0, dcopFlag=1, initJctFlag=1),
             dict(newtonIter=1, dcopFlag=1),
             dict(newtonIter=1, dcopFlag=1, tranopFlag=1),
             dict(newtonIter=0, initTranFlag=1, beginIntegrationFlag=1, lastTimeStep=5e-10, **STEP),
             dict(newtonIter=1, beginIntegrationFlag=1, lastTimeStep=5e-10, **STEP),
             dict(newtonIter=1, lastTimeStep=4e-10, **STEP),
             dict(newtonIter=2, lastTimeStep=4e-10, **STEP)]


@pytest.mark.parametrize("topology", ["private", "shared"])
def test_bjt_parity(topology):
    n = 20000
    spec = bjt_spec(n, seed=401, temp_spread=True, lambertw_frac=0.25)
    lead = (np.arange(n) % 4 == 0).astype(np.int32)
    print(f"bjt parity {topology} n={n}", flush=True)
    p = Pair("bjt", spec, topology=topology, seed=402, lead=lead, ground_frac=0.05,
             noise_runs=NOISE_RUNS)
    run_parity(p, np.random.default_rng(403), iters=BJT_ITERS)


def test_bjt_initial_conditions_and_off():
    """IC1/IC2 given (initJct starts from them), a quarter of the instances
    OFF (zero junction voltages under initJct and initFix)."""
    n = 8000
    spec = bjt_spec(n, seed=404, ic=True, off_frac=0.25)
    print(f"bjt ic/off n={n}", flush=True)
    p = Pair("bjt", spec, seed=405, noise_runs=NOISE_RUNS)
    iters = [dict(newtonIter=0, dcopFlag=1, initJctFlag=1),
             dict(newtonIter=1, dcopFlag=1, initFixFlag=1),
             dict(newtonIter=2, dcopFlag=1)]
    run_parity(p, np.random.default_rng(406), iters=iters)


def test_bjt_inputop_flags():
    """inputOP: junctions start at vcrit unless every terminal's flag is set
    (ground counts as unset)."""
    n = 8000
    spec = bjt_spec(n, seed=407)
    print(f"bjt inputOP n={n}", flush=True)
    p = Pair("bjt", spec, topology="shared", seed=408, ground_frac=0.05, noise_runs=NOISE_RUNS)
    rng = np.random.default_rng(409)
    p.set_inputs(flagSol=(rng.random(p.topo.n_nodes) < 0.9).astype(np.float64))
    run_parity(p, rng, iters=[dict(newtonIter=0, dcopFlag=1, initJctFlag=1, inputOPFlag=1)])


def test_bjt_parity_limiter_off():
    import oracle
    n = 8000
    spec = bjt_spec(n, seed=410, lambertw_frac=0.25)
    print(f"bjt limiter off n={n}", flush=True)
    p = Pair("bjt", spec, options=oracle.default_options(voltageLimiterFlag=0), seed=411,
             noise_runs=NOISE_RUNS)
    run_parity(p, np.random.default_rng(412), iters=BJT_ITERS)


@pytest.mark.parametrize("topology", ["private", "blocked"])
def test_bjt_fused_equals_split(topology):
    spec = bjt_spec(6000, seed=413)
    print(f"bjt fused vs split {topology} n=6000", flush=True)
    lead = (np.arange(6000) % 3 == 0).astype(np.int32)
    twin_drive("bjt", spec, 414, np.random.default_rng(415), fused_b=True, topology=topology, lead=lead,
               iters=BJT_ITERS)


def test_bjt_run_to_run_bit_identical():
    spec = bjt_spec(6000, seed=416)
    print("bjt determinism n=6000", flush=True)
    twin_drive("bjt", spec, 417, np.random.default_rng(418), fused_b=False, topology="shared",
               iters=BJT_ITERS)


@pytest.mark.parametrize("topology", ["private", "shared"])
def test_bjt_new_excess_phase_parity(topology):
    """newExcessPhase = 1: every card (one with PTF, three without: both
    stamp forms in one group), DC, operating point and transient
    iterations."""
    import oracle
    n = 12000
    spec = bjt_spec(n, seed=420, temp_spread=True)
    lead = (np.arange(n) % 5 == 0).astype(np.int32)
    print(f"bjt new excess phase {topology} n={n}", flush=True)
    p = Pair("bjt", spec, topology=topology, seed=421, lead=lead, ground_frac=0.05,
             options=oracle.default_options(newExcessPhase=1), noise_runs=NOISE_RUNS)
    run_parity(p, np.random.default_rng(422), iters=BJT_ITERS)


def test_bjt_new_excess_phase_fused_equals_split():
    import oracle
    spec = bjt_spec(6000, seed=423)
    print("bjt new excess phase fused vs split n=6000", flush=True)
    twin_drive("bjt", spec, 424, np.random.default_rng(425), fused_b=True, iters=BJT_ITERS,
               options=oracle.default_options(newExcessPhase=1))
