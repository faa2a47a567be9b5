"""The oracle's updateState over OpenMP threads (the CPU baseline's
all-cores run) gives the serial run's bits: instances are independent in
that loop, as in the reference's _OMP loop (N_DEV_MOSFET_B4.C:16184), and
the loads that add into shared rows stay serial."""
import numpy as np
import pytest

import oracle
from harness import Pair
from specs import bjt_spec, bsim4_spec, diode_spec, mosfet1_spec

SPECS = {
    "diode": lambda: diode_spec(3000, seed=401),
    "mosfet1": lambda: mosfet1_spec(2000, seed=402),
    "bjt": lambda: bjt_spec(2000, seed=403),
    "bsim4": lambda: bsim4_spec(500, seed=404),
}
ITERS = [dict(newtonIter=0, dcopFlag=1, initJctFlag=1),
         dict(newtonIter=1, transientFlag=1)]


def _run(model, threads):
    oracle.set_threads(threads)
    try:
        p = Pair(model, SPECS[model](), topology="shared", seed=405, gpu=False, ground_frac=0.02)
        rng = np.random.default_rng(406)
        p.set_inputs(currSto=rng.uniform(-0.5, 1.5, p.topo.n_sto),
                     currSta=rng.uniform(-1e-15, 1e-15, p.topo.n_sta))
        outs = []
        for flags in ITERS:
            p.set_inputs(nextSol=rng.uniform(-0.5, 1.5, p.topo.n_nodes))
            res, _ = p.load(**flags)
            outs.append(res)
        return outs
    finally:
        oracle.set_threads(1)


@pytest.mark.parametrize("model", list(SPECS))
def test_threaded_update_state_bit_identical(model):
    a, b = _run(model, 1), _run(model, 4)
    for it, (ra, rb) in enumerate(zip(a, b)):
        for k in ra:
            assert np.array_equal(ra[k].view(np.int64), rb[k].view(np.int64)), (it, k)
