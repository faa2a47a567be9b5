"""Pins of the oracle's independent voltage source (N_DEV_Vsrc.C with the
waveforms of N_DEV_SourceData.C), CPU only: each waveform against its
definition (SPICE PULSE, SIN, EXP), the DC value at the operating point and
the waveform in transient, the solver-state defaults of parameters left not
given, and the branch equations: F = (i, -i, v+ - v- - src), dF/dx = +-1."""
import math

import numpy as np
import pytest

from harness import Pair
from specs import vsource_spec

N = 64


def _pulse(t, v1, v2, td, tr, tf, pw, per):
    t = t - td
    if t > per and per != 0.0:
        t -= per * math.floor(t / per)
    if t <= 0 or t > tr + pw + tf:
        return v1
    if t < tr:
        return v1 + (v2 - v1) * t / tr
    if t <= tr + pw:
        return v2
    return v2 + (v1 - v2) * (t - tr - pw) / tf


def _sin(t, v0, va, freq, td, theta, phase):
    t = t - td
    if t <= 0:
        return v0 + va * math.sin(2 * math.pi * phase / 360)
    return v0 + va * math.sin(2 * math.pi * (freq * t + phase / 360)) * math.exp(-t * theta)


def _exp(t, v1, v2, td1, tau1, td2, tau2):
    if t <= td1:
        return v1
    out = v1 + (v2 - v1) * (1 - math.exp(-(t - td1) / tau1))
    if t > td2:
        out += (v1 - v2) * (1 - math.exp(-(t - td2) / tau2))
    return out


def _expected(p, i, t, tstep, tstop):
    g = lambda k, d: d if np.isnan(p[k][i]) else p[k][i]
    kind = p["TRANSIENTSOURCETYPE"][i]
    if kind == 2:
        return _pulse(t, p["V1"][i], p["V2"][i], g("TD", 0.0), g("TR", tstep), g("TF", tstep),
                      g("PW", tstop), g("PER", tstop))
    if kind == 0:
        return _sin(t, p["V0"][i], p["VA"][i], g("FREQ", 1.0 / tstop), g("TD", 0.0),
                    g("THETA", 0.0), p["PHASE"][i])
    if kind == 1:
        return _exp(t, p["V1"][i], p["V2"][i], g("TD1", 0.0), g("TAU1", tstep),
                    g("TD2", g("TD1", 0.0) + tstep), g("TAU2", tstep))
    return p["DCV0"][i]


@pytest.fixture(scope="module")
def pair():
    return Pair("vsource", vsource_spec(N, seed=5), topology="shared", seed=6, gpu=False,
                ground_frac=0.2)


def test_waveforms_and_dc_selection(pair):
    p = pair.spec["inst_params"]
    tstep, tstop = 1e-12, 2e-9
    common = dict(startingTimeStep=tstep, finalTime=tstop, bpTol=0.0)
    rng = np.random.default_rng(7)
    for t in np.concatenate([[0.0], rng.uniform(0.0, 3e-9, 40)]):
        pair.load(newtonIter=1, transientFlag=1, currTime=float(t), **common)
        got = pair.orc.read("source")
        want = [_expected(p, i, t, tstep, tstop) for i in range(N)]
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-15, err_msg=f"t={t}")
    # the operating point uses DCV0 whatever the waveform; tranOP the waveform at t
    pair.load(newtonIter=1, dcopFlag=1, currTime=5e-10, **common)
    np.testing.assert_array_equal(pair.orc.read("source"), p["DCV0"])
    pair.load(newtonIter=1, dcopFlag=1, tranopFlag=1, currTime=0.0, **common)
    want = [_expected(p, i, 0.0, tstep, tstop) for i in range(N)]
    np.testing.assert_allclose(pair.orc.read("source"), want, rtol=1e-13, atol=1e-15)


def test_branch_equations(pair):
    rng = np.random.default_rng(8)
    x = rng.uniform(-1, 1, pair.topo.n_nodes)
    pair.set_inputs(nextSol=x)
    ro, _ = pair.load(newtonIter=1, transientFlag=1, currTime=3e-10, startingTimeStep=1e-12,
                      finalTime=2e-9)
    src = pair.orc.read("source")
    lids = pair.topo.node_lids.reshape(N, 3)
    v = lambda l: np.where(l >= 0, x[np.maximum(l, 0)], 0.0)
    f = np.zeros(pair.topo.n_nodes)
    dense = np.zeros((pair.topo.n_nodes, pair.topo.n_nodes))
    for (pp, nn, br), s in zip(lids, src):
        ib = x[br]
        for row, val in ((pp, ib), (nn, -ib)):
            if row >= 0:
                f[row] += val
        f[br] += v(np.array(pp)) - v(np.array(nn)) - s
        for r, c, val in ((pp, br, 1.0), (nn, br, -1.0), (br, pp, 1.0), (br, nn, -1.0)):
            if r >= 0 and c >= 0:
                dense[r, c] += val
    np.testing.assert_allclose(ro["f"], f, rtol=1e-14, atol=1e-15)
    rows = np.repeat(np.arange(pair.topo.n_nodes), np.diff(pair.topo.row_ptr))
    got = np.zeros_like(dense)
    got[rows, pair.topo.col_idx] = ro["dFdx"]
    np.testing.assert_array_equal(got, dense)
