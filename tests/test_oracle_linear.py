"""Pins the resistor and capacitor oracle (oracle/or_linear.c) without the
reference binary: Ohm's law and Q = C V with the quadratic temperature factor
(N_DEV_Resistor.C:848-865, N_DEV_Capacitor.C:271-280), the sheet-resistance
and semiconductor-capacitance geometry (N_DEV_Resistor.C:340-355,
N_DEV_Capacitor.C:196-244), the IC= branch equations at the operating point
and in transient (N_DEV_Capacitor.C:1860-2135), finite-difference Jacobians
and KCL.  The expected values are computed here from the device laws with
numpy, in a different operation order from the oracle's, and compared to a
few ulps."""
import numpy as np
import pytest

import oracle
from harness import Pair
from specs import capacitor_spec, resistor_spec

CTOK = 273.15
TNOM_K = 300.15


def _slot(topo, r, c):
    lo, hi = topo.row_ptr[r], topo.row_ptr[r + 1]
    j = np.searchsorted(topo.col_idx[lo:hi], c)
    assert topo.col_idx[lo + j] == c
    return lo + j


def _pair(model, spec, lead=None, **opt):
    return Pair(model, spec, options=oracle.default_options(**opt), gpu=False, lead=lead)


def _temp_factor(spec, cards_key, mult_key):
    """Per-instance multiplier * (1 + TC1 dT + TC2 dT^2), dT in kelvin."""
    n = spec["n"]
    cards = spec["models"]
    mi = spec["model_index"]
    ip = spec["inst_params"]
    f = np.empty(n)
    for i in range(n):
        card = cards[mi[i]][1]
        temp = ip["TEMP"][i] + CTOK if "TEMP" in ip else TNOM_K
        tnom = card["TNOM"] + CTOK if "TNOM" in card else TNOM_K
        d = temp - tnom
        tc1 = ip["TC1"][i] if "TC1" in ip else card.get("TC1", 0.0)
        tc2 = ip["TC2"][i] if "TC2" in ip else card.get("TC2", 0.0)
        f[i] = card.get(mult_key, 1.0) * (1.0 + tc1 * d + tc2 * d * d)
    return f


def _drive(p, rng, **flags):
    x = rng.uniform(-0.5, 1.5, p.topo.n_nodes)
    p.set_inputs(nextSol=x)
    flags.setdefault("newtonIter", 1)
    flags.setdefault("transientFlag", 1)
    res, _ = p.load(**flags)
    return x, res


@pytest.mark.parametrize("sheet", [False, True])
def test_resistor_ohms_law(sheet):
    n = 300
    spec = resistor_spec(n, seed=401, temp_spread=True, sheet=sheet)
    lead = np.ones(n, np.int32)
    p = _pair("resistor", spec, lead=lead)
    x, res = _drive(p, np.random.default_rng(402))
    lids = p.topo.node_lids.reshape(-1, 2)
    ip = spec["inst_params"]
    if sheet:
        R = np.empty(n)
        for i in range(n):
            card = spec["models"][spec["model_index"][i]][1]
            w = ip["W"][i]
            nar = card.get("NARROW", 0.0)
            R[i] = card["RSH"] * (ip["L"][i] - nar) / (w - nar)
    else:
        R = ip["R"]
    G = 1.0 / (R * _temp_factor(spec, "R", "R"))
    V = x[lids[:, 0]] - x[lids[:, 1]]
    I = V * G
    np.testing.assert_allclose(res["f"][lids[:, 0]], I, rtol=4e-16, atol=0)
    np.testing.assert_allclose(res["f"][lids[:, 1]], -I, rtol=4e-16, atol=0)
    np.testing.assert_allclose(res["storeLeadF"], I, rtol=4e-16, atol=0)
    t = p.topo
    for i in range(0, n, 7):
        a, b = lids[i]
        blk = [[res["dFdx"][_slot(t, r, c)] for c in (a, b)] for r in (a, b)]
        np.testing.assert_allclose(blk, G[i] * np.array([[1, -1], [-1, 1]]), rtol=4e-16)


def test_resistor_defaults():
    """R absent and no RSH: 1000 ohm; W absent: the model's DEFW."""
    n = 8
    spec = {"models": [(1, {})], "n": n, "model_index": np.zeros(n, np.int32),
            "inst_params": {"L": np.full(n, 1e-6)}}
    p = _pair("resistor", spec)
    np.testing.assert_array_equal(p.orc.read("R"), np.full(n, 1000.0))
    spec2 = {"models": [(1, {"RSH": 10.0, "DEFW": 4e-6})], "n": n,
             "model_index": np.zeros(n, np.int32), "inst_params": {"L": np.full(n, 2e-6)}}
    p2 = _pair("resistor", spec2)
    np.testing.assert_allclose(p2.orc.read("R"), np.full(n, 10.0 * 2e-6 / 4e-6), rtol=1e-15)


@pytest.mark.parametrize("mode", ["plain", "semi", "age"])
def test_capacitor_charge(mode):
    n = 300
    spec = capacitor_spec(n, seed=403, temp_spread=True, semi=mode == "semi", age=mode == "age")
    lead = np.ones(n, np.int32)
    p = _pair("capacitor", spec, lead=lead)
    x, res = _drive(p, np.random.default_rng(404))
    lids = p.topo.node_lids.reshape(-1, 2)
    ip = spec["inst_params"]
    if mode == "semi":
        base = np.empty(n)
        for i in range(n):
            card = spec["models"][spec["model_index"][i]][1]
            nar = card.get("NARROW", 0.0)
            L, W = ip["L"][i], ip["W"][i]
            base[i] = card["CJ"] * (L - nar) * (W - nar) + 2 * card.get("CJSW", 0.0) * (L + W - 2 * nar)
    elif mode == "age":
        base = np.where(ip["AGE"] >= 1, ip["C"] * (1 - ip["D"] * np.log10(ip["AGE"])), ip["C"])
    else:
        base = ip["C"]
    C = base * _temp_factor(spec, "C", "C")
    np.testing.assert_allclose(p.orc.read("C"), C, rtol=6e-16)
    V = x[lids[:, 0]] - x[lids[:, 1]]
    Q = C * V
    np.testing.assert_allclose(res["q"][lids[:, 0]], Q, rtol=6e-16, atol=0)
    np.testing.assert_allclose(res["q"][lids[:, 1]], -Q, rtol=6e-16, atol=0)
    np.testing.assert_allclose(res["nextSta"], Q, rtol=6e-16, atol=0)
    np.testing.assert_allclose(res["storeLeadQ"], Q, rtol=6e-16, atol=0)
    np.testing.assert_array_equal(res["storeLeadF"], 0.0)
    np.testing.assert_array_equal(res["f"], 0.0)
    t = p.topo
    for i in range(0, n, 7):
        a, b = lids[i]
        blk = [[res["dQdx"][_slot(t, r, c)] for c in (a, b)] for r in (a, b)]
        np.testing.assert_allclose(blk, C[i] * np.array([[1, -1], [-1, 1]]), rtol=6e-16)


def test_capacitor_ic_branch():
    """IC=: at the operating point the branch holds V(pos) - V(neg) = IC and
    its current enters the terminals; in transient the branch equation is
    i_bra = 0 and the capacitor is a plain charge."""
    n = 50
    spec = capacitor_spec(n, seed=405, ic=True)
    lead = np.ones(n, np.int32)
    p = _pair("capacitor", spec, lead=lead)
    lids = p.topo.node_lids.reshape(-1, 3)
    C = p.orc.read("C")
    IC = spec["inst_params"]["IC"]
    t = p.topo
    x, res = _drive(p, np.random.default_rng(406), newtonIter=1, transientFlag=0, dcopFlag=1)
    vp, vn, ib = x[lids[:, 0]], x[lids[:, 1]], x[lids[:, 2]]
    np.testing.assert_array_equal(res["f"][lids[:, 0]], ib)
    np.testing.assert_array_equal(res["f"][lids[:, 1]], -ib)
    np.testing.assert_allclose(res["f"][lids[:, 2]], vp - vn - IC, rtol=1e-15, atol=1e-16)
    np.testing.assert_allclose(res["q"][lids[:, 0]], C * IC, rtol=6e-16)
    np.testing.assert_allclose(res["nextSta"], C * IC, rtol=6e-16)
    np.testing.assert_array_equal(res["storeLeadF"], 0.0)
    for i in range(0, n, 5):
        a, b, br = lids[i]
        assert res["dFdx"][_slot(t, a, br)] == 1.0 and res["dFdx"][_slot(t, b, br)] == -1.0
        assert res["dFdx"][_slot(t, br, a)] == 1.0 and res["dFdx"][_slot(t, br, b)] == -1.0
        assert res["dFdx"][_slot(t, br, br)] == 0.0
        assert res["dQdx"][_slot(t, a, a)] == 0.0
    x, res = _drive(p, np.random.default_rng(407), newtonIter=1, transientFlag=1)
    vp, vn = x[lids[:, 0]], x[lids[:, 1]]
    np.testing.assert_array_equal(res["f"], 0.0)
    np.testing.assert_allclose(res["q"][lids[:, 0]], C * (vp - vn), rtol=6e-16)
    for i in range(0, n, 5):
        a, b, br = lids[i]
        assert res["dFdx"][_slot(t, br, br)] == 1.0 and res["dFdx"][_slot(t, a, br)] == 0.0
        assert res["dQdx"][_slot(t, a, a)] == C[i] and res["dQdx"][_slot(t, a, b)] == -C[i]


@pytest.mark.parametrize("model", ["resistor", "capacitor"])
def test_fd_jacobian_and_kcl(model):
    """Central differences of F (resistor) and Q (capacitor) against the
    loaded Jacobian, on a shared topology; the rows of each load sum to 0."""
    n = 40
    spec = resistor_spec(n, seed=408) if model == "resistor" else capacitor_spec(n, seed=408)
    p = Pair(model, spec, topology="shared", seed=409, gpu=False)
    key, jac = ("f", "dFdx") if model == "resistor" else ("q", "dQdx")
    rng = np.random.default_rng(410)
    x0 = rng.uniform(-0.5, 1.5, p.topo.n_nodes)
    p.set_inputs(nextSol=x0)
    res, _ = p.load(newtonIter=1, transientFlag=1)
    assert abs(res[key].sum()) <= 1e-12 * np.abs(res[key]).sum()
    t = p.topo
    J = np.zeros((t.n_nodes, t.n_nodes))
    for r in range(t.n_nodes):
        for s in range(t.row_ptr[r], t.row_ptr[r + 1]):
            J[r, t.col_idx[s]] = res[jac][s]
    h = 1e-3
    for c in range(t.n_nodes):
        xp, xm = x0.copy(), x0.copy()
        xp[c] += h
        xm[c] -= h
        p.set_inputs(nextSol=xp)
        fp = p.load(newtonIter=1, transientFlag=1)[0][key]
        p.set_inputs(nextSol=xm)
        fm = p.load(newtonIter=1, transientFlag=1)[0][key]
        fd = (fp - fm) / (2 * h)
        scale = np.abs(J[:, c]).max() + 1e-300
        assert np.abs(fd - J[:, c]).max() <= 1e-9 * scale, c
