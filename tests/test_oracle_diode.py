"""Pins the diode oracle (oracle/or_diode.c) without the reference binary:
textbook closed forms, finite-difference Jacobians (the TESTJAC check of
NumericalJacobian::testDAEMatrices, N_DEV_NumericalJacobian.C:175-582, here
with central differences and tighter tolerances), KCL conservation, the
pnjlim special cases (N_DEV_DeviceSupport.C:214-252) and the temperature law
(N_DEV_Diode.C:1376-1593)."""
import numpy as np
import pytest

import oracle
from harness import Pair, splitmix64
from specs import DIODE_CARDS, diode_spec

KT_Q = 1.3806226e-23 / 1.6021918e-19
T0 = 300.15


def _pair(spec, **opt):
    o = oracle.default_options(**opt)
    return Pair("diode", spec, options=o, gpu=False)


def _spec(n, card, level=1, area=None):
    return {"models": [(level, card)], "n": n, "model_index": np.zeros(n, np.int32),
            "inst_params": {"AREA": area if area is not None else np.ones(n)}}


def _load(p, vpos, vneg, **flags):
    p.set_inputs(nextSol=_interleave(p, vpos, vneg))
    flags.setdefault("newtonIter", 1)
    flags.setdefault("transientFlag", 1)
    res, _ = p.load(**flags)
    return res


def _interleave(p, vpos, vneg):
    x = np.zeros(p.topo.n_nodes)
    lids = p.topo.node_lids.reshape(-1, 2) if p.counts[1].sum() == 0 else None
    assert lids is not None
    x[lids[:, 0]] = vpos
    x[lids[:, 1]] = vneg
    return x


def test_forward_region_closed_form():
    n = 64
    card = {"IS": 3e-14, "N": 1.2}
    p = _pair(_spec(n, card), voltageLimiterFlag=0)
    v = np.linspace(-0.05, 0.8, n)
    res = _load(p, v, np.zeros(n))
    lids = p.topo.node_lids.reshape(-1, 2)
    vte = 1.2 * KT_Q * T0
    Id = 3e-14 * (np.exp(v / vte) - 1.0) + 1e-12 * v
    np.testing.assert_allclose(-res["f"][lids[:, 1]], Id, rtol=1e-12, atol=1e-27)
    np.testing.assert_allclose(res["f"][lids[:, 0]], Id, rtol=1e-12, atol=1e-27)


def test_linear_reverse_and_breakdown_closed_form():
    n = 64
    card = {"IS": 1e-14, "BV": 2.0, "IBV": 1e-3}
    p = _pair(_spec(n, card), voltageLimiterFlag=0)
    vte = KT_Q * T0
    v = np.linspace(-1.2, -4 * vte, n)   # linear reverse region, above the knee (~1.35 V)
    res = _load(p, np.zeros(n), -v)
    lids = p.topo.node_lids.reshape(-1, 2)
    arg = (3.0 * vte / (v * np.e)) ** 3
    Id = -1e-14 * (1.0 + arg) + 1e-12 * v
    np.testing.assert_allclose(-res["f"][lids[:, 1]], Id, rtol=1e-12, atol=1e-27)
    # breakdown: the knee tBrkdwnV is placed so that the current at V = -BV
    # is -IBV (plus the gmin leak)
    tbv = p.orc.read("tBrkdwnV")
    assert np.all((tbv > 3 * vte) & (tbv <= 2.0))
    res = _load(p, np.zeros(n), np.full(n, 2.0))
    Idb = -res["f"][lids[:, 1]]
    np.testing.assert_allclose(Idb, -1e-3 - 2e-12, rtol=1e-9)


def test_depletion_charge_closed_form():
    n = 64
    cj, vj, m, tt = 2e-12, 0.8, 0.45, 3e-9
    card = {"IS": 1e-14, "CJO": cj, "VJ": vj, "M": m, "TT": tt, "FC": 0.5}
    p = _pair(_spec(n, card), voltageLimiterFlag=0)
    v = np.linspace(-2.0, 0.39, n)          # below FC*VJ
    res = _load(p, v, np.zeros(n))
    lids = p.topo.node_lids.reshape(-1, 2)
    vte = KT_Q * T0
    Id = np.where(v >= -3 * vte, 1e-14 * (np.exp(v / vte) - 1), -1e-14 * (1 + (3 * vte / (v * np.e)) ** 3)) + 1e-12 * v
    Qd = tt * Id + vj * cj * (1 - (1 - v / vj) ** (1 - m)) / (1 - m)
    np.testing.assert_allclose(res["q"][lids[:, 0]], Qd, rtol=1e-11, atol=1e-30)
    np.testing.assert_allclose(res["q"][lids[:, 1]], -Qd, rtol=1e-11, atol=1e-30)


def test_saturation_current_temperature_law():
    n = 16
    card = {"IS": 1e-14, "N": 1.1, "EG": 1.11, "XTI": 3.0}
    temps = np.linspace(-40.0, 125.0, n)
    spec = _spec(n, card)
    spec["inst_params"]["TEMP"] = temps
    p = _pair(spec)
    T = temps + 273.15
    vt = KT_Q * T
    want = 1e-14 * np.exp((T / T0 - 1.0) * 1.11 / (1.1 * vt) + 3.0 / 1.1 * np.log(T / T0))
    np.testing.assert_allclose(p.orc.read("tSatCur"), want, rtol=1e-13)


FD_CARDS = DIODE_CARDS[:4] + [
    # level 2 without IKF/ISR: Khi = 1, Kgen = 0, where its Jacobian is exact
    (2, {"IS": 1e-14, "RS": 2.0, "CJO": 1e-12, "TT": 2e-9, "BV": 3.0, "IBV": 1e-4,
         "TBV1": 1e-3, "TRS1": 2e-3, "TNOM": 25.0}),
]


def test_level2_conductance_quirk():
    """Level 2 with IKF or ISR: the reference's Gd is not dId/dVd.  Its
    high-injection term uses DKhi = +0.5*Khi*Gd1/(IKF+Inorm) (the derivative
    has a minus sign) and DKgen = -M*(1-Vd/VJ)/(VJ*Kgen) (N_DEV_Diode.C
    :1201-1214).  Pin the current by its closed form and the conductance by
    its exact offset from the true derivative."""
    n = 40
    vt = KT_Q * T0
    ikf, isat = 1e-2, 1e-14
    p = _pair(_spec(n, {"IS": isat, "IKF": ikf}, level=2), voltageLimiterFlag=0)
    v = np.linspace(0.3, 0.8, n)
    res = _load(p, v, np.zeros(n))
    lids = p.topo.node_lids.reshape(-1, 2)
    inorm = isat * (np.exp(v / vt) - 1) + 1e-12 * v
    gd1 = isat * np.exp(v / vt) / vt + 1e-12
    khi = np.sqrt(ikf / (ikf + inorm))
    np.testing.assert_allclose(-res["f"][lids[:, 1]], inorm * khi, rtol=1e-12)
    true_gd = gd1 * khi - inorm * 0.5 * khi * gd1 / (ikf + inorm)
    quirk = inorm * khi * gd1 / (ikf + inorm)          # twice the sign-flipped term
    gd = p.orc.read("Gd")
    np.testing.assert_allclose(gd, true_gd + quirk, rtol=1e-12)
    # recombination term
    m, vj, isr = 0.5, 1.0, 1e-12
    p = _pair(_spec(n, {"IS": isat, "ISR": isr, "NR": 2.0, "M": m, "VJ": vj}, level=2),
              voltageLimiterFlag=0)
    res = _load(p, v, np.zeros(n))
    irec = isr * (np.exp(v / (2 * vt)) - 1)
    pot = p.orc.read("tJctPot")
    kgen = (((1 - v / pot) ** 2 + 0.005) ** m) ** 0.5
    np.testing.assert_allclose(-res["f"][lids[:, 1]], inorm + irec * kgen, rtol=1e-12)
    gd2 = isr * np.exp(v / (2 * vt)) / (2 * vt)
    dkgen_ref = -m * (1 - v / pot) / (pot * kgen)
    np.testing.assert_allclose(p.orc.read("Gd"), gd1 + gd2 * kgen + irec * dkgen_ref, rtol=1e-11)


@pytest.mark.parametrize("card_idx", range(len(FD_CARDS)))
def test_fd_jacobian(card_idx):
    """Central-difference dF/dx and dQ/dx against the analytic stamps."""
    level, card = FD_CARDS[card_idx]
    n = 200
    rng = splitmix64(7 + card_idx)
    spec = {"models": [(level, card)], "n": n, "model_index": np.zeros(n, np.int32),
            "inst_params": {"AREA": np.exp(0.2 * rng.standard_normal(n))}}
    p = _pair(spec, voltageLimiterFlag=0)
    x0 = rng.uniform(-0.5, 1.5, p.topo.n_nodes)
    p.set_inputs(nextSol=x0)
    base, _ = p.load(newtonIter=1, transientFlag=1)
    J = {"f": base["dFdx"], "q": base["dQdx"]}
    rp, ci = p.topo.row_ptr, p.topo.col_idx
    nodes_per = p.counts[0] + p.counts[1]
    offs = np.concatenate([[0], np.cumsum(nodes_per)])
    maxk = int(nodes_per.max())
    for k in range(maxk):
        has = nodes_per > k
        cols = p.topo.node_lids[offs[:-1][has] + k]
        h = 1e-6 * (1 + np.abs(x0[cols]))
        xp = x0.copy(); xp[cols] += h
        xm = x0.copy(); xm[cols] -= h
        p.set_inputs(nextSol=xp); rp_, _ = p.load(newtonIter=1, transientFlag=1)
        p.set_inputs(nextSol=xm); rm_, _ = p.load(newtonIter=1, transientFlag=1)
        hx = np.zeros(p.topo.n_nodes); hx[cols] = h
        for vec in ("f", "q"):
            fd = (rp_[vec] - rm_[vec])
            # analytic column entries: for each row r, J[r, col] * 2h
            ana = np.zeros(p.topo.n_nodes)
            for r in range(p.topo.n_nodes):
                for s in range(rp[r], rp[r + 1]):
                    c = ci[s]
                    if hx[c] != 0.0:
                        ana[r] += J[vec][s] * 2 * hx[c]
            scale = np.maximum(np.abs(fd), np.abs(ana))
            # rounding of F itself bounds what a difference can resolve
            fmag = np.maximum(np.abs(rp_[vec]), np.abs(rm_[vec]))
            ok = np.abs(fd - ana) <= 1e-5 * scale + 1e-14 * fmag + 1e-22
            assert ok.all(), (vec, k, np.nonzero(~ok)[0][:5], fd[~ok][:5], ana[~ok][:5])


def test_kcl_conservation():
    """Each instance's F and Q contributions sum to zero over its terminals."""
    n = 500
    spec = diode_spec(n, seed=11)
    p = _pair(spec)
    rng = splitmix64(12)
    p.set_inputs(nextSol=rng.uniform(-0.5, 1.5, p.topo.n_nodes),
                 nextSto=rng.uniform(-0.5, 1.5, p.topo.n_sto))
    res, _ = p.load(newtonIter=1, transientFlag=1)
    nodes_per = p.counts[0] + p.counts[1]
    offs = np.concatenate([[0], np.cumsum(nodes_per)])
    for vec in ("f", "q", "dFdxdVp", "dQdxdVp"):
        sums = np.add.reduceat(res[vec][p.topo.node_lids], offs[:-1])
        mags = np.maximum.reduceat(np.abs(res[vec][p.topo.node_lids]), offs[:-1])
        assert np.all(np.abs(sums) <= 1e-12 * mags + 1e-300), vec


def test_pnjlim_special_cases():
    """No limiting when |dV| <= 2 vt or V <= vcrit; log damping otherwise."""
    n = 6
    p = _pair(_spec(n, {"IS": 1e-14}))
    vt = KT_Q * T0
    vcrit = p.orc.read("tVcrit")[0]
    vold = np.array([0.5, 0.5, 0.5, 0.0, -1.0, 0.7])
    vnew = np.array([0.5 + 1.5 * vt, 0.5 + 10 * vt, vcrit - 0.01, 1.0, 1.0, 0.6])
    # first call stores Vd member; emulate the previous iteration through sto
    p.set_inputs(nextSol=_interleave(p, vold, np.zeros(n)))
    p.load(newtonIter=1, transientFlag=1)   # sets member Vd = vold (no limiting from 0? see below)
    got_old = p.orc.read("Vd")
    p.set_inputs(nextSol=_interleave(p, vnew, np.zeros(n)))
    p.load(newtonIter=2, transientFlag=1)
    got = p.orc.read("Vd")
    for j in range(n):
        vo, vn = got_old[j], vnew[j]
        if vn > vcrit and abs(vn - vo) > 2 * vt:
            want = vo + vt * np.log(1 + (vn - vo) / vt) if vo > 0 else vt * np.log(vn / vt)
        else:
            want = vn
        assert got[j] == pytest.approx(want, rel=1e-14, abs=1e-15), j
