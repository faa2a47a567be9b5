"""Pins of the Gummel-Poon BJT oracle restatement, CPU only.

Central-difference Jacobians over the parity cards (junctions inside the
exponential branch, where the reference differentiates exactly); the two
places where the reference's Jacobian is approximate -- the chord
conductance of a junction biased below -5 vt and the base resistance whose
bias dependence is not differentiated; forward-active collector current with
the Early effect in closed form; saturation-current and beta temperature laws;
an alias name (VA) acting as its primary (VAF).
"""
import numpy as np
import pytest

import oracle
from fdcheck import fd_check
from harness import Pair, splitmix64
from specs import BJT_CARDS, bjt_spec

TRAN = dict(newtonIter=1, transientFlag=1, pdt=2e9, currTimeStep=1e-10, lastTimeStep=1e-10)
DCOP = dict(newtonIter=1, dcopFlag=1)
KOQ = 1.3806226e-23 / 1.6021918e-19


def _pair(spec, **opts):
    o = dict(voltageLimiterFlag=0)
    o.update(opts)
    return Pair("bjt", spec, gpu=False, options=oracle.default_options(**o))


def _small_bias(p, seed):
    """|node voltage| <= 60 mV: both junctions in the exponential branch."""
    return splitmix64(seed).uniform(-0.06, 0.06, p.topo.n_nodes)


@pytest.mark.parametrize("card", range(len(BJT_CARDS)))
def test_fd_jacobian(card):
    # base-current crowding (IRB) is left out: its bias dependence is not
    # differentiated by the reference (test_base_resistance_...)
    level, params = BJT_CARDS[card]
    params = {k: v for k, v in params.items() if k not in ("IRB", "RBM")}
    spec = bjt_spec(48, seed=7 + card, cards=[(level, params)], temp_spread=True)
    p = _pair(spec)
    x = _small_bias(p, 11 + card)
    p.set_inputs(nextSto=np.zeros(p.topo.n_sto))
    for flags in (TRAN, DCOP):
        bad = fd_check(p, x, flags)
        assert not bad, (flags, bad[:5])


def test_fd_catches_planted_jacobian_error():
    spec = bjt_spec(24, seed=5, cards=BJT_CARDS[:1])
    p = _pair(spec)
    load = p.load

    def planted(**kw):
        ro, rg = load(**kw)
        ro["dFdx"] = ro["dFdx"] * 1.001
        return ro, rg
    p.load = planted
    assert len(fd_check(p, _small_bias(p, 6), TRAN, vecs=("f",))) > 10


def test_reverse_chord_conductance():
    """Below -5 vt the junction current is linearised as i = g v with the
    chord g = -IS/v + gmin (N_DEV_BJT.C:2780-2786), not the derivative gmin,
    so the Jacobian departs from dF/dx there."""
    card = (1, {"TYPE": 1.0, "IS": 1e-9, "BF": 50.0})
    spec = bjt_spec(24, seed=8, cards=[card])
    p = _pair(spec)
    x = splitmix64(9).uniform(-0.06, 0.06, p.topo.n_nodes)
    offs = np.concatenate([[0], np.cumsum(p.counts[0] + p.counts[1])])[:-1]
    x[p.topo.node_lids[offs + 1]] = -0.5        # base well below emitter and collector
    assert fd_check(p, x, DCOP, vecs=("f",))


def test_base_resistance_modulation_not_differentiated():
    """With IRB the base resistance follows the base current, but only its
    dependence on vBB' is stamped (diBrdvCp = diBrdvEp = 0, N_DEV_BJT.C:
    3198-3201); with a constant base resistance the check passes.  Forward
    active bias (vBE 0.5..0.7 V, |vBC| < 0.1 V) keeps both junctions in the
    exponential branch."""
    crowd = (1, {"TYPE": 1.0, "IS": 1e-14, "BF": 50.0, "RB": 100.0, "IRB": 1e-6, "RBM": 5.0})
    const = (1, {"TYPE": 1.0, "IS": 1e-14, "BF": 50.0, "RB": 100.0})
    for card, approx in ((crowd, True), (const, False)):
        p = _pair(bjt_spec(24, seed=10, cards=[card]))
        rng = splitmix64(12)
        offs = np.concatenate([[0], np.cumsum(p.counts[0] + p.counts[1])])[:-1]
        lids = p.topo.node_lids
        n = len(offs)
        vb = rng.uniform(0.5, 0.7, n)
        x = np.zeros(p.topo.n_nodes)
        x[lids[offs + 1]] = vb                                   # base
        x[lids[offs + 4]] = vb - rng.uniform(0.0, 1e-3, n)       # base' (RB drop)
        x[lids[offs + 0]] = vb + rng.uniform(-0.1, 0.1, n)       # collector
        assert bool(fd_check(p, x, DCOP, vecs=("f",))) == approx


@pytest.mark.parametrize("dtype", [1.0, -1.0])
def test_forward_active_early_effect(dtype):
    """Emitter at 0, base at vBE, collector at vBE - vBC with vBC = -2 V:
    iC = (iBE - iBC) / qB - iBC / BR with iBE = IS (e^{vBE/vt} - 1) + gmin vBE,
    the chord iBC = -IS + gmin vBC, qB = 1 / (1 - vBC / VAF)."""
    n = 40
    isat, bf, br, vaf = 1e-15, 80.0, 2.0, 40.0
    card = (1, {"TYPE": dtype, "IS": isat, "BF": bf, "BR": br, "VAF": vaf, "TNOM": 27.0})
    spec = {"models": [card], "n": n, "model_index": np.zeros(n, np.int32),
            "inst_params": {"AREA": np.ones(n)}}
    p = _pair(spec)
    rng = splitmix64(13)
    vbe = rng.uniform(0.3, 0.75, n)
    vbc = -2.0
    offs = np.concatenate([[0], np.cumsum(p.counts[0] + p.counts[1])])[:-1]
    lids = p.topo.node_lids
    x = np.zeros(p.topo.n_nodes)
    x[lids[offs + 1]] = dtype * vbe
    x[lids[offs + 0]] = dtype * (vbe - vbc)
    p.set_inputs(nextSol=x)
    ro, _ = p.load(**DCOP)
    vt = KOQ * (27.0 + 273.15)
    gmin = 1e-12
    ibe = isat * np.expm1(vbe / vt) + gmin * vbe
    ibc = -isat + gmin * vbc
    qb = 1.0 / (1.0 - vbc / vaf)
    ic = (ibe - ibc) / qb - ibc / br
    np.testing.assert_allclose(ro["f"][lids[offs + 0]], dtype * ic, rtol=1e-10)
    ib = ibe / bf + ibc / br
    np.testing.assert_allclose(ro["f"][lids[offs + 1]], -dtype * ib * -1.0, rtol=1e-10)


def test_temperature_laws():
    """IS(T) = IS exp((T/TNOM - 1) EG/vt + XTI ln(T/TNOM)), BF(T) = BF (T/TNOM)^XTB."""
    n = 30
    card = (1, {"TYPE": 1.0, "IS": 1e-16, "BF": 100.0, "EG": 1.15, "XTI": 3.5, "XTB": 1.5,
                "TNOM": 30.0})
    spec = bjt_spec(n, seed=14, cards=[card], temp_spread=True)
    p = _pair(spec)
    T = spec["inst_params"]["TEMP"] + 273.15
    Tn = 30.0 + 273.15
    vt = KOQ * T
    want_is = 1e-16 * np.exp((T / Tn - 1.0) * 1.15 / vt + 3.5 * np.log(T / Tn))
    np.testing.assert_allclose(p.orc.read("tSatCur"), want_is, rtol=1e-12)
    np.testing.assert_allclose(p.orc.read("tBetaF"), 100.0 * (T / Tn) ** 1.5, rtol=1e-12)


def test_alias_name_acts_as_primary():
    """VA sets the same forward Early voltage the evaluation uses as VAF."""
    base = {"TYPE": 1.0, "IS": 1e-16, "BF": 100.0, "CJE": 1e-12, "TF": 1e-10}
    pa = _pair(bjt_spec(30, seed=15, cards=[(1, dict(base, VAF=45.0))]))
    pb = _pair(bjt_spec(30, seed=15, cards=[(1, dict(base, VA=45.0))]))
    x = splitmix64(16).uniform(-0.2, 0.8, pa.topo.n_nodes)
    pa.set_inputs(nextSol=x)
    pb.set_inputs(nextSol=x)
    ra, _ = pa.load(**TRAN)
    rb, _ = pb.load(**TRAN)
    for k in ("f", "q", "dFdx", "dQdx"):
        np.testing.assert_array_equal(ra[k], rb[k], err_msg=k)


# ---- the new-DAE excess phase (DeviceOptions newExcessPhase = 1) ----------
# Two internal unknowns per instance, the phase-shifted transfer current i_fx
# and its time derivative (N_DEV_BJT.C:1032-1036, :1149-1153); the collector
# current takes i_fx in transient (auxDAECalculations :1928-1980).
EP_CARD = BJT_CARDS[3]          # PTF = 30 degrees, TF = 0.3 ns


def _ep_rows(p):
    """(Ifx, dIfx) LIDs of every instance: its last two internal nodes."""
    nodes_per = p.counts[0] + p.counts[1]
    ends = np.cumsum(nodes_per)
    return p.topo.node_lids[ends - 2], p.topo.node_lids[ends - 1]


def _ep_pair(n=32, seed=21, card=EP_CARD):
    p = _pair(bjt_spec(n, seed=seed, cards=[card], temp_spread=True), newExcessPhase=1)
    p.set_inputs(nextSto=np.zeros(p.topo.n_sto))
    return p


def test_new_excess_phase_fd_jacobian():
    p = _ep_pair()
    x = _small_bias(p, 31)
    ifx, difx = _ep_rows(p)
    assert not fd_check(p, x, TRAN), "transient Jacobian"
    # at the operating point the reference keeps a unit dIfx diagonal while
    # the dIfx row of F is zero, and Q's Ifx row (i_fx) has no dQ/dx stamp:
    # those two rows are the reference's own inconsistency, the rest is exact
    ex_f = np.zeros(p.topo.n_nodes, bool)
    ex_f[difx] = True
    ex_q = np.zeros(p.topo.n_nodes, bool)
    ex_q[ifx] = True
    assert not fd_check(p, x, DCOP, vecs=("f",), exclude=ex_f)
    assert not fd_check(p, x, DCOP, vecs=("q",), exclude=ex_q)


def test_new_excess_phase_filter_is_second_order_bessel():
    """The Ifx/dIfx block of the DAE is the excess-phase filter
    H(s) = 3 / (td^2 s^2 + 3 td s + 3), td = PTF (radians) * TF: the
    determinant of s dQ/dx + dF/dx over the block has those coefficients."""
    p = _ep_pair(n=4)
    x = _small_bias(p, 32)
    p.set_inputs(nextSol=x)
    ro, _ = p.load(**TRAN)
    ifx, difx = _ep_rows(p)
    rp, ci = p.topo.row_ptr, p.topo.col_idx
    td = EP_CARD[1]["PTF"] * np.pi / 180.0 * EP_CARD[1]["TF"]

    def entry(vals, r, c):
        k = rp[r] + np.nonzero(ci[rp[r]:rp[r + 1]] == c)[0]
        return float(vals[k].sum()) if len(k) else 0.0
    for a, b in zip(ifx, difx):
        G = np.array([[entry(ro["dFdx"], a, a), entry(ro["dFdx"], a, b)],
                      [entry(ro["dFdx"], b, a), entry(ro["dFdx"], b, b)]])
        Cm = np.array([[entry(ro["dQdx"], a, a), entry(ro["dQdx"], a, b)],
                       [entry(ro["dQdx"], b, a), entry(ro["dQdx"], b, b)]])
        # det(s C + G) = c2 s^2 + c1 s + c0
        c2 = np.linalg.det(Cm)
        c0 = np.linalg.det(G)
        c1 = Cm[0, 0] * G[1, 1] + G[0, 0] * Cm[1, 1] - Cm[0, 1] * G[1, 0] - G[0, 1] * Cm[1, 0]
        np.testing.assert_allclose([c2, c1, c0], [td * td, 3 * td, 3.0], rtol=1e-12)


def test_new_excess_phase_steady_state_is_the_dc_current():
    """With i_fx at its steady state iBE/qB and di_fx = 0, the transient
    loads carry the operating point's terminal currents, and both excess
    rows vanish."""
    p = _ep_pair(n=16)
    x = _small_bias(p, 33)
    ifx, difx = _ep_rows(p)
    p.set_inputs(nextSol=x)
    p.load(**DCOP)
    iBE = p.orc.read("iBE")
    qB = p.orc.read("qB")
    x[ifx] = iBE / qB
    x[difx] = 0.0
    p.set_inputs(nextSol=x)
    dc, _ = p.load(**DCOP)
    tr, _ = p.load(**TRAN)
    mask = np.ones(p.topo.n_nodes, bool)
    mask[ifx] = mask[difx] = False
    np.testing.assert_allclose(tr["f"][mask], dc["f"][mask], rtol=1e-12,
                               atol=1e-12 * np.max(np.abs(dc["f"][mask])))
    scale = np.abs(iBE / qB).max()
    assert np.max(np.abs(tr["f"][ifx])) == 0.0
    assert np.max(np.abs(tr["f"][difx])) <= 1e-12 * 3 * scale


def test_new_excess_phase_off_without_ptf():
    """PTF = 0 (td = 0): the excess rows are identities (F = i_fx, di_fx;
    unit Jacobian diagonal) and the terminal currents are the card's
    without excess phase."""
    card = (1, {k: v for k, v in EP_CARD[1].items() if k != "PTF"})
    p = _ep_pair(n=16, card=card)
    q = _pair(bjt_spec(16, seed=21, cards=[card], temp_spread=True))
    q.set_inputs(nextSto=np.zeros(q.topo.n_sto))
    x = _small_bias(p, 34)
    ifx, difx = _ep_rows(p)
    keep = np.ones(p.topo.n_nodes, bool)
    keep[ifx] = keep[difx] = False
    p.set_inputs(nextSol=x)
    q.set_inputs(nextSol=x[keep])
    rp_, _ = p.load(**TRAN)
    rq_, _ = q.load(**TRAN)
    np.testing.assert_array_equal(rp_["f"][ifx], x[ifx])
    np.testing.assert_array_equal(rp_["f"][difx], x[difx])
    np.testing.assert_array_equal(rp_["f"][keep], rq_["f"])
