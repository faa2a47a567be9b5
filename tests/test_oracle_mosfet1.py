"""Pins of the Level-1 MOSFET oracle restatement, CPU only.

Central-difference Jacobians over every parity card with the body junction
reverse biased; the SPICE body-effect linearisation that makes the Jacobian
approximate under forward body bias; the Shichman-Hodges drain current in
cutoff, saturation and the linear region; the linearised reverse junction
current; the DC Meyer gate charge in saturation.
"""
import numpy as np
import pytest

import oracle
from fdcheck import fd_check
from harness import Pair, splitmix64
from specs import MOSFET1_CARDS, mosfet1_spec

TRAN = dict(newtonIter=1, transientFlag=1)
DCOP = dict(newtonIter=1, dcopFlag=1)
VT300 = 1.3806226e-23 * 300.15 / 1.6021918e-19


def _pair(spec, **opts):
    o = dict(voltageLimiterFlag=0)
    o.update(opts)
    return Pair("mosfet1", spec, gpu=False, options=oracle.default_options(**o))


def _terminals(p):
    offs = np.concatenate([[0], np.cumsum(p.counts[0] + p.counts[1])])[:-1]
    return offs, p.topo.node_lids


def _reverse_body(p, x, dtype):
    offs, lids = _terminals(p)
    x = x.copy()
    x[lids[offs + 3]] = -0.6 if dtype > 0 else 1.6
    return x


@pytest.mark.parametrize("card", range(len(MOSFET1_CARDS)))
@pytest.mark.parametrize("geometry", [False, True])
def test_fd_jacobian_f(card, geometry):
    spec = mosfet1_spec(48, seed=7 + card, cards=[MOSFET1_CARDS[card]], geometry=geometry,
                        temp_spread=geometry)
    p = _pair(spec)
    x = _reverse_body(p, splitmix64(11 + card).uniform(-0.5, 1.5, p.topo.n_nodes),
                      MOSFET1_CARDS[card][1]["TYPE"])
    for flags in (TRAN, DCOP):
        bad = fd_check(p, x, flags, vecs=("f",))
        assert not bad, (flags, bad[:5])


@pytest.mark.parametrize("card", range(len(MOSFET1_CARDS)))
def test_fd_jacobian_q(card):
    """Q against its stamped dQ/dx at a point whose stored voltages are the
    present ones (the Meyer charge is integrated from the stored point)."""
    spec = mosfet1_spec(48, seed=17 + card, cards=[MOSFET1_CARDS[card]])
    p = _pair(spec)
    x = splitmix64(19 + card).uniform(-0.5, 1.5, p.topo.n_nodes)
    p.set_inputs(nextSol=x)
    ro, _ = p.load(**TRAN)
    p.set_inputs(currSto=ro["nextSto"], currSta=ro["nextSta"])
    assert not fd_check(p, x, TRAN, vecs=("q",))


def test_forward_body_bias_jacobian_is_approximate():
    """For vbs > 0 the reference extends sqrt(phi - vbs) by its tangent but
    differentiates with gamma / (2 sarg) (N_DEV_MOSFET1.C:2922-2936), so the
    body-effect entries depart from the derivative of F there."""
    spec = mosfet1_spec(48, seed=7, cards=[MOSFET1_CARDS[0]], geometry=False)
    p = _pair(spec)
    offs, lids = _terminals(p)
    x = splitmix64(11).uniform(-0.5, 1.5, p.topo.n_nodes)
    x[lids[offs + 3]] = 1.6   # bulk above source and drain: forward body bias
    assert fd_check(p, x, TRAN, vecs=("f",))


def _one_card(card, n, w, l, **inst):
    params = {"W": w, "L": l}
    params.update(inst)
    return {"models": [(1, card)], "n": n, "model_index": np.zeros(n, np.int32),
            "inst_params": params}


@pytest.mark.parametrize("dtype", [1.0, -1.0])
def test_square_law(dtype):
    """GAMMA = 0, TEMP = TNOM: Id = KP W/L (1 + LAMBDA vds) vgst^2 / 2 in
    saturation, KP W/L (1 + LAMBDA vds) vds (vgst - vds/2) in the linear
    region, 0 in cutoff; the drain row also carries the reverse-biased
    drain junction, linearised as (IS/vt + gmin) vbd."""
    n = 60
    rng = splitmix64(23)
    w = 1e-6 * np.exp(0.2 * rng.standard_normal(n))
    l = 2e-7 * np.exp(0.2 * rng.standard_normal(n))
    kp, vto, lam, isat = 1e-4, 0.5, 0.05, 1e-14
    card = {"TYPE": dtype, "KP": kp, "VTO": dtype * vto, "LAMBDA": lam, "IS": isat, "TNOM": 27.0}
    p = _pair(_one_card(card, n, w, l))
    vgs = rng.uniform(0.0, 1.5, n)
    vds = rng.uniform(0.05, 1.5, n)
    offs, lids = _terminals(p)
    x = np.zeros(p.topo.n_nodes)
    x[lids[offs + 0]] = dtype * vds
    x[lids[offs + 1]] = dtype * vgs
    p.set_inputs(nextSol=x)
    ro, _ = p.load(**DCOP)
    vgst = vgs - vto
    beta = kp * w / l
    sat = vgst <= vds
    ids = np.where(vgst <= 0, 0.0,
                   np.where(sat, beta * (1 + lam * vds) * vgst ** 2 * 0.5,
                            beta * (1 + lam * vds) * vds * (vgst - 0.5 * vds)))
    cbd = (isat / VT300 + 1e-12) * (-vds)
    want = dtype * (ids - cbd)
    np.testing.assert_allclose(ro["f"][lids[offs + 0]], want, rtol=1e-9, atol=1e-22)
    assert (vgst <= 0).any() and (sat & (vgst > 0)).any() and (~sat).any()


def test_meyer_charge_dc_saturation():
    """DC, saturation: capgs = Cox/3, capgd = capgb = 0, so the gate row of Q
    is (2 Cox/3 + CGSO W) vgs + CGDO W vgd + CGBO L vgb with
    Cox = 3.9 eps0 / TOX * L * W."""
    n = 40
    rng = splitmix64(29)
    w = 1e-6 * np.exp(0.2 * rng.standard_normal(n))
    l = 2e-7 * np.exp(0.2 * rng.standard_normal(n))
    tox, cgso, cgdo, cgbo = 3e-9, 2e-10, 3e-10, 1e-10
    card = {"TYPE": 1.0, "KP": 1e-4, "VTO": 0.3, "TOX": tox, "CGSO": cgso, "CGDO": cgdo,
            "CGBO": cgbo, "TNOM": 27.0}
    p = _pair(_one_card(card, n, w, l))
    vgs = rng.uniform(0.5, 1.0, n)
    vds = vgs + rng.uniform(0.1, 0.5, n)        # vds > vgst: saturation
    offs, lids = _terminals(p)
    x = np.zeros(p.topo.n_nodes)
    x[lids[offs + 0]] = vds
    x[lids[offs + 1]] = vgs
    p.set_inputs(nextSol=x)
    ro, _ = p.load(**DCOP)
    cox = 3.9 * 8.854214871e-12 / tox * l * w
    want = (2 * cox / 3 + cgso * w) * vgs + cgdo * w * (vgs - vds) + cgbo * l * vgs
    np.testing.assert_allclose(ro["q"][lids[offs + 1]], want, rtol=1e-12, atol=1e-30)
