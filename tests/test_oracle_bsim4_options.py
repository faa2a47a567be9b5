"""Pins of the BSIM4 option branches of the oracle restatement, CPU only.

Central-difference Jacobian checks (the reference's TESTJAC check) on the
option cards whose stamped Jacobian is the derivative of their F and Q, and
laws in closed form, each from the model's defining equations
(N_DEV_MOSFET_B4.C line ranges cited per test):

* source/drain edge tunnelling Igs(vgs), Igd(vgd) (:10410-10476);
* the channel tunnelling current split at Vds = 0, Igcs = Igcd = Igc/2
  (:10346-10402: Pigcd * Vdseff vanishes);
* bias-dependent source/drain resistance Rs(V), Rd(V) and the series
  conductance 1/(1/gs + Rs) (:9991-10062);
* capMod 0 accumulation and depletion gate charges (:10718-10777);
* the gate electrode current through grgeltd (rgateMod 1, :11827-11831) and
  grgeltd's geometry law (:7319-7329);
* the body resistance network currents (rbodyMod 1, :11909-11916) and
  grbdb = GBMIN + 1/RBDB (:7273-7274);
* the NQS time constant gtau = gcrg / CoxWL * 1e-9 (:11626-11637).

Reference behaviour the checks expose, kept by the oracle:

* capMod 0 differentiates its charges in Vbseff without the Vbseff(Vbs)
  smoothing chain (cgsb/cbsb use dVbseff_dVb = 1 from the DC Vbseff's
  regime), so its body-side columns are approximate;
* the gate tunnelling Jacobian applies dVbseff_dVb twice on the Pigcd path
  and differentiates the accumulation/depletion split of Voxdepinv
  piecewise: its gate row is approximate (other rows exact);
* trnqsMod: the F vector carries no charge-node current while dF/dx carries
  the NQS conductances (the reference marks NQS as not implemented,
  :17106-17109).
"""
import numpy as np
import pytest

import oracle
from fdcheck import fd_check
from harness import Pair, splitmix64
from specs import BSIM4_OPT_CARDS, bsim4_spec

TRAN = dict(newtonIter=1, transientFlag=1)
DCOP = dict(newtonIter=1, dcopFlag=1)
EXACT = ["rdsp", "cap1x0", "cap1x1", "cap1x5", "rg1", "rg2", "rg3", "rb1", "rb2", "acnqs", "mtrl"]


def _pair(spec, **opts):
    o = dict(voltageLimiterFlag=0)
    o.update(opts)
    return Pair("bsim4", spec, gpu=False, options=oracle.default_options(**o))


@pytest.mark.parametrize("name", EXACT)
@pytest.mark.parametrize("geometry", [False, True])
def test_fd_jacobian_options(name, geometry):
    spec = bsim4_spec(40, seed=7 + EXACT.index(name), cards=[BSIM4_OPT_CARDS[name]],
                      geometry=geometry, nf_mix=geometry, temp_spread=geometry)
    p = _pair(spec)
    x = splitmix64(11 + EXACT.index(name)).uniform(-0.5, 1.5, p.topo.n_nodes)
    bad = fd_check(p, x, TRAN)
    assert not bad, bad[:5]


def test_fd_catches_planted_error_in_option_rows():
    """A 0.1% error planted in the body-network rows of dF/dx is caught."""
    spec = bsim4_spec(24, seed=5, cards=[BSIM4_OPT_CARDS["rb1"]])
    p = _pair(spec)
    load = p.load

    def planted(**kw):
        ro, rg = load(**kw)
        ro["dFdx"] = ro["dFdx"] * 1.001
        return ro, rg
    p.load = planted
    x = splitmix64(6).uniform(-0.5, 1.5, p.topo.n_nodes)
    assert len(fd_check(p, x, TRAN, vecs=("f",))) > 20


@pytest.mark.parametrize("name", ["igc1", "igc2", "igc2b", "rds", "mtrlp"])
def test_fd_tunnelling_cards_without_tunnelling(name):
    """The gate tunnelling cards with igcMod = igbMod = 0 (thin oxide,
    tempMod 2, edge flat-band offset, bias-dependent Rs/Rd) have exact
    Jacobians: the cards' differences from central differences come from
    the reference's approximate tunnelling Jacobian alone, which the
    closed-form tunnelling pins below check instead."""
    card = (54, dict(BSIM4_OPT_CARDS[name][1], IGCMOD=0.0, IGBMOD=0.0))
    spec = bsim4_spec(40, seed=17, cards=[card], geometry=False)
    p = _pair(spec)
    x = splitmix64(19).uniform(-0.5, 1.5, p.topo.n_nodes)
    bad = fd_check(p, x, TRAN)
    assert not bad, bad[:5]


def _terminals(p, vd, vg, vs, vb, prime=True):
    """Private topology: terminals D, G, S, B; drain/source primes (when the
    card creates them) at their terminal's voltage."""
    x = np.zeros(p.topo.n_nodes)
    nodes_per = p.counts[0] + p.counts[1]
    offs = np.concatenate([[0], np.cumsum(nodes_per)])[:-1]
    lids = p.topo.node_lids
    for k, v in enumerate((vd, vg, vs, vb)):
        x[lids[offs + k]] = v
    if prime:
        has = nodes_per >= 6
        x[lids[offs[has] + 4]] = np.broadcast_to(vd, offs.shape)[has]
        x[lids[offs[has] + 5]] = np.broadcast_to(vs, offs.shape)[has]
    p.set_inputs(nextSol=x)
    return offs, lids


def _card(name, **extra):
    return (54, dict(BSIM4_OPT_CARDS[name][1], **extra))


def test_edge_tunnelling_closed_form():
    """Igs = AechvbEdgeS vgs Vge exp(BechvbEdge (AIGS + (AIGS CIGS - BIGS) Vge
    - BIGS CIGS Vge^2)), Vge = sqrt((vgs - VFBSD - VFBSDOFF)^2 + 1e-4), and
    the same law for Igd in vgd; the AIGS/BIGS/CIGS defaults are the NMOS
    ones (0.43, 0.054, 0.075) and VFBSDOFF = 0."""
    n = 48
    rng = np.random.default_rng(3)
    spec = bsim4_spec(n, seed=3, cards=[_card("igc1", IGBMOD=0.0)], geometry=False)
    p = _pair(spec)
    vg = rng.uniform(0.2, 1.4, n)
    vd = rng.uniform(0.0, 1.0, n)
    _terminals(p, vd, vg, 0.0, 0.0)
    p.load(**DCOP)
    r = p.orc.read
    vfbsd = r("vfbsd")
    aigs, bigs, cigs = 0.43, 0.054, 0.075
    for v, pre, got in ((vg, r("AechvbEdgeS"), r("Igs")), (vg - vd, r("AechvbEdgeD"), r("Igd"))):
        vge = np.sqrt((v - vfbsd) ** 2 + 1e-4)
        t5 = r("BechvbEdge") * (aigs + (aigs * cigs - bigs) * vge - bigs * cigs * vge * vge)
        want = pre * v * vge * np.exp(np.clip(t5, -34.0, 34.0))
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=0)


def test_channel_tunnelling_split_at_zero_vds():
    """At Vds = 0, Vdseff = 0 makes T7 = -Pigcd Vdseff = 0, so Igcs and Igcd
    are Igc (1e-4) / (2e-4) each: equal halves, for PIGCD given or not."""
    n = 32
    rng = np.random.default_rng(4)
    for extra in ({}, {"PIGCD": 1.5}):
        spec = bsim4_spec(n, seed=4, cards=[_card("igc1", **extra)], geometry=False)
        p = _pair(spec)
        v = rng.uniform(0.3, 1.4, n)
        _terminals(p, 0.0, v, 0.0, 0.0)
        p.load(**DCOP)
        igcs, igcd = p.orc.read("Igcs"), p.orc.read("Igcd")
        assert np.all(igcs > 0)
        np.testing.assert_array_equal(igcs, igcd)


@pytest.mark.parametrize("dtype", [1.0, -1.0])
def test_bias_dependent_series_resistance(dtype):
    """rdsMod 1: gstot = gs / (1 + gs Rs), Rs = RSWMIN' + RS0'/2 (T2 +
    sqrt(T2^2 + 0.01)), T2 = 1 / (1 + PRWG vgs_eff) - PRWB vbs, vgs_eff =
    (T0 + sqrt(T0^2 + 1e-4)) / 2, T0 = vgs - VFBSD; likewise gdtot in vgd,
    vbd.  Terminal voltages in the device's type convention."""
    n = 40
    rng = np.random.default_rng(5)
    card = (54, dict(BSIM4_OPT_CARDS["rdsp"][1], TYPE=dtype))
    spec = bsim4_spec(n, seed=5, cards=[card], geometry=True)
    p = _pair(spec)
    vg, vd, vb = rng.uniform(-0.2, 1.2, n), rng.uniform(0.0, 1.0, n), rng.uniform(-0.4, 0.0, n)
    _terminals(p, dtype * vd, dtype * vg, 0.0, dtype * vb)
    p.load(**DCOP)
    r = p.orc.read
    prwg, prwb = 0.3, 0.05

    def series(vgx, vbx, g, r0, rmin):
        t0 = vgx - r("vfbsd")
        vge = 0.5 * (t0 + np.sqrt(t0 * t0 + 1e-4))
        t2 = 1.0 / (1.0 + prwg * vge) - prwb * vbx
        rs = rmin + 0.5 * r0 * (t2 + np.sqrt(t2 * t2 + 0.01))
        return g / (1.0 + g * rs)
    np.testing.assert_allclose(r("gstot"), series(vg, vb, r("sourceConductance"), r("rs0"), r("rswmin")),
                               rtol=1e-12)
    np.testing.assert_allclose(r("gdtot"), series(vg - vd, vb - vd, r("drainConductance"), r("rd0"),
                                                  r("rdwmin")), rtol=1e-12)


@pytest.mark.parametrize("region", ["accumulation", "depletion"])
def test_capmod0_gate_charge(region):
    """capMod 0 (no poly depletion, Vbs < 0 so Vbseff = Vbs): accumulation
    Qg = CoxWL (Vgs - Vbs - VFBCV); depletion Qg = CoxWL K1ox (sqrt(K1ox^2/4
    + Vgs - Vbs - VFBCV) - K1ox/2); the intrinsic bulk charge is -Qg and the
    drain charge 0 in both."""
    n = 32
    rng = np.random.default_rng(6)
    spec = bsim4_spec(n, seed=6, cards=[BSIM4_OPT_CARDS["cap0x0"]], geometry=False)
    p = _pair(spec)
    vb = rng.uniform(-0.6, -0.3, n)
    # VFBCV defaults to -1: accumulation below Vgs - Vbs = -1, depletion just above
    vg = vb + (rng.uniform(-1.6, -1.1, n) if region == "accumulation" else rng.uniform(-0.95, -0.7, n))
    _terminals(p, 0.05, vg, 0.0, vb)
    p.load(**TRAN)
    r = p.orc.read
    arg1 = vg - vb - r("vfbcv")
    if region == "accumulation":
        want = r("CoxWL") * arg1
    else:
        k = r("k1ox")
        want = r("CoxWL") * k * (np.sqrt(0.25 * k * k + arg1) - 0.5 * k)
    qg_intrinsic = r("qgate") - r("qgdo") - r("qgso") - r("qgb")
    np.testing.assert_allclose(qg_intrinsic, want, rtol=1e-12, atol=1e-30)


def test_gate_electrode_current():
    """rgateMod 1: Igate = grgeltd (Vge - Vgp) with grgeltd = 1 / (RSHG (XGW
    + Weff_CJ / 3 / NGCON) / (NGCON NF (L - XGL))) (XGW = XGL = 0, NGCON =
    NF = 1 by default)."""
    n = 24
    rng = np.random.default_rng(7)
    spec = bsim4_spec(n, seed=7, cards=[BSIM4_OPT_CARDS["rg1"]], geometry=False)
    p = _pair(spec)
    x = rng.uniform(-0.5, 1.5, p.topo.n_nodes)
    p.set_inputs(nextSol=x)
    p.load(**DCOP)
    r = p.orc.read
    nodes_per = p.counts[0] + p.counts[1]
    offs = np.concatenate([[0], np.cumsum(nodes_per)])[:-1]
    lids = p.topo.node_lids
    vge, vgp = x[lids[offs + 1]], x[lids[offs + 4]]
    L = spec["inst_params"]["L"]
    g = 1.0 / (5.0 * (r("weffCJ") / 3.0) / L)
    np.testing.assert_allclose(r("grgeltd"), g, rtol=1e-13)
    np.testing.assert_allclose(r("Igate"), g * (vge - vgp), rtol=1e-12, atol=1e-30)


def test_body_network_currents():
    """rbodyMod 1: grbdb = GBMIN + 1/RBDB; the drain-body and source-body
    currents are grbdb (Vdb - Vb) and grbsb (Vsb - Vb)."""
    n = 24
    spec = bsim4_spec(n, seed=8, cards=[BSIM4_OPT_CARDS["rb1"]], geometry=False)
    p = _pair(spec)
    x = np.random.default_rng(8).uniform(-0.5, 1.5, p.topo.n_nodes)
    p.set_inputs(nextSol=x)
    p.load(**DCOP)
    r = p.orc.read
    np.testing.assert_allclose(r("grbdb"), 1e-12 + 1.0 / 40.0, rtol=1e-15)
    np.testing.assert_allclose(r("grbsb"), 1e-12 + 1.0 / 50.0, rtol=1e-15)
    nodes_per = p.counts[0] + p.counts[1]
    offs = np.concatenate([[0], np.cumsum(nodes_per)])[:-1]
    lids = p.topo.node_lids
    # internal node order after D, G, S, B: body prime, source body, drain body
    vb, vsb, vdb = x[lids[offs + 3]], x[lids[offs + 5]], x[lids[offs + 6]]
    np.testing.assert_allclose(r("Idbb"), r("grbdb") * (vdb - vb), rtol=1e-12, atol=1e-30)
    np.testing.assert_allclose(r("Isbb"), r("grbsb") * (vsb - vb), rtol=1e-12, atol=1e-30)


def test_nqs_time_constant_and_missing_charge_current():
    """trnqsMod: gtau = XRCRG1 (XRCRG2 Vtm beta + Ids) / CoxWL x 1e-9 is
    linear in XRCRG1 (doubling it doubles gtau exactly) and positive in a
    transient call; the F vector has no charge-node row while dF/dx loads
    the NQS conductances on it (the reference's incomplete NQS)."""
    n = 16
    x = None
    gt = []
    for k1 in (12.0, 24.0):
        spec = bsim4_spec(n, seed=9, cards=[(54, dict(BSIM4_OPT_CARDS["nqs"][1], XRCRG1=k1))],
                          geometry=False)
        p = _pair(spec)
        if x is None:
            x = np.random.default_rng(9).uniform(-0.5, 1.5, p.topo.n_nodes)
        p.set_inputs(nextSol=x)
        ro, _ = p.load(**TRAN)
        gt.append(p.orc.read("gtau"))
    assert np.all(gt[0] > 0)
    np.testing.assert_array_equal(gt[1], 2.0 * gt[0])
    nodes_per = p.counts[0] + p.counts[1]
    offs = np.concatenate([[0], np.cumsum(nodes_per)])[:-1]
    qrow = p.topo.node_lids[offs + 4]
    assert np.all(ro["f"][qrow] == 0.0)
    rp, ci = p.topo.row_ptr, p.topo.col_idx
    diag = np.array([ro["dFdx"][rp[q] + np.nonzero(ci[rp[q]:rp[q + 1]] == q)[0][0]] for q in qrow])
    assert np.all(diag > 0.0)


@pytest.mark.parametrize("dtype", [1.0, -1.0])
def test_material_model_setup_laws(dtype):
    """mtrlMod 1 (:5703-5781, :6817-6846): Coxe = 3.9 eps0 / EOT; the
    substrate band gap Eg0 = BG0SUB - TBGASUB Tnom^2 / (Tnom + TBGBSUB) and
    intrinsic density ni = NI0SUB (Tnom/300.15)^1.5 exp((Eg(300.15) - Eg0) /
    (2 Vtm0)); the source/drain flat band VFBSD = PHIG - (EASUB + Eg0/2 -
    type min(Vtm0 ln(NSD/ni), Eg0/2)) with NSD = 1e20."""
    card = (54, dict(BSIM4_OPT_CARDS["mtrlp"][1], TYPE=dtype, IGCMOD=0.0))
    spec = bsim4_spec(8, seed=10, cards=[card], geometry=False)
    p = _pair(spec)
    p.set_inputs(nextSol=np.zeros(p.topo.n_nodes))
    p.load(**DCOP)
    r = p.orc.read
    eot, phig, easub, bg0, tbga, tbgb, ni0 = 1.5e-9, 4.9, 4.05, 1.1, 7.02e-4, 1108.0, 1.2e10
    np.testing.assert_array_equal(r("coxe"), 3.9 * 8.854214871e-12 / eot)
    tnom = 27.0 + 273.15
    vtm0 = 8.617087e-5 * tnom
    eg0 = bg0 - tbga * tnom * tnom / (tnom + tbgb)
    t0 = bg0 - tbga * 90090.0225 / (300.15 + tbgb)
    ni = ni0 * (tnom / 300.15) * np.sqrt(tnom / 300.15) * np.exp((t0 - eg0) / (2.0 * vtm0))
    t0 = min(vtm0 * np.log(1e20 / ni), 0.5 * eg0)
    want = phig - (easub + 0.5 * eg0 - dtype * t0)
    np.testing.assert_allclose(r("vfbsd"), want, rtol=1e-12)
