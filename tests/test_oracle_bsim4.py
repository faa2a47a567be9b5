"""Pins of the BSIM4 (level 54) oracle restatement, CPU only.

The oracle's stamped Jacobians are checked against central differences of its
own F and Q (the reference's TESTJAC check, N_DEV_NumericalJacobian.C:175-582)
on every card of the parity set; its junction currents against the diode
equation in closed form; its type handling by the NMOS/PMOS mirror identity;
its drain/source handling by the symmetry of a symmetric device.  A 0.1%
error planted in the Jacobian must be caught by the difference check.
"""
import numpy as np
import pytest

import oracle
from fdcheck import fd_check
from harness import Pair, splitmix64
from specs import BSIM4_CARDS, BSIM4_VTL_CARD, bsim4_spec

TRAN = dict(newtonIter=1, transientFlag=1)
DCOP = dict(newtonIter=1, dcopFlag=1)
NOLIM = dict(voltageLimiterFlag=0)


def _pair(spec, **opts):
    o = dict(NOLIM)
    o.update(opts)
    return Pair("bsim4", spec, gpu=False, options=oracle.default_options(**o))


def _fd(p, seed, flags, **kw):
    x = splitmix64(seed).uniform(-0.5, 1.5, p.topo.n_nodes)
    return fd_check(p, x, flags, **kw)


@pytest.mark.parametrize("card", range(len(BSIM4_CARDS)))
@pytest.mark.parametrize("geometry", [False, True])
def test_fd_jacobian(card, geometry):
    spec = bsim4_spec(48, seed=7 + card, cards=[BSIM4_CARDS[card]], geometry=geometry,
                      nf_mix=geometry, temp_spread=geometry)
    p = _pair(spec)
    bad = _fd(p, 11 + card, TRAN)
    assert not bad, bad[:5]


def test_fd_jacobian_dcop():
    spec = bsim4_spec(48, seed=3, cards=BSIM4_CARDS[:2])
    p = _pair(spec)
    assert not _fd(p, 4, DCOP)


def test_fd_catches_planted_jacobian_error():
    spec = bsim4_spec(24, seed=5, cards=BSIM4_CARDS[:2])
    p = _pair(spec)
    load = p.load

    def planted(**kw):
        ro, rg = load(**kw)
        ro["dFdx"] = ro["dFdx"] * 1.001
        return ro, rg
    p.load = planted
    assert len(_fd(p, 6, TRAN, vecs=("f",))) > 20


@pytest.mark.parametrize("extra,exact", [
    ({"LAMBDA": 0.2}, True),
    ({"MOBMOD": 2.0}, True),
    ({"VTL": 2e5}, False),
    ({"LAMBDA": 0.2, "MOBMOD": 2.0}, False),
])
def test_approximate_jacobian_options(extra, exact):
    """Two option sets where the reference's stamped Jacobian is not the
    derivative of its F: the source-end velocity limit differentiates
    Fsevl = (1 + (vs/vtl)^6)^(-1/6) holding Coxeff, Weff and the Vbseff chain
    fixed (N_DEV_MOSFET_B4.C:9905-9945), and the LAMBDA term chains
    dueff_dVd, which under mobMod 2 leaves out the Vgsteff path
    (:9300-9320, :9500-9540).  Each option alone otherwise passes.  GPU
    parity covers these cards exactly."""
    spec = bsim4_spec(48, seed=9, cards=[(54, dict(extra, TYPE=1.0))], geometry=False)
    bad = _fd(_pair(spec), 10, TRAN)
    assert (not bad) == exact, bad[:3]


def _set_terminals(p, vd, vg, vs, vb):
    """Private topology: local nodes 0..3 of each instance are D, G, S, B."""
    x = np.zeros(p.topo.n_nodes)
    offs = np.concatenate([[0], np.cumsum(p.counts[0] + p.counts[1])])[:-1]
    lids = p.topo.node_lids
    for k, v in enumerate((vd, vg, vs, vb)):
        x[lids[offs + k]] = v
    p.set_inputs(nextSol=x)
    return offs, lids


@pytest.mark.parametrize("dtype", [1.0, -1.0])
def test_junction_diode_closed_form(dtype):
    """D = G = S = 0, body at V: the channel carries nothing, so the body row
    is the two junction diodes: I = Is (exp(V/(NJ vt)) - 1) + gmin V each,
    with Is = AS * JSS (sidewall densities zero, TEMP = TNOM)."""
    n = 64
    rng = splitmix64(21)
    w = 90e-9 * np.exp(0.2 * rng.standard_normal(n))
    l = 45e-9 * np.exp(0.2 * rng.standard_normal(n))
    area_s = w * 100e-9
    area_d = w * 120e-9
    jss = 2e-4
    spec = {"models": [(54, {"TYPE": dtype, "JSS": jss, "JSWS": 0.0, "JSWGS": 0.0, "TNOM": 27.0})],
            "n": n, "model_index": np.zeros(n, np.int32),
            "inst_params": {"W": w, "L": l, "AS": area_s, "AD": area_d,
                            "PS": 2 * (w + 100e-9), "PD": 2 * (w + 120e-9)}}
    p = _pair(spec)
    v = rng.uniform(-0.3, 0.6, n)
    offs, lids = _set_terminals(p, 0.0, 0.0, 0.0, dtype * v)
    ro, _ = p.load(**TRAN)
    vt = 8.617087e-5 * (27.0 + 273.15)
    gmin = 1e-12
    diode = lambda a: a * jss * np.expm1(v / vt) + gmin * v
    want_b = dtype * (diode(area_s) + diode(area_d))
    got_b = ro["f"][lids[offs + 3]]
    np.testing.assert_allclose(got_b, want_b, rtol=1e-9, atol=1e-24)
    np.testing.assert_allclose(ro["f"][lids[offs + 1]], 0.0, atol=1e-30)
    np.testing.assert_allclose(ro["f"][lids[offs + 0]], -dtype * diode(area_d), rtol=1e-9, atol=1e-24)


def test_nmos_pmos_mirror():
    """A PMOS card equal to an NMOS card (type-dependent defaults given) with
    every voltage negated loads -F, -Q and the same Jacobians.  Mobility model
    2 is left out: its vtfbphi1 is 2x or 2.5x the flat-band distance by type
    (N_DEV_MOSFET_B4.C:7190-7199)."""
    common = {"U0": 0.03, "EU": 1.67, "VDDEOT": 1.5, "JSS": 1e-4, "CGSL": 1e-11}
    n = 40
    nspec = bsim4_spec(n, seed=31, cards=[(54, dict(common, TYPE=1.0, VTH0=0.4))])
    pspec = bsim4_spec(n, seed=31, cards=[(54, dict(common, TYPE=-1.0, VTH0=-0.4))])
    pspec["inst_params"]["DELVTO"] = -pspec["inst_params"]["DELVTO"]   # a signed shift of VTH0
    pn, pp = _pair(nspec), _pair(pspec)
    x = splitmix64(32).uniform(-0.5, 1.5, pn.topo.n_nodes)
    pn.set_inputs(nextSol=x)
    pp.set_inputs(nextSol=-x)
    rn, _ = pn.load(**TRAN)
    rp, _ = pp.load(**TRAN)
    for k in ("f", "q"):
        np.testing.assert_allclose(rp[k], -rn[k], rtol=1e-12, atol=1e-300, err_msg=k)
    for k in ("dFdx", "dQdx"):
        np.testing.assert_allclose(rp[k], rn[k], rtol=1e-12, atol=1e-300, err_msg=k)


def test_drain_source_symmetry():
    """With equal drain and source geometry, exchanging the drain and source
    voltages exchanges the drain and source rows of F."""
    n = 40
    rng = splitmix64(41)
    w = 90e-9 * np.exp(0.2 * rng.standard_normal(n))
    l = 45e-9 * np.exp(0.2 * rng.standard_normal(n))
    geo = {"W": w, "L": l, "AS": w * 1e-7, "AD": w * 1e-7, "PS": 2 * (w + 1e-7), "PD": 2 * (w + 1e-7)}
    spec = {"models": [(54, {"TYPE": 1.0})], "n": n, "model_index": np.zeros(n, np.int32),
            "inst_params": geo}
    p = _pair(spec)
    vd, vg, vs, vb = (rng.uniform(-0.5, 1.5, n) for _ in range(4))
    offs, lids = _set_terminals(p, vd, vg, vs, vb)
    r1, _ = p.load(**DCOP)
    _set_terminals(p, vs, vg, vd, vb)
    r2, _ = p.load(**DCOP)
    d, s = lids[offs + 0], lids[offs + 2]
    scale = np.maximum(np.abs(r1["f"][d]), np.abs(r1["f"][s])) + 1e-18
    np.testing.assert_array_less(np.abs(r2["f"][s] - r1["f"][d]) / scale, 1e-9)
    np.testing.assert_array_less(np.abs(r2["f"][d] - r1["f"][s]) / scale, 1e-9)


@pytest.mark.parametrize("bin_unit", [1, 0])
def test_binning_closed_form(bin_unit):
    """Size-dependent parameter binning (the reference's size-dependent block
    in updateTemperature, N_DEV_MOSFET_B4.C:5640-7734): P = P0 + PL/Leff +
    PW/Weff + PP/(Leff Weff), lengths in microns with BINUNIT = 1; here VTH0
    and its L/W/P coefficients, plus an instance DELVTO.  With LINT = WINT =
    LL = WW = XL = XW = 0 and NF = 1, Leff = L and Weff = W."""
    n = 32
    rng = splitmix64(31)
    w = 90e-9 * np.exp(0.3 * rng.standard_normal(n))
    l = 45e-9 * np.exp(0.3 * rng.standard_normal(n))
    dv = rng.uniform(-0.05, 0.05, n)
    s = 1e-6 if bin_unit == 1 else 1.0
    card = {"TYPE": 1.0, "VTH0": 0.4, "LVTH0": 0.01 / s, "WVTH0": 0.02 / s,
            "PVTH0": 0.003 / (s * s), "BINUNIT": float(bin_unit)}
    spec = {"models": [(54, card)], "n": n, "model_index": np.zeros(n, np.int32),
            "inst_params": {"W": w, "L": l, "DELVTO": dv}}
    p = _pair(spec)
    np.testing.assert_allclose(p.orc.read("leff"), l, rtol=1e-15)
    np.testing.assert_allclose(p.orc.read("weff"), w, rtol=1e-15)
    want = 0.4 + (0.01 / s) * (s / l) + (0.02 / s) * (s / w) + (0.003 / (s * s)) * (s * s / (l * w)) + dv
    np.testing.assert_allclose(p.orc.read("vth0"), want, rtol=1e-14)


@pytest.mark.parametrize("u0,up", [(0.03, 0.0), (400.0, 0.1)])
def test_mobility_temperature_law(u0, up):
    """tempMod 0 low-field mobility: u0temp = U0 (1 - UP exp(-Leff/LP))
    (T/TNOM)^UTE, U0 above 1 taken in cm^2/Vs (divided by 1e4), at an
    instance TEMP of 85 C against TNOM = 27 C."""
    n = 16
    rng = splitmix64(32)
    w = 90e-9 * np.exp(0.2 * rng.standard_normal(n))
    l = 45e-9 * np.exp(0.2 * rng.standard_normal(n))
    card = {"TYPE": 1.0, "U0": u0, "UTE": -1.3, "UP": up, "LP": 1e-8, "TNOM": 27.0}
    spec = {"models": [(54, card)], "n": n, "model_index": np.zeros(n, np.int32),
            "inst_params": {"W": w, "L": l, "TEMP": np.full(n, 85.0)}}
    p = _pair(spec)
    u = u0 / 1e4 if u0 > 1.0 else u0
    want = u * (1.0 - up * np.exp(-l / 1e-8)) * ((85.0 + 273.15) / (27.0 + 273.15)) ** -1.3
    np.testing.assert_allclose(p.orc.read("u0temp"), want, rtol=1e-13)


@pytest.mark.parametrize("temp_mod", [0, 1, 2])
def test_temperature_mode_laws(temp_mod):
    """tempMod 0 scales UA..UD, VSAT and RDSW additively in T/TNOM - 1;
    tempMod 1 and 2 multiplicatively in T - TNOM; tempMod 2 takes the
    threshold's temperature term as -KT1 (T/TNOM - 1) instead of
    (KT1 + KT1L/Leff)(T/TNOM - 1) (N_DEV_MOSFET_B4.C:6700-6740, 7050-7055):
    the flat-band factor of tempMod 1 exceeds tempMod 2's by exactly
    2 KT1 (T/TNOM - 1)."""
    n = 8
    rng = splitmix64(40)
    w = 1e-6 * np.exp(0.2 * rng.standard_normal(n))
    l = 1e-7 * np.exp(0.2 * rng.standard_normal(n))
    card = {"TYPE": 1.0, "TNOM": 27.0, "UA": 2e-9, "UB": 1e-18, "UC": 5e-11, "UD": 0.01,
            "UA1": 3e-3, "UB1": -2e-3, "UC1": 1e-3, "UD1": 5e-3, "VSAT": 8e4, "AT": 2e-3,
            "RDSW": 150.0, "PRT": 0.4, "KT1": -0.11, "TEMPMOD": float(temp_mod)}
    T, Tn = 85.0 + 273.15, 27.0 + 273.15

    def build(c):
        spec = {"models": [(54, c)], "n": n, "model_index": np.zeros(n, np.int32),
                "inst_params": {"W": w, "L": l, "TEMP": np.full(n, 85.0)}}
        return _pair(spec)
    p = build(card)
    t0, dt = T / Tn - 1.0, T - Tn
    for name in ("UA", "UB", "UC", "UD"):
        base, k1 = card[name], card[name + "1"]
        want = base + k1 * t0 if temp_mod == 0 else base * (1.0 + k1 * dt)
        np.testing.assert_allclose(p.orc.read(name.lower()), want, rtol=1e-13, err_msg=name)
    want = 8e4 - 2e-3 * t0 if temp_mod == 0 else 8e4 * (1.0 - 2e-3 * dt)
    np.testing.assert_allclose(p.orc.read("vsattemp"), want, rtol=1e-14)
    # rds0 = RDSW(T) / (Weff,CJ in um)^WR with WR = 1 and NF = 1
    rdsw = 150.0 + 0.4 * t0 if temp_mod == 0 else 150.0 * (1.0 + 0.4 * dt)
    np.testing.assert_allclose(p.orc.read("rds0"), rdsw / (p.orc.read("weffCJ") * 1e6),
                               rtol=1e-13)
    if temp_mod == 1:
        q = build(dict(card, TEMPMOD=2.0))
        np.testing.assert_allclose(p.orc.read("vfbzbfactor") - q.orc.read("vfbzbfactor"),
                                   2 * -0.11 * t0, rtol=1e-9)
