"""Pins of the oracle functions that the Jacobian-difference and symmetry
checks cannot fix, each by what the mathematics or the model's defining
law fixes independently of the restated code (CPU only).

* SPICE support: Lambert W by w e^w = x; pnjlim by current continuity
  (the limited voltage carries the linearised current); fetlim and limvds
  by their fixed points, step bounds, monotonicity and no-overshoot.
* BSIM4 eval (N_DEV_MOSFET_B4.C:7744-11919): the threshold voltage at zero
  bias with the short-channel, narrow-width and DIBL terms off is VTH0; the
  narrow-width term scales as 1/(Weff + W0); the DIBL shift is linear in
  Vds and in ETA0; the KT1/KT1L temperature law; no channel current at
  Vds = 0; the deep-subthreshold slope d ln Ids/dVgs = 1/(n Vt) with
  n = 1 + CIT/Coxe; Vdseff is the smooth minimum that solves
  (Vdsat - Vdseff)(Vds - Vdseff) = DELTA Vdseff; the linear-region current
  has the form beta Vgsteff (1 - Abulk Vdseff/(2(Vgsteff + 2 Vt)))
  Vdseff/(1 + Vdseff/EsatL) with beta = ueff Coxeff Weff/Leff when CLM,
  DIBL, SCBE, DITS and Rds are off; the stamped F and Q of every instance
  sum to zero over its nodes (current and charge conservation).
* BSIM4 setup (:5640-7734): Leff/Weff from LINT/WINT and the LL/LW/LWL,
  WL/WW/WWL laws; vsat(T) = VSAT - AT (T/TNOM - 1); the junction saturation
  current's temperature law, measured through the body-row current.
* MOSFET1 and BJT junction capacitance temperature laws: PB(T) in its
  closed form, CJ(T) by the grading law, both the identity at T = TNOM.
* BJT old-style excess phase (:2370-2466): vanishing delay gives the loads
  without excess phase; the filter's step response converges, at first
  order in h/td, to that of td^2/3 y'' + td y' + y = x.

Each check is shown to fail on a planted error where the bar could look
loose.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
from harness import Pair, splitmix64

KBOQ = 8.617087e-5            # N_DEV_MOSFET_B4.C CONSTKboQ
EPS0 = 8.854214871e-12        # N_DEV_Const.h CONSTperm0
Q, BOLTZ, REF = 1.6021918e-19, 1.3806226e-23, 300.15
KOVERQ = BOLTZ / Q
EG300 = 1.1150877


def _eg(t):
    return 1.16 - 7.02e-4 * t * t / (t + 1108.0)


# ---------------------------------------------------------------------------
# SPICE support functions
# ---------------------------------------------------------------------------
def _lib():
    L = oracle.lib()
    L.or_lambertw.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_int),
                              C.POINTER(C.c_double)]
    L.or_lambertw.restype = None
    L.or_pnjlim.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double,
                            C.POINTER(C.c_int)]
    L.or_pnjlim.restype = C.c_double
    L.or_fetlim.argtypes = [C.c_double, C.c_double, C.c_double]
    L.or_fetlim.restype = C.c_double
    L.or_limvds.argtypes = [C.c_double, C.c_double]
    L.or_limvds.restype = C.c_double
    return L


def _lambertw(x):
    w, ierr, xi = C.c_double(), C.c_int(), C.c_double()
    _lib().or_lambertw(x, C.byref(w), C.byref(ierr), C.byref(xi))
    return w.value, ierr.value


def test_lambertw_inverts_w_exp_w():
    turn = -math.exp(-1.0)
    xs = np.concatenate([turn + np.logspace(-12, -1, 40), np.linspace(-0.35, -1e-6, 60),
                         np.logspace(-300, 0, 60), np.linspace(1.0, 50.0, 60),
                         np.logspace(1, 300, 60)])
    for x in xs:
        w, ierr = _lambertw(x)
        assert ierr == 0, x
        assert w >= -1.0
        # w e^w = x to the precision that w carries
        assert abs(w * math.exp(w) - x) <= 1e-14 * abs(x) + 1e-14 * abs(w * math.exp(w)) \
            * max(1.0, abs(w)), (x, w)
    assert _lambertw(0.0) == (0.0, 0)
    assert abs(_lambertw(math.e)[0] - 1.0) <= 2e-16
    assert abs(_lambertw(2.0 * math.exp(2.0))[0] - 2.0) <= 4e-16
    assert abs(_lambertw(-math.exp(-1.0) + 1e-300)[0] + 1.0) <= 1e-7
    w, ierr = _lambertw(turn - 1e-3)       # below the branch point: no real solution
    assert ierr == 1 and w == -1.0


def test_pnjlim_keeps_the_linearised_current():
    """A limited step lands where the exponential carries the current the
    junction's tangent at vold predicts: exp((v - vold)/vt) = 1 + (vnew -
    vold)/vt for vold > 0, exp(v/vt) = vnew/vt for vold <= 0."""
    L = _lib()
    rng = splitmix64(51)
    vt, vcrit = 0.025852, 0.62
    for _ in range(2000):
        vold = rng.uniform(-1.0, 0.9)
        vnew = rng.uniform(vcrit, 3.0)
        lim = C.c_int()
        v = L.or_pnjlim(vnew, vold, vt, vcrit, C.byref(lim))
        if abs(vnew - vold) <= 2 * vt:
            assert v == vnew and lim.value == 0
            continue
        assert lim.value == 1
        if vold > 0:
            arg = 1 + (vnew - vold) / vt
            if arg <= 0:                # a step down past the tangent's zero: vcrit
                assert v == vcrit
                continue
            assert math.exp((v - vold) / vt) == pytest.approx(arg, rel=1e-12)
        else:
            assert math.exp(v / vt) == pytest.approx(vnew / vt, rel=1e-12)
        # the log-damped step never overshoots
        assert (v - vold) * (vnew - vold) >= 0 and abs(v - vold) <= abs(vnew - vold)


def test_fetlim_bounds():
    L = _lib()
    rng = splitmix64(52)
    for _ in range(20000):
        vto = rng.uniform(-1.0, 1.0)
        vold = vto + rng.uniform(-6.0, 8.0)
        vnew = vold + rng.uniform(-10.0, 10.0)
        v = L.or_fetlim(vnew, vold, vto)
        assert L.or_fetlim(vold, vold, vto) == vold                    # fixed point
        assert min(vold, vnew) <= v <= max(vold, vnew)                 # no overshoot
        hi = abs(2 * (vold - vto)) + 2
        if vold < vto or vold >= vto + 3.5:                            # off / fully on
            assert abs(v - vold) <= max(hi, hi / 2 + 2) + 1e-12
        else:                                                          # middle region
            assert vto - 0.5 <= v <= vto + 4.0 or v == vnew
        # monotone in vnew
        assert L.or_fetlim(vnew + 0.01, vold, vto) >= v - 1e-15


def test_limvds_bounds():
    L = _lib()
    rng = splitmix64(53)
    for _ in range(20000):
        vold = rng.uniform(-2.0, 8.0)
        vnew = rng.uniform(-4.0, 30.0)
        v = L.or_limvds(vnew, vold)
        if vold >= 3.5:
            assert 2.0 <= v or v == vnew
            assert v <= max(3.0 * vold + 2.0, vnew) and (v <= 3.0 * vold + 2.0 or v == vnew)
            if 3.5 <= vnew <= vold:
                assert v == vnew
        else:
            assert min(-0.5, vnew) <= v <= 4.0
            if vnew < vold:
                assert v >= -0.5                                       # a fall stops at -0.5
            if -0.5 <= vnew <= 4.0:
                assert v == vnew                                       # identity band
        # monotone in vnew on either side of vold (below -0.5 a fall is
        # stopped at -0.5 while a rise is not, so it jumps at vnew = vold)
        if (vnew + 0.01 - vold) * (vnew - vold) > 0:
            assert L.or_limvds(vnew + 0.01, vold) >= v


# ---------------------------------------------------------------------------
# BSIM4 evaluation
# ---------------------------------------------------------------------------
QUIET = {"K3": 0.0, "DVT0": 0.0, "DVT0W": 0.0, "LPE0": 0.0, "DVTP0": 0.0, "TNOM": 27.0}


def _b4(card, n, seed, w=None, l=None, extra=None, temp=None):
    rng = splitmix64(seed)
    w = 90e-9 * np.exp(0.2 * rng.standard_normal(n)) if w is None else np.broadcast_to(w, (n,))
    l = 45e-9 * np.exp(0.2 * rng.standard_normal(n)) if l is None else np.broadcast_to(l, (n,))
    params = {"W": np.asarray(w, float), "L": np.asarray(l, float)}
    if extra:
        params.update(extra)
    if temp is not None:
        params["TEMP"] = np.full(n, float(temp))
    spec = {"models": [(54, card)], "n": n, "model_index": np.zeros(n, np.int32),
            "inst_params": params}
    return Pair("bsim4", spec, gpu=False, options=oracle.default_options(voltageLimiterFlag=0))


def _terminals(p, vd, vg, vs, vb):
    x = np.zeros(p.topo.n_nodes)
    offs = np.concatenate([[0], np.cumsum(p.counts[0] + p.counts[1])])[:-1]
    lids = p.topo.node_lids
    for k, v in enumerate((vd, vg, vs, vb)):
        x[lids[offs + k]] = v
    p.set_inputs(nextSol=x)


DC = dict(newtonIter=1, dcopFlag=1)


def _vth(p, vd=0.0, vg=0.6, vb=0.0):
    _terminals(p, vd, vg, 0.0, vb)
    p.load(**DC)
    return p.orc.read("von")


@pytest.mark.parametrize("dtype", [1.0, -1.0])
def test_vth_is_vth0_at_zero_bias(dtype):
    n = 48
    dv = splitmix64(61).uniform(-0.05, 0.05, n)
    p = _b4(dict(QUIET, TYPE=dtype, VTH0=0.45 * dtype), n, 62, extra={"DELVTO": dv})
    got = _vth(p, vg=0.6 * dtype)
    np.testing.assert_allclose(got, dtype * (0.45 * dtype + dv), rtol=0, atol=1e-14)


def test_vth_narrow_width_term():
    """Vth - VTH0 = K3 TOXE phi / (Weff + W0) at zero bias: the product with
    Weff + W0 is one constant per card, and K3 scales it."""
    n = 48
    shifts = []
    for k3 in (40.0, 80.0):
        p = _b4(dict(QUIET, TYPE=1.0, VTH0=0.45, K3=k3, W0=2.5e-7), n, 63)
        d = (_vth(p) - 0.45) * (p.orc.read("weff") + 2.5e-7)
        assert np.all(d > 0)
        np.testing.assert_allclose(d, d[0], rtol=1e-11)
        shifts.append(d[0])
    assert shifts[1] == pytest.approx(2.0 * shifts[0], rel=1e-12)
    # the bar sees a width law off by a 1e-9 relative term
    bad = d * (1.0 + 1e-9 * np.arange(n))
    assert not np.allclose(bad, bad[0], rtol=1e-11, atol=0)


def test_vth_dibl_shift_linear_in_vds_and_eta0():
    n = 32
    shift = {}
    for eta in (0.08, 0.16):
        p = _b4(dict(QUIET, TYPE=1.0, VTH0=0.45, ETA0=eta, ETAB=0.0), n, 64)
        v0 = _vth(p, vd=0.0)
        s1 = _vth(p, vd=0.1) - v0
        s3 = _vth(p, vd=0.3) - v0
        assert np.all(s1 < 0)
        np.testing.assert_allclose(s3 / s1, 3.0, rtol=1e-10)
        shift[eta] = s1
    np.testing.assert_allclose(shift[0.16] / shift[0.08], 2.0, rtol=1e-10)


def test_vth_temperature_law():
    """tempMod 0: Vth(T) = VTH0 + (KT1 + KT1L/Leff)(T/TNOM - 1) at zero bias."""
    n = 32
    kt1, kt1l = -0.11, 2e-9
    p = _b4(dict(QUIET, TYPE=1.0, VTH0=0.45, KT1=kt1, KT1L=kt1l, KT2=0.0), n, 65, temp=85.0)
    want = 0.45 + (kt1 + kt1l / p.orc.read("leff")) * ((85.0 + 273.15) / (27.0 + 273.15) - 1.0)
    np.testing.assert_allclose(_vth(p), want, rtol=0, atol=2e-15)


def test_no_channel_current_at_zero_vds():
    n = 64
    rng = splitmix64(66)
    p = _b4({"TYPE": 1.0}, n, 67)
    _terminals(p, 0.0, rng.uniform(0.0, 1.2, n), 0.0, rng.uniform(-0.3, 0.0, n))
    p.load(**DC)
    assert np.all(p.orc.read("cdrain") == 0.0)


@pytest.mark.parametrize("cit", [0.0, 2e-3])
def test_deep_subthreshold_slope(cit):
    """Deep below threshold Vgsteff is an exponential of Vgst/(n Vt), so
    gm / Ids = 1 / (n Vt) with n = 1 + CIT/Coxe when NFACTOR and CDSC* are
    zero; Coxe = EPSROX eps0 / TOXE.  At Vgst = -1 V the departures are below
    1e-7: the exponential's own corrections (x/2 in log(1 + x), x ~ 1e-8,
    and the denominator's m*/(n Coxe/Cdep0 e^(...)) ~ 1e-8), and the rounding
    of the reference's log(1.0 + x) (~1e-16/x), which grows deeper down.
    Below m* Vgst = -34 n Vt the model holds Vgsteff at a floor: gm = 0."""
    n = 24
    card = dict(QUIET, TYPE=1.0, NFACTOR=0.0, CDSC=0.0, CDSCB=0.0, CDSCD=0.0, CIT=cit,
                TOXE=3e-9, EPSROX=3.9)
    p = _b4(card, n, 68)
    vth = _vth(p, vd=0.5)
    _terminals(p, 0.5, vth - 1.0, 0.0, 0.0)
    p.load(**DC)
    ids, gm = p.orc.read("cdrain"), p.orc.read("gm")
    nfac = 1.0 + cit / (3.9 * EPS0 / 3e-9)
    vt = KBOQ * 300.15
    np.testing.assert_allclose(gm / ids * nfac * vt, 1.0, rtol=2e-7)
    assert np.all(np.abs(gm / ids * (1.0 + 1e-5) * nfac * vt - 1.0) > 2e-7)   # n off by 1e-5
    _terminals(p, 0.5, vth - 2.4, 0.0, 0.0)
    p.load(**DC)
    assert np.all(p.orc.read("gm") == 0.0) and np.all(p.orc.read("cdrain") > 0.0)


def test_vdseff_smooth_minimum():
    n = 400
    rng = splitmix64(69)
    p = _b4({"TYPE": 1.0, "DELTA": 0.01}, n, 70)
    vds = rng.uniform(0.01, 1.5, n)
    _terminals(p, vds, rng.uniform(0.3, 1.5, n), 0.0, rng.uniform(-0.5, 0.0, n))
    p.load(**DC)
    vdsat, vdseff, mode = p.orc.read("vdsat"), p.orc.read("Vdseff"), p.orc.read("mode")
    keep = (mode == 1) & (vdseff < vds)
    assert keep.sum() > 300
    lhs = (vdsat - vdseff) * (vds - vdseff)
    scale = 0.01 * vdsat + np.abs(vdsat * vds)
    np.testing.assert_array_less(np.abs(lhs - 0.01 * vdseff)[keep], 1e-13 * scale[keep])
    assert np.all(vdseff[keep] <= np.minimum(vds, vdsat)[keep])
    # a DELTA off by 1e-6 relative is outside the bar
    assert np.any(np.abs(lhs - 0.01 * (1 + 1e-6) * vdseff)[keep] > 1e-13 * scale[keep])


def test_linear_region_current_form():
    """Long channel, CLM/DIBL/SCBE/DITS and Rds off: cdrain / [Vgsteff (1 -
    Abulk Vdseff / (2 (Vgsteff + 2 Vt))) Vdseff / (1 + Vdseff/EsatL)] is
    beta = ueff Coxeff Weff/Leff, independent of Vds, and beta Leff/(ueff
    Weff) = Coxeff is the same for every W and L at the same bias."""
    card = {"TYPE": 1.0, "PCLM": 0.0, "PDIBLC1": 0.0, "PDIBLC2": 0.0, "PDIBLCB": 0.0,
            "PVAG": 0.0, "PSCBE2": 0.0, "PDITS": 0.0, "RDSW": 0.0, "RDSWMIN": 0.0,
            "ETA0": 0.0, "DVT0": 0.0, "DVT0W": 0.0, "K3": 0.0, "LPE0": 0.0}
    n = 16
    rng = splitmix64(71)
    w = rng.uniform(5e-6, 20e-6, n)
    l = rng.uniform(5e-6, 20e-6, n)
    p = _b4(card, n, 72, w=w, l=l)
    vt = KBOQ * 300.15
    betas = []
    for vds in (0.005, 0.01, 0.02, 0.04):
        _terminals(p, vds, 1.0, 0.0, 0.0)
        p.load(**DC)
        g = p.orc.read("Vgsteff")
        a = p.orc.read("Abulk")
        vde = p.orc.read("Vdseff")
        form = g * (1.0 - a * vde / (2.0 * (g + 2.0 * vt))) * vde / (1.0 + vde / p.orc.read("EsatL"))
        betas.append(p.orc.read("cdrain") / form)
    betas = np.asarray(betas)
    np.testing.assert_allclose(betas / betas[0], 1.0, rtol=1e-10)
    coxeff = betas[0] * p.orc.read("leff") / (p.orc.read("ueff") * p.orc.read("weff"))
    np.testing.assert_allclose(coxeff, coxeff[0], rtol=1e-9)
    assert 0.5 * 3.9 * EPS0 / 3e-9 < coxeff[0] < 3.9 * EPS0 / 3e-9 * 1.0000001


def test_mobility_degradation_law():
    """mobMod 0: ueff = u0(T) / (1 + T3 (UA + UB T3) + UD (TOXE/(Vgsteff +
    2 sqrt(Vth^2 + 1e-4)))^2 Vth^2), T3 = (Vgsteff + 2 Vth)/TOXE, at Vbs = 0."""
    n = 32
    rng = splitmix64(81)
    ua, ub, ud, toxe = 2e-9, 1.5e-18, 1e14 * 1e-19, 2.5e-9
    p = _b4({"TYPE": 1.0, "MOBMOD": 0.0, "UA": ua, "UB": ub, "UC": 0.0, "UD": ud,
             "TOXE": toxe}, n, 82)
    _terminals(p, rng.uniform(0.05, 1.2, n), rng.uniform(0.2, 1.5, n), 0.0, 0.0)
    p.load(**DC)
    g, vth = p.orc.read("Vgsteff"), p.orc.read("von")
    t3 = (g + 2 * vth) / toxe
    coul = ud * (toxe / (g + 2 * np.sqrt(vth * vth + 1e-4))) ** 2 * vth * vth
    want = p.orc.read("u0temp") / (1.0 + t3 * (ua + ub * t3) + coul)
    np.testing.assert_allclose(p.orc.read("ueff"), want, rtol=1e-13)


def test_gidl_current_law():
    """Igidl = AGIDL WeffCJ T1 exp(-BGIDL/T1) (-Vbd^3)/(CGIDL - Vbd^3),
    T1 = (Vds - Vgs - EGIDL) / (3 TOXE), for Vbd <= 0 (no poly depletion)."""
    n = 40
    rng = splitmix64(83)
    agidl, bgidl, cgidl, egidl, toxe = 1e-9, 1e9, 0.5, 0.8, 3e-9
    p = _b4({"TYPE": 1.0, "AGIDL": agidl, "BGIDL": bgidl, "CGIDL": cgidl, "EGIDL": egidl,
             "TOXE": toxe}, n, 84)
    vd, vg = rng.uniform(1.2, 2.5, n), rng.uniform(-0.6, 0.0, n)
    _terminals(p, vd, vg, 0.0, 0.0)
    p.load(**DC)
    np.testing.assert_allclose(p.orc.read("weffCJ"), p.orc.read("weff"), rtol=1e-15)
    t1 = (vd - vg - egidl) / (3 * toxe)
    want = agidl * p.orc.read("weffCJ") * t1 * np.exp(-bgidl / t1) * vd ** 3 / (cgidl + vd ** 3)
    np.testing.assert_allclose(p.orc.read("Igidl"), want, rtol=1e-13)


def test_charge_partition_symmetric_at_zero_vds():
    """capMod 2 with the 40/60 partition (XPART = 0): at Vds = 0 the channel
    charge splits equally between drain and source, in inversion and below."""
    n = 32
    rng = splitmix64(85)
    p = _b4({"TYPE": 1.0, "XPART": 0.0}, n, 86)
    vs = rng.uniform(-0.2, 0.2, n)
    _terminals(p, vs, rng.uniform(-0.3, 1.5, n), vs, rng.uniform(-0.5, 0.0, n))
    p.load(newtonIter=1, transientFlag=1, pdt=1e12)
    qs, qd = p.orc.read("qsrc"), p.orc.read("qdrn")
    assert np.any(np.abs(qs) > 1e-18)
    np.testing.assert_allclose(qd, qs, rtol=1e-11, atol=1e-30)


def test_stamped_currents_and_charges_are_conserved():
    """Every current and charge of an instance flows between its own nodes:
    its F and Q contributions sum to zero over its nodes (private topology)."""
    from specs import BSIM4_CARDS, bsim4_spec
    spec = bsim4_spec(200, seed=73, cards=BSIM4_CARDS, geometry=True, nf_mix=True)
    p = Pair("bsim4", spec, gpu=False, options=oracle.default_options(voltageLimiterFlag=0))
    per = (p.counts[0] + p.counts[1]).astype(np.int64)
    p.set_inputs(nextSol=splitmix64(74).uniform(-0.5, 1.5, p.topo.n_nodes),
                 currSto=splitmix64(75).uniform(-0.5, 1.5, p.topo.n_sto))
    ro, _ = p.load(newtonIter=1, transientFlag=1, pdt=1e12)
    starts = np.concatenate([[0], np.cumsum(per)])
    for k in ("f", "q"):
        tot = np.add.reduceat(ro[k], starts[:-1])
        mag = np.add.reduceat(p.mag[k], starts[:-1])
        np.testing.assert_array_less(np.abs(tot), 1e-13 * mag + 1e-300, err_msg=k)


# ---------------------------------------------------------------------------
# BSIM4 setup
# ---------------------------------------------------------------------------
def test_effective_geometry_laws():
    """Leff = L + XL - 2 (LINT + LL/Lnew^LLN + LW/Wnew^LWN + LWL/(Lnew^LLN Wnew^LWN)),
    Weff = W/NF + XW - 2 (WINT + WL/Lnew^WLN + WW/Wnew^WWN + WWL/(Lnew^WLN Wnew^WWN))."""
    g = {"LINT": 5e-9, "WINT": 8e-9, "LL": 2e-18, "LLN": 1.3, "LW": 1e-17, "LWN": 1.1,
         "LWL": 3e-26, "WL": 4e-18, "WLN": 1.2, "WW": 1e-17, "WWN": 1.05, "WWL": 2e-26,
         "XL": 2e-9, "XW": 3e-9}
    n = 40
    rng = splitmix64(76)
    w = 200e-9 * np.exp(0.3 * rng.standard_normal(n))
    l = 60e-9 * np.exp(0.3 * rng.standard_normal(n))
    nf = rng.integers(1, 4, n).astype(float)
    p = _b4(dict(g, TYPE=1.0), n, 77, w=w, l=l, extra={"NF": nf})
    ln, wn = l + g["XL"], w / nf + g["XW"]
    dl = g["LINT"] + g["LL"] / ln ** g["LLN"] + g["LW"] / wn ** g["LWN"] + \
        g["LWL"] / (ln ** g["LLN"] * wn ** g["LWN"])
    dw = g["WINT"] + g["WL"] / ln ** g["WLN"] + g["WW"] / wn ** g["WWN"] + \
        g["WWL"] / (ln ** g["WLN"] * wn ** g["WWN"])
    np.testing.assert_allclose(p.orc.read("leff"), ln - 2 * dl, rtol=1e-14)
    np.testing.assert_allclose(p.orc.read("weff"), wn - 2 * dw, rtol=1e-14)


def test_saturation_velocity_temperature_law():
    n = 16
    p = _b4({"TYPE": 1.0, "VSAT": 8e4, "AT": 3.3e4, "TNOM": 27.0}, n, 78, temp=110.0)
    want = 8e4 - 3.3e4 * ((110.0 + 273.15) / (27.0 + 273.15) - 1.0)
    np.testing.assert_allclose(p.orc.read("vsattemp"), want, rtol=1e-15)


@pytest.mark.parametrize("temp", [-20.0, 85.0])
def test_junction_current_temperature_law(temp):
    """Body row of a device with D = G = S = 0: two junctions carrying
    A JS(T) (exp(V/(NJ Vt)) - 1) + gmin V, JS(T) = JSS exp((Eg(TNOM)/Vt(TNOM)
    - Eg(T)/Vt(T) + XTIS ln(T/TNOM)) / NJ)."""
    n = 48
    rng = splitmix64(79)
    w = 90e-9 * np.exp(0.2 * rng.standard_normal(n))
    jss, xtis, nj = 2e-4, 3.0, 1.2
    card = {"TYPE": 1.0, "JSS": jss, "JSWS": 0.0, "JSWGS": 0.0, "XTIS": xtis, "NJS": nj,
            "TNOM": 27.0}
    p = _b4(card, n, 80, w=w, l=45e-9,
            extra={"AS": w * 1e-7, "AD": w * 1.2e-7, "PS": 2 * (w + 1e-7),
                   "PD": 2 * (w + 1.2e-7)}, temp=temp)
    T, Tn = temp + 273.15, 300.15
    js = jss * np.exp((_eg(Tn) / (KBOQ * Tn) - _eg(T) / (KBOQ * T) + xtis * np.log(T / Tn)) / nj)
    v = rng.uniform(-0.3, 0.5, n)
    _terminals(p, 0.0, 0.0, 0.0, v)
    ro, _ = p.load(newtonIter=1, transientFlag=1)
    diode = lambda a: a * js * np.expm1(v / (nj * KBOQ * T)) + 1e-12 * v
    offs = np.concatenate([[0], np.cumsum(p.counts[0] + p.counts[1])])[:-1]
    got = ro["f"][p.topo.node_lids[offs + 3]]
    np.testing.assert_allclose(got, diode(w * 1e-7) + diode(w * 1.2e-7), rtol=1e-9, atol=1e-26)


# ---------------------------------------------------------------------------
# MOSFET1 and BJT junction temperature laws
# ---------------------------------------------------------------------------
def _pb_closed(pb, t, tnom):
    """SPICE junction potential: (T/TNOM) PB - 3 Vt ln(T/TNOM) - (T/TNOM) Eg(TNOM) + Eg(T)."""
    vt = KOVERQ * t
    return (t / tnom) * pb - 3.0 * vt * np.log(t / tnom) - (t / tnom) * _eg(tnom) + _eg(t)


def _pbfact(t):
    """pbfact(T) = Eg(T) - Eg300 T/REF - 3 Vt ln(T/REF) in closed form."""
    return _eg(t) - EG300 * t / REF - 3.0 * KOVERQ * t * np.log(t / REF)


@pytest.mark.parametrize("tnom,temp", [(50.0, 100.0), (27.0, -40.0), (60.0, 60.0)])
def test_mosfet1_junction_temperature_laws(tnom, temp):
    pb, cj, mj, cjsw, mjsw = 0.8, 2e-4, 0.45, 3e-10, 0.3
    card = {"TYPE": 1.0, "VTO": 0.5, "KP": 1e-4, "PB": pb, "CJ": cj, "MJ": mj, "CJSW": cjsw,
            "MJSW": mjsw, "TNOM": tnom}
    n = 8
    spec = {"models": [(1, card)], "n": n, "model_index": np.zeros(n, np.int32),
            "inst_params": {"W": np.full(n, 1e-6), "L": np.full(n, 1e-6),
                            "TEMP": np.full(n, temp)}}
    p = Pair("mosfet1", spec, gpu=False)
    T, Tn = temp + 273.15, tnom + 273.15
    tpb = _pb_closed(pb, T, Tn)
    np.testing.assert_allclose(p.orc.read("tBulkPot"), tpb, rtol=1e-12)
    # the potential extrapolated to the reference temperature, with the band
    # gap there taken as the constant Eg300 (1.1150877 V; Eg(300.15 K) from
    # the band-gap law is 1.1150910 V)
    pb0 = _pb_closed(pb, REF, Tn) - _eg(REF) + EG300
    for m, c0, name in ((mj, cj, "tCj"), (mjsw, cjsw, "tCjsw")):
        want = c0 * (1 + m * (4e-4 * (T - REF) - (tpb - pb0) / pb0)) / \
            (1 + m * (4e-4 * (Tn - REF) - (pb - pb0) / pb0))
        np.testing.assert_allclose(p.orc.read(name), want, rtol=1e-11)
    if tnom == temp:
        np.testing.assert_allclose(p.orc.read("tBulkPot"), pb, rtol=1e-14)
        np.testing.assert_allclose(p.orc.read("tCj"), cj, rtol=1e-14)


@pytest.mark.parametrize("tnom,temp", [(50.0, 100.0), (27.0, -40.0), (60.0, 60.0)])
def test_bjt_junction_temperature_laws(tnom, temp):
    """The reference's BJT law takes pbfact at T where SPICE takes it at TNOM
    (N_DEV_BJT.C:1658): PB(T) = (T/TNOM) PB + pbfact(T) (1 - T/TNOM), which
    equals PB at T = TNOM."""
    vje, cje, mje, vjc, cjc, mjc = 0.75, 1e-12, 0.33, 0.7, 5e-13, 0.4
    card = {"TYPE": 1.0, "IS": 1e-16, "BF": 100.0, "VJE": vje, "CJE": cje, "MJE": mje,
            "VJC": vjc, "CJC": cjc, "MJC": mjc, "TNOM": tnom}
    n = 8
    spec = {"models": [(1, card)], "n": n, "model_index": np.zeros(n, np.int32),
            "inst_params": {"AREA": np.ones(n), "TEMP": np.full(n, temp)}}
    p = Pair("bjt", spec, gpu=False)
    T, Tn = temp + 273.15, tnom + 273.15
    for pb, c0, m, pot, cap in ((vje, cje, mje, "tBEPot", "tBECap"),
                                (vjc, cjc, mjc, "tBCPot", "tBCCap")):
        tpb = (T / Tn) * pb + _pbfact(T) * (1.0 - T / Tn)
        np.testing.assert_allclose(p.orc.read(pot), tpb, rtol=1e-12)
        pbo = (pb - _pbfact(T)) / (Tn / REF)
        want = c0 * (1 + m * (4e-4 * (T - REF) - (tpb - pbo) / pbo)) / \
            (1 + m * (4e-4 * (Tn - REF) - (pb - pbo) / pbo))
        np.testing.assert_allclose(p.orc.read(cap), want, rtol=1e-11)
        if tnom == temp:
            np.testing.assert_allclose(p.orc.read(pot), pb, rtol=1e-14)
            np.testing.assert_allclose(p.orc.read(cap), c0, rtol=1e-14)


# ---------------------------------------------------------------------------
# BJT old-style excess phase
# ---------------------------------------------------------------------------
BJT_XP = {"TYPE": 1.0, "IS": 1e-16, "BF": 100.0, "VAF": 50.0, "TF": 1e-10, "CJE": 1e-12,
          "CJC": 5e-13}


def _bjt_pair(ptf, n=4):
    spec = {"models": [(1, dict(BJT_XP, PTF=ptf))], "n": n, "model_index": np.zeros(n, np.int32),
            "inst_params": {"AREA": np.ones(n)}}
    return Pair("bjt", spec, gpu=False, options=oracle.default_options(voltageLimiterFlag=0))


def _bjt_bias(p, vbe, vce):
    """Private nodes C, B, E, S (internal nodes collapsed without RB/RC/RE)."""
    x = np.zeros(p.topo.n_nodes)
    offs = np.concatenate([[0], np.cumsum(p.counts[0] + p.counts[1])])[:-1]
    lids = p.topo.node_lids
    x[lids[offs + 0]] = vce
    x[lids[offs + 1]] = vbe
    p.set_inputs(nextSol=x)


def test_excess_phase_vanishing_delay():
    """td -> 0: the filter passes iBE/qB through, so the loads equal those
    without excess phase."""
    h = 1e-12
    step = dict(transientFlag=1, pdt=1.0 / h, currTimeStep=h, lastTimeStep=h)
    a, b = _bjt_pair(0.0), _bjt_pair(1e-9)
    for p in (a, b):
        _bjt_bias(p, np.array([0.6, 0.65, 0.7, 0.72]), 1.0)
    for flags in (dict(newtonIter=0, beginIntegrationFlag=1, **step), dict(newtonIter=1, **step)):
        ra, _ = a.load(**flags)
        rb, _ = b.load(**flags)
        np.testing.assert_allclose(rb["f"], ra["f"], rtol=1e-7, atol=1e-20)


def _step_response(h, td_over, steps):
    """(t, e): the excess-phase state after a base-emitter step, normalised
    so that e = 1 before the step and 0 at the new steady state."""
    tf = BJT_XP["TF"]
    ptf = td_over * 180.0 / math.pi        # td = PTF / (180/pi) * TF = td_over * TF
    p = _bjt_pair(ptf, n=1)
    td = td_over * tf
    flags = dict(transientFlag=1, pdt=1.0 / h, currTimeStep=h, lastTimeStep=h)
    # steady state at the new bias (begin integration initialises the filter there)
    _bjt_bias(p, np.array([0.72]), 1.0)
    rB, _ = p.load(newtonIter=0, beginIntegrationFlag=1, **flags)
    xB = rB["nextSta"][6]
    _bjt_bias(p, np.array([0.66]), 1.0)
    rA, _ = p.load(newtonIter=0, beginIntegrationFlag=1, **flags)
    xA = rA["nextSta"][6]
    _bjt_bias(p, np.array([0.72]), 1.0)
    p.set_inputs(currSta=rA["currSta"], lastSta=rA["lastSta"])
    ys = [xA]
    for k in range(steps):
        r, _ = p.load(newtonIter=1, **flags)
        ys.append(r["nextSta"][6])
        p.set_inputs(lastSta=p.h["currSta"].copy(), currSta=r["nextSta"])
    t = np.arange(steps + 1) * h
    return t / td, (np.asarray(ys) - xB) / (xA - xB)


def test_excess_phase_step_response():
    """The state follows y'' + (3/td) y' + (3/td^2)(y - x) = 0, whose step
    response is e(t) = exp(-1.5 t/td)(cos(sqrt3 t/(2 td)) + sqrt3 sin(sqrt3
    t/(2 td))); the backward-difference recurrence converges to it at first
    order in h/td."""
    errs = []
    for h_over in (1 / 100, 1 / 200, 1 / 400):
        td_over = 3.0
        h = h_over * td_over * BJT_XP["TF"]
        u, e = _step_response(h, td_over, int(round(6.0 / h_over)))
        r3 = math.sqrt(3.0)
        exact = np.exp(-1.5 * u) * (np.cos(r3 * u / 2) + r3 * np.sin(r3 * u / 2))
        errs.append(np.max(np.abs(e - exact)))
        assert e[0] == 1.0 and abs(e[-1]) < 0.01
    assert errs[-1] < 0.01
    assert 1.7 < errs[0] / errs[1] < 2.3 and 1.7 < errs[1] / errs[2] < 2.3
