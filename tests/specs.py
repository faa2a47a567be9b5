"""Seeded synthetic circuit specs for the four models (parity-test inputs)."""
import numpy as np

from harness import splitmix64


def lognormal(rng, median, sigma, n):
    return median * np.exp(sigma * rng.standard_normal(n))


# Diode cards spanning the evaluation's branches: default (no charge),
# depletion + diffusion charge, series resistance (internal node), breakdown,
# level 2 (recombination, high injection, temperature coefficients).
DIODE_CARDS = [
    (1, {}),
    (1, {"IS": 2e-14, "N": 1.05, "CJO": 2e-12, "VJ": 0.75, "M": 0.4, "TT": 5e-9, "FC": 0.5}),
    (1, {"IS": 1e-13, "RS": 10.0, "CJO": 1e-12, "TT": 1e-9, "BV": 5.0, "IBV": 1e-5}),
    (1, {"IS": 5e-15, "BV": 1.2, "IBV": 1e-3, "IRF": 1.5, "CJO": 3e-13, "M": 0.33}),
    (2, {"IS": 1e-14, "ISR": 1e-12, "NR": 2.0, "IKF": 1e-2, "CJO": 1e-12, "TT": 2e-9,
         "RS": 2.0, "BV": 3.0, "IBV": 1e-4, "TBV1": 1e-3, "TRS1": 2e-3, "TNOM": 25.0}),
]


def diode_spec(n, seed=20261203, cards=None, temp_spread=False):
    rng = splitmix64(seed)
    cards = cards if cards is not None else DIODE_CARDS
    mi = rng.integers(0, len(cards), size=n).astype(np.int32)
    params = {"AREA": lognormal(rng, 1.0, 0.2, n)}
    if temp_spread:
        params["TEMP"] = rng.uniform(-20.0, 120.0, n)
    return {"models": cards, "n": n, "model_index": mi, "inst_params": params}


def terminal_voltages(rng, n, lo=-0.5, hi=1.5):
    return rng.uniform(lo, hi, n)


# BSIM4 cards spanning the evaluation's branches.  Card 0/1 are the default
# NMOS/PMOS (the device-eval microbenchmark cards); the others switch on
# internal drain/source nodes (sheet resistance), junction models dioMod 0
# and 2 with breakdown and trap-assisted tunnelling, GIDL/GISL, gate
# polysilicon depletion, mobility models 1 and 2, charge partition 50/50 and
# 0/100, cvchargeMod 1, tempMod 1 and 2 with TNOM != TEMP (whose
# temperature coefficients are relative: AT, UA1.. given as such), velocity
# saturation and the velocity-overshoot term (separate card, below), binning terms, stress (LOD) and well proximity.
BSIM4_CARDS = [
    (54, {"TYPE": 1.0}),
    (54, {"TYPE": -1.0}),
    (54, {"TYPE": 1.0, "RSH": 20.0, "DIOMOD": 2.0, "JSS": 1e-4, "JSWS": 1e-10, "JSWGS": 1e-10,
          "CJS": 1e-3, "CJSWS": 1e-10, "CJSWGS": 1e-10, "BVS": 3.0, "XJBVS": 1.0,
          "AGIDL": 1e-9, "BGIDL": 1e9, "CGIDL": 0.5, "EGIDL": 0.8, "NGATE": 1e20,
          "CGSL": 1e-11, "CKAPPAS": 0.6, "TOXE": 1.8e-9, "TOXP": 1.5e-9, "TEMPMOD": 2.0,
          "AT": 1e-3, "UA1": 1e-3, "UB1": -1e-3, "UC1": 1e-3, "PRT": 1e-3}),
    (54, {"TYPE": -1.0, "DIOMOD": 0.0, "MOBMOD": 1.0, "XPART": 0.5, "CVCHARGEMOD": 1.0,
          "TEMPMOD": 1.0, "AT": 1e-3, "UA1": 1e-3, "UB1": -1e-3, "UC1": 1e-3, "PRT": 1e-3,
          "TNOM": 40.0, "JTSS": 1e-7, "NJTS": 15.0, "XTSS": 0.02, "VTSS": 10.0,
          "ACDE": 0.8, "MOIN": 12.0, "NOFF": 1.2, "VOFFCV": -0.05}),
    (54, {"TYPE": 1.0, "MOBMOD": 2.0, "XPART": 1.0,
          "LVTH0": 1e-9, "WVTH0": 2e-9, "PVTH0": 1e-17, "KU0": 1e-8, "KVTH0": 1e-8,
          "SAREF": 1e-6, "SBREF": 1e-6, "WEB": 1e-3, "WEC": 1e-4, "KVTH0WE": 1e-3,
          "PDITS": 0.1, "PDITSD": 0.1, "PDITSL": 1e6, "FPROUT": 0.2, "DVTP0": 1e-8, "DVTP1": 0.05}),
]
# Source-end velocity limit (VTL given).  The reference's Fsevl derivatives
# leave out the bias dependence of Coxeff, so its Jacobian is approximate
# there (N_DEV_MOSFET_B4.C:9905-9945): parity holds, a finite-difference
# check does not.
BSIM4_VTL_CARD = (54, {"TYPE": 1.0, "VTL": 2e5, "LAMBDA": 0.2, "MOBMOD": 2.0})

# Option cards: gate tunnelling (igcMod 1/2, igbMod, PIGCD given or not; a
# thin oxide so the currents matter), bias-dependent source/drain resistance
# (rdsMod 1), the capMod 0 and 1 charge models under each charge partition,
# the gate resistance networks (rgateMod 1-3), the body resistance networks
# (rbodyMod 1/2) and the NQS charge node (trnqsMod; acnqsMod).
BSIM4_OPT_CARDS = {
    "igc1": (54, {"TYPE": 1.0, "IGCMOD": 1.0, "IGBMOD": 1.0, "TOXE": 1.2e-9, "TOXP": 1.0e-9}),
    "igc2": (54, {"TYPE": -1.0, "IGCMOD": 2.0, "IGBMOD": 1.0, "PIGCD": 2.0, "TOXE": 1.5e-9,
                  "TOXP": 1.3e-9, "TEMPMOD": 2.0, "NGATE": 1e20, "VFBSDOFF": 0.02}),
    "igc2b": (54, {"TYPE": 1.0, "IGCMOD": 2.0, "TOXE": 1.1e-9, "TOXP": 0.9e-9, "AIGSD": 0.012,
                   "POXEDGE": 1.1}),
    "rds": (54, {"TYPE": 1.0, "RDSMOD": 1.0, "RSH": 20.0, "RDSW": 150.0, "RSW": 80.0, "RDW": 90.0,
                 "RSWMIN": 10.0, "RDWMIN": 12.0, "PRWG": 0.5, "PRWB": 0.1, "IGCMOD": 1.0,
                 "TOXE": 1.2e-9, "TOXP": 1.0e-9}),
    "rdsp": (54, {"TYPE": -1.0, "RDSMOD": 1.0, "RSH": 15.0, "RSW": 60.0, "RDW": 70.0, "PRWG": 0.3,
                  "PRWB": 0.05}),
    "cap0x0": (54, {"TYPE": 1.0, "CAPMOD": 0.0, "XPART": 0.0}),
    "cap0x5": (54, {"TYPE": -1.0, "CAPMOD": 0.0, "XPART": 0.5}),
    "cap0x1": (54, {"TYPE": 1.0, "CAPMOD": 0.0, "XPART": 1.0}),
    "cap1x0": (54, {"TYPE": 1.0, "CAPMOD": 1.0, "XPART": 0.0}),
    "cap1x1": (54, {"TYPE": -1.0, "CAPMOD": 1.0, "XPART": 1.0, "CVCHARGEMOD": 1.0}),
    "cap1x5": (54, {"TYPE": 1.0, "CAPMOD": 1.0, "XPART": 0.5}),
    "rg1": (54, {"TYPE": 1.0, "RGATEMOD": 1.0, "RSHG": 5.0}),
    "rg2": (54, {"TYPE": -1.0, "RGATEMOD": 2.0, "RSHG": 5.0, "XRCRG1": 12.0, "XRCRG2": 1.0}),
    "rg3": (54, {"TYPE": 1.0, "RGATEMOD": 3.0, "RSHG": 5.0, "CAPMOD": 1.0}),
    "rb1": (54, {"TYPE": 1.0, "RBODYMOD": 1.0, "RBDB": 40.0, "RBSB": 50.0, "RBPB": 30.0,
                 "RBPS": 20.0, "RBPD": 25.0}),
    "rb2": (54, {"TYPE": -1.0, "RBODYMOD": 2.0, "RBPS0": 40.0, "RBPD0": 45.0, "RBPBX0": 60.0,
                 "RBPBY0": 70.0, "RBSBX0": 80.0, "RBSBY0": 90.0, "RBDBX0": 85.0, "RBDBY0": 95.0}),
    "nqs": (54, {"TYPE": 1.0, "TRNQSMOD": 1.0, "XRCRG1": 12.0, "XRCRG2": 1.0}),
    "nqsrg3": (54, {"TYPE": -1.0, "TRNQSMOD": 1.0, "RGATEMOD": 3.0, "RBODYMOD": 1.0, "RSHG": 4.0}),
    "acnqs": (54, {"TYPE": 1.0, "ACNQSMOD": 1.0, "RGATEMOD": 2.0, "RSHG": 5.0}),
    "mtrl": (54, {"TYPE": 1.0, "MTRLMOD": 1.0, "EOT": 1.2e-9, "VDDEOT": 1.0, "PHIG": 4.2,
                  "EPSRGATE": 11.7, "NGATE": 1e20, "NSUB": 6e16, "AGIDL": 1e-9, "BGIDL": 1e9,
                  "CGIDL": 0.5, "EGIDL": 0.8, "AGISL": 1e-9, "BGISL": 1e9, "CGISL": 0.5}),
    "mtrlp": (54, {"TYPE": -1.0, "MTRLMOD": 1.0, "EOT": 1.5e-9, "PHIG": 4.9, "EPSRSUB": 12.9,
                   "BG0SUB": 1.1, "NI0SUB": 1.2e10, "MOBMOD": 0.0, "IGCMOD": 1.0}),
}


def bsim4_spec(n, seed=20261203, cards=None, geometry=True, temp_spread=False, nf_mix=False,
               sizes=None):
    """W/L lognormal (sigma 0.2) around 90 nm / 45 nm.  `geometry` adds the
    source/drain geometry, stress and multiplier instance parameters; `sizes`
    draws (W, L) from that many distinct pairs so size-dependent records are
    shared between instances."""
    rng = splitmix64(seed)
    cards = cards if cards is not None else BSIM4_CARDS
    mi = rng.integers(0, len(cards), size=n).astype(np.int32)
    if sizes:
        w_set = lognormal(rng, 90e-9, 0.2, sizes)
        l_set = lognormal(rng, 45e-9, 0.2, sizes)
        pick = rng.integers(0, sizes, size=n)
        params = {"W": w_set[pick], "L": l_set[pick]}
    else:
        params = {"W": lognormal(rng, 90e-9, 0.2, n), "L": lognormal(rng, 45e-9, 0.2, n)}
    if geometry:
        w = params["W"]
        params.update({
            "AS": w * 100e-9, "AD": w * 120e-9, "PS": 2 * (w + 100e-9), "PD": 2 * (w + 120e-9),
            "NRS": rng.uniform(0.5, 2.0, n), "NRD": rng.uniform(0.5, 2.0, n),
            "SA": lognormal(rng, 0.5e-6, 0.3, n), "SB": lognormal(rng, 0.5e-6, 0.3, n),
            "SCA": rng.uniform(0.0, 5.0, n), "M": rng.integers(1, 3, size=n).astype(float),
            "DELVTO": rng.uniform(-0.02, 0.02, n),
        })
    if nf_mix:
        params["NF"] = rng.integers(1, 3, size=n).astype(float)
    if temp_spread:
        params["TEMP"] = rng.uniform(-20.0, 120.0, n)
    return {"models": cards, "n": n, "model_index": mi, "inst_params": params}


# Level-1 MOSFET cards: the microbenchmark NMOS/PMOS pair, drain/source
# resistances (internal nodes) with a zero-bias junction model, NSUB-derived
# threshold and body factor with lateral diffusion, explicit zero-bias
# junction capacitances and unequal grading, and sheet resistance with
# saturation-current density.
MOSFET1_CARDS = [
    (1, {"TYPE": 1.0, "VTO": 0.45, "KP": 2e-4, "GAMMA": 0.4, "PHI": 0.7, "LAMBDA": 0.05,
         "TOX": 2e-9, "CGSO": 3e-10, "CGDO": 3e-10, "CJ": 1e-3, "MJ": 0.5, "CJSW": 3e-10,
         "MJSW": 0.33}),
    (1, {"TYPE": -1.0, "VTO": -0.45, "KP": 1e-4, "GAMMA": 0.4, "PHI": 0.7, "LAMBDA": 0.05,
         "TOX": 2e-9, "CGSO": 3e-10, "CGDO": 3e-10, "CJ": 1e-3, "MJ": 0.5, "CJSW": 3e-10,
         "MJSW": 0.33}),
    (1, {"TYPE": 1.0, "VTO": 0.5, "KP": 1e-4, "RD": 50.0, "RS": 30.0, "IS": 1e-15, "PB": 0.9,
         "TOX": 3e-9, "CGBO": 1e-10, "LAMBDA": 0.1, "FC": 0.4}),
    (1, {"TYPE": -1.0, "NSUB": 1e17, "TOX": 4e-9, "LD": 2e-9, "UO": 250.0, "CBD": 2e-15,
         "CBS": 3e-15, "MJ": 0.45, "MJSW": 0.3, "CJSW": 1e-10, "TNOM": 40.0}),
    (1, {"TYPE": 1.0, "VTO": 0.4, "KP": 3e-4, "GAMMA": 0.3, "RSH": 25.0, "JS": 1e-4,
         "CJ": 5e-4, "MJ": 0.4, "CJSW": 2e-10, "MJSW": 0.4, "TOX": 2.5e-9, "CGSO": 2e-10}),
]


def mosfet1_spec(n, seed=20261203, cards=None, geometry=True, temp_spread=False):
    rng = splitmix64(seed)
    cards = cards if cards is not None else MOSFET1_CARDS
    mi = rng.integers(0, len(cards), size=n).astype(np.int32)
    params = {"W": lognormal(rng, 90e-9, 0.2, n), "L": lognormal(rng, 45e-9, 0.2, n)}
    if geometry:
        w = params["W"]
        params.update({"AD": w * 120e-9, "AS": w * 100e-9, "PD": 2 * (w + 120e-9),
                       "PS": 2 * (w + 100e-9), "NRD": rng.uniform(0.5, 2.0, n),
                       "NRS": rng.uniform(0.5, 2.0, n), "M": rng.integers(1, 3, size=n).astype(float)})
    if temp_spread:
        params["TEMP"] = rng.uniform(-20.0, 120.0, n)
    return {"models": cards, "n": n, "model_index": mi, "inst_params": params}


# Gummel-Poon BJT cards: the microbenchmark NPN (Early effect, high-current
# roll-off, B-E leakage, series resistances, junction and transit-time
# charges); a PNP with reverse Early voltage, B-C leakage, base-resistance
# current crowding (IRB), substrate capacitance and forward-transit-time
# bias dependence; an NPN through the alias names with temperature exponents
# and the leakage set by C2/C4; and an NPN with excess phase.
BJT_CARDS = [
    (1, {"TYPE": 1.0, "IS": 1e-16, "BF": 100.0, "BR": 1.0, "VAF": 50.0, "IKF": 0.01,
         "ISE": 1e-14, "NE": 1.5, "RB": 10.0, "RC": 1.0, "RE": 0.5, "CJE": 1e-12,
         "CJC": 5e-13, "TF": 1e-10, "TR": 1e-8}),
    (1, {"TYPE": -1.0, "IS": 2e-16, "BF": 60.0, "BR": 3.0, "VAR": 10.0, "IKR": 5e-3,
         "ISC": 1e-14, "NC": 1.8, "RB": 50.0, "IRB": 1e-4, "RBM": 5.0, "CJS": 2e-13,
         "MJS": 0.4, "VJS": 0.6, "TF": 2e-10, "XTF": 2.0, "VTF": 3.0, "ITF": 0.05,
         "CJE": 8e-13, "MJE": 0.4, "CJC": 3e-13, "XCJC": 0.6, "FC": 0.6}),
    (1, {"TYPE": 1.0, "BFM": 150.0, "VA": 80.0, "IK": 0.02, "JLE": 1e-14, "NLE": 1.6,
         "XTB": 1.5, "XTI": 3.5, "EG": 1.15, "TNOM": 40.0, "C4": 10.0, "PE": 0.8, "ME": 0.4,
         "CJE": 1e-12, "NK": 0.6}),
    (1, {"TYPE": 1.0, "IS": 5e-17, "BF": 80.0, "VAF": 40.0, "TF": 3e-10, "PTF": 30.0,
         "CJE": 5e-13, "CJC": 2e-13, "RC": 2.0}),
]


def bjt_spec(n, seed=20261203, cards=None, temp_spread=False, lambertw_frac=0.0,
             off_frac=0.0, ic=False):
    rng = splitmix64(seed)
    cards = cards if cards is not None else BJT_CARDS
    mi = rng.integers(0, len(cards), size=n).astype(np.int32)
    params = {"AREA": lognormal(rng, 1.0, 0.2, n)}
    if temp_spread:
        params["TEMP"] = rng.uniform(-20.0, 120.0, n)
    if lambertw_frac:
        params["LAMBERTW"] = (rng.random(n) < lambertw_frac).astype(np.float64)
    if off_frac:
        params["OFF"] = (rng.random(n) < off_frac).astype(np.float64)
    if ic:
        params["IC1"] = rng.uniform(0.3, 0.8, n)
        params["IC2"] = rng.uniform(0.0, 3.0, n)
    return {"models": cards, "n": n, "model_index": mi, "inst_params": params}


RESISTOR_CARDS = [
    (1, {}),
    (1, {"TC1": 1.5e-3, "TC2": 2e-6, "TNOM": 30.0, "R": 1.25}),
    (1, {"TC1": -4e-4}),
]


def resistor_spec(n, seed=20261203, cards=None, temp_spread=False, sheet=False):
    """Resistors with R lognormal around 1 kOhm, or (sheet) R from RSH and
    per-instance L/W on cards that give RSH, NARROW and DEFW."""
    rng = splitmix64(seed)
    if sheet:
        cards = cards if cards is not None else [
            (1, {"RSH": 50.0, "NARROW": 1e-7}), (1, {"RSH": 120.0, "DEFW": 2e-6, "TC1": 1e-3})]
        params = {"L": lognormal(rng, 2e-6, 0.3, n)}
        if n:
            params["W"] = lognormal(rng, 1e-6, 0.2, n)
    else:
        cards = cards if cards is not None else RESISTOR_CARDS
        params = {"R": lognormal(rng, 1000.0, 0.5, n)}
    mi = rng.integers(0, len(cards), size=n).astype(np.int32)
    if temp_spread:
        params["TEMP"] = rng.uniform(-20.0, 120.0, n)
    return {"models": cards, "n": n, "model_index": mi, "inst_params": params}


CAPACITOR_CARDS = [
    (1, {}),
    (1, {"C": 1.2, "TC1": 2e-4, "TC2": 1e-6, "TNOM": 20.0}),
]


def capacitor_spec(n, seed=20261203, cards=None, temp_spread=False, ic=False, semi=False,
                   age=False):
    """Capacitors with C lognormal around 1 fF (or, semi, C from CJ/CJSW and
    per-instance L/W); ic adds IC= values (a branch unknown each); age adds
    AGE/D."""
    rng = splitmix64(seed)
    if semi:
        cards = cards if cards is not None else [
            (1, {"CJ": 1e-3, "CJSW": 1e-10, "NARROW": 5e-8}), (1, {"CJ": 2e-3, "DEFW": 3e-6})]
        params = {"L": lognormal(rng, 2e-6, 0.3, n), "W": lognormal(rng, 1e-6, 0.2, n)}
    else:
        cards = cards if cards is not None else CAPACITOR_CARDS
        params = {"C": lognormal(rng, 1e-15, 0.3, n)}
    mi = rng.integers(0, len(cards), size=n).astype(np.int32)
    if ic:
        params["IC"] = rng.uniform(-0.5, 1.5, n)
    if age:
        params["AGE"] = rng.uniform(0.5, 1e4, n)
        params["D"] = rng.uniform(0.01, 0.05, n)
    if temp_spread:
        params["TEMP"] = rng.uniform(-20.0, 120.0, n)
    return {"models": cards, "n": n, "model_index": mi, "inst_params": params}


def vsource_spec(n, seed=20261203, kinds=(-1, 0, 1, 2, 6), defaults_frac=0.2):
    """Voltage sources with a mix of waveforms (TRANSIENTSOURCETYPE: -1 none,
    0 SIN, 1 EXP, 2 PULSE, 6 DC) and random parameters; a fraction of the
    timing parameters is left not given (NaN) to take the solver-state
    defaults."""
    rng = splitmix64(seed)
    kind = rng.choice(np.asarray(kinds, np.float64), size=n)
    p = {"DCV0": rng.uniform(-1.0, 2.0, n), "TRANSIENTSOURCETYPE": kind,
         "V0": rng.uniform(-0.5, 0.5, n), "VA": rng.uniform(0.1, 1.0, n),
         "FREQ": rng.uniform(1e8, 5e9, n), "THETA": rng.uniform(0.0, 1e9, n),
         "PHASE": rng.uniform(-90.0, 90.0, n), "V1": rng.uniform(-0.5, 0.5, n),
         "V2": rng.uniform(0.5, 1.5, n), "TD": rng.uniform(0.0, 2e-10, n),
         "TR": rng.uniform(1e-11, 1e-10, n), "TF": rng.uniform(1e-11, 1e-10, n),
         "PW": rng.uniform(1e-10, 5e-10, n), "PER": rng.uniform(8e-10, 1.2e-9, n),
         "TD1": rng.uniform(0.0, 1e-10, n), "TAU1": rng.uniform(1e-11, 2e-10, n),
         "TD2": rng.uniform(3e-10, 6e-10, n), "TAU2": rng.uniform(1e-11, 2e-10, n)}
    p["TRANSIENTSOURCETYPE"][kind == -1] = np.nan
    for k in ("TD", "TR", "TF", "PW", "PER", "FREQ", "THETA", "TD1", "TAU1", "TD2", "TAU2"):
        p[k][rng.random(n) < defaults_frac] = np.nan
    return {"models": [(1, {})], "n": n, "model_index": np.zeros(n, np.int32),
            "inst_params": p}
