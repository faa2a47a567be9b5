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
