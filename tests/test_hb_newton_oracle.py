"""The oracle's harmonic-balance Newton, pinned by what the mathematics fixes:

* a linear RC low-pass driven by a sinusoid: the HB solution is the phasor
  solution, V_out(k) = V_in(k) / (1 + j k w R C), exactly (to the Newton
  tolerance) at every harmonic;
* a memoryless diode-resistor circuit (no charges): with 2H + 1 samples and
  2H + 1 unknown real spectral components, HB's residual vanishing means
  KCL holds at every sample, so each sample's voltage is the scalar root of
  (V_in - V) / R = IS (exp(V / Vt) - 1) + GMIN V;
* a planted sign error in the j w k Q term fails the RC check.
"""
import numpy as np
import pytest
from scipy.optimize import brentq

import oracle.hb as ohb
from hb_newton_oracle import OracleHB
from xyce_amd import default_options
from xyce_amd.hb import sample_steps
from xyce_amd.netlist_reader import Netlist

T = 1e-9
RC_TEXT = "rc low-pass\nV1 in 0 DC 0\nR1 in out 1000\nC1 out 0 1e-12\n"
DIODE_TEXT = "diode clipper\n.model dm d is=1e-14\nV1 in 0 DC 0\nR1 in a 1000\nD1 a 0 dm\n"
# Boltzmann's constant and the electron charge as the device models take them
KB, QE, TNOM = 1.3806226e-23, 1.6021918e-19, 300.15


def _opts():
    return default_options(voltageLimiterFlag=0)


def rc_expected(N, amp=0.1, dc=0.5, R=1000.0, C=1e-12):
    """Spectrum of V(out) in the permuted layout for V(in) = dc + amp sin(w t)."""
    H = (N - 1) // 2
    vin = np.zeros(H + 1, complex)
    vin[0] = dc
    vin[1] = -0.5j * amp           # (1/N) sum_s x_s e^{-j w k t_s} of amp sin(w t)
    w = 2 * np.pi / T
    return np.array([vin[k] / (1 + 1j * k * w * R * C) for k in range(H + 1)])


def _rc_solution(N, monkey=None):
    nl = Netlist(RC_TEXT + sample_steps({"V1": lambda t: 0.5 + 0.1 * np.sin(2 * np.pi * t / T)},
                                        N, T))
    hb = OracleHB(nl, T, _opts())
    if monkey:
        monkey()
    xf, its = hb.solve(np.zeros((hb.n, 2 * N)))
    return nl, xf, its


@pytest.mark.parametrize("N", [5, 9])
def test_rc_phasor(N):
    nl, xf, its = _rc_solution(N)
    out = nl.node("out")
    want = rc_expected(N)
    got = xf[out, 0::2][:(N + 1) // 2] + 1j * xf[out, 1::2][:(N + 1) // 2]
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-13)
    assert its <= 3           # linear: one Newton step, one to confirm


def test_rc_phasor_catches_wrong_jwq_sign(monkeypatch):
    good = ohb.add_jwq
    flipped = lambda ff, qf, period: good(ff, -np.asarray(qf), period)
    nl, xf, _ = _rc_solution(5, monkey=lambda: monkeypatch.setattr(ohb, "add_jwq", flipped))
    out = nl.node("out")
    got = xf[out, 2] + 1j * xf[out, 3]
    assert abs(got - rc_expected(5)[1]) > 1e-3


def diode_samples(N, amp=1.5, R=1000.0, Is=1e-14, gmin=1e-12):
    t = np.arange(N) * T / N
    vin = amp * np.sin(2 * np.pi * t / T)
    vt = KB * TNOM / QE
    f = lambda v, vi: (vi - v) / R - Is * np.expm1(v / vt) - gmin * v
    return np.array([brentq(f, -2.0, 2.0, args=(vi,), xtol=1e-15) for vi in vin])


def test_memoryless_diode_samples_are_kcl_roots():
    N = 9
    nl = Netlist(DIODE_TEXT + sample_steps({"V1": lambda t: 1.5 * np.sin(2 * np.pi * t / T)},
                                           N, T))
    hb = OracleHB(nl, T, _opts())
    xf, _ = hb.solve(np.zeros((hb.n, 2 * N)))
    xt = ohb.ift_permuted(xf)
    np.testing.assert_allclose(xt[:, nl.node("a")], diode_samples(N), rtol=0, atol=1e-10)
