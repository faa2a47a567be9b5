"""The device harmonic-balance Newton-Krylov solve (xyce_amd.hb.HBSolver):
every time sample's copy of the circuit loaded in one pass, matrix-free
Jacobian products through xa_csr_spmv and the transforms, flexible GMRES
with the FP32 block-Jacobi preconditioner.

* RC low-pass: the phasor solution at every harmonic (closed form);
* memoryless diode clipper: every sample at the scalar KCL root;
* a 25-transistor BSIM4 differential amplifier with its RC filter and
  sources: the same fixed point as the oracle's exact HB Newton (dense
  Jacobian, LAPACK), waveforms within 1e-9 V.
"""
import numpy as np
import pytest

import oracle.hb as ohb
from hb_newton_oracle import OracleHB
from test_hb_newton_oracle import DIODE_TEXT, RC_TEXT, T, diode_samples, rc_expected

pytestmark = pytest.mark.gpu


def _solver(text, N, sources):
    import xyce_amd
    from xyce_amd.hb import HBSolver, sample_steps
    from xyce_amd.netlist_reader import Circuit, Netlist
    nl = Netlist(text + sample_steps(sources, N, T))
    ctx = xyce_amd.Context(0)
    ckt = Circuit(ctx, nl, step=-2, options=xyce_amd.default_options(voltageLimiterFlag=0))
    return nl, HBSolver(ckt, T)


@pytest.mark.parametrize("N", [9, 65])
def test_rc_phasor(N):
    import torch
    nl, hb = _solver(RC_TEXT, N, {"V1": lambda t: 0.5 + 0.1 * np.sin(2 * np.pi * t / T)})
    xf, st = hb.solve(torch.zeros((hb.n, 2 * N), dtype=torch.float64, device="cuda"),
                      vtol=1e-12, krylov_tol=1e-13, verbose=True)
    xf = xf.cpu().numpy()
    out = nl.node("out")
    H = (N - 1) // 2
    got = xf[out, 0::2][:H + 1] + 1j * xf[out, 1::2][:H + 1]
    np.testing.assert_allclose(got, rc_expected(N), rtol=0, atol=1e-12)
    print(f"N={N}: Newton {st['newton']}, matvecs {st['matvecs']}", flush=True)


@pytest.mark.parametrize("N", [9, 33])
def test_memoryless_diode_kcl_roots(N):
    import torch
    nl, hb = _solver(DIODE_TEXT, N, {"V1": lambda t: 1.5 * np.sin(2 * np.pi * t / T)})
    xf, st = hb.solve(torch.zeros((hb.n, 2 * N), dtype=torch.float64, device="cuda"),
                      vtol=1e-12, krylov_tol=1e-12, verbose=True)
    xt = ohb.ift_permuted(xf.cpu().numpy())
    np.testing.assert_allclose(xt[:, nl.node("a")], diode_samples(N), rtol=0, atol=1e-10)
    print(f"N={N}: Newton {st['newton']}, final |R| {st['final_residual_max']:.2e}", flush=True)


def test_diffamp_matches_oracle_hb():
    import torch
    import xyce_amd
    from xyce_amd import netlists
    N = 9
    text, v0 = netlists.diffamp_netlist(25)
    drive = lambda t: 0.01 * np.sin(2 * np.pi * t / T)
    nl, hb = _solver(text, N, {"VINP": lambda t: 0.5 + drive(t), "VINN": lambda t: 0.5 - drive(t)})
    x0 = np.zeros(hb.n)
    for k, name in enumerate(nl.node_names):
        x0[k] = v0[int(name[1:])]
    xt0 = np.tile(x0, (N, 1))
    xf0 = ohb.dft_permuted(xt0)
    print(f"amplifier: {hb.n} unknowns per sample, {N} samples", flush=True)
    xf, st = hb.solve(torch.from_numpy(xf0).cuda(), vtol=1e-11, krylov_tol=1e-11, verbose=True)
    got = ohb.ift_permuted(xf.cpu().numpy())
    ref_hb = OracleHB(nl, T, xyce_amd.default_options(voltageLimiterFlag=0))
    xf_o, its = ref_hb.solve(xf0, vtol=1e-12)
    want = ohb.ift_permuted(xf_o)
    err = np.max(np.abs(got - want))
    print(f"device Newton {st['newton']} (matvecs {st['matvecs']}), oracle Newton {its}; "
          f"max |v_device - v_oracle| {err:.2e} V", flush=True)
    assert err < 1e-9
    # the drive reaches the output filter
    filt = [nl.node(f"n{len(v0) - 2}"), nl.node(f"n{len(v0) - 1}")]
    assert np.ptp(want[:, filt[0]]) > 1e-4
