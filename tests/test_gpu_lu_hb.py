"""The batched FP32 LU on the matrix cores and the harmonic-balance path on
the GPU.

* LU: P A = L U reconstructs every matrix of a batch (random matrices with
  zero diagonal entries, so the pivoting is exercised) to FP32 accuracy, the
  solve's backward error is at FP32 level, the identity padding stays
  untouched, and a zero pivot is reported in info.
* HB transforms: the permuted forward / inverse transforms and the j w k Q
  term against the oracle's restatement (oracle/hb.py), to 1e-13 of the
  vectors' size.
* HB load: one load of every sample of a small differential amplifier at
  once -- the time-domain outputs against the oracle's loads at the same
  samples (assert_parity), the frequency-domain vectors against the oracle's
  transforms of those outputs.
* HB block solve: the block-Jacobi preconditioner factored in FP32 and
  refined in FP64, against an FP64 complex solve of every harmonic's system;
  the FP32-only residual, the refined residual, the refinement steps and the
  forward error are printed per harmonic.  Small netlist against numpy, and
  the full 129-sample, 2000-transistor amplifier against torch's FP64 solve.
"""
import time

import numpy as np
import pytest
import torch

from circuit import CircuitPair, Part
from harness import assert_parity
from oracle import hb as ohb
import xyce_amd
from xyce_amd import Context, default_solver_state, make_vectors, netlists
from xyce_amd.hb import HBPlan, HBPrecond, lu_f32_factor, lu_f32_solve

pytestmark = pytest.mark.gpu
# conductance to ground added on every diagonal of the amplifier's Jacobians:
# the gate-only nodes (bias, inputs) have no conductance of their own, and
# their sources are not in this BSIM4-only netlist
SHUNT = 1e-5


def _padded(mats, ld):
    a = np.zeros((len(mats), ld, ld), np.float32)
    for b, m in enumerate(mats):
        n = len(m)
        a[b, :n, :n] = m
        a[b, np.arange(n, ld), np.arange(n, ld)] = 1.0
    return torch.from_numpy(a).cuda()


def _apply_ipiv(A, ipiv, rows):
    A = A.copy()
    for i in range(rows):
        p = int(ipiv[i])
        if p != i:
            A[[i, p]] = A[[p, i]]
    return A


@pytest.mark.parametrize("m", [37, 300, 1100])
def test_lu_f32_factor_and_solve(m):
    rng = np.random.default_rng(100 + m)
    batch, ld = 3, (m + 63) // 64 * 64
    mats = []
    for _ in range(batch):
        A = rng.standard_normal((m, m))
        A[np.arange(0, m, 7), np.arange(0, m, 7)] = 0.0   # zero diagonals: pivoting needed
        mats.append(A)
    ctx = Context(0)
    print(f"LU batch {batch} x {m} (ld {ld})", flush=True)
    lu = _padded(mats, ld)
    ipiv, info = lu_f32_factor(ctx, lu, m)
    torch.cuda.synchronize()
    assert info.cpu().numpy().tolist() == [0] * batch
    LUh = lu.cpu().numpy().astype(np.float64)
    piv = ipiv.cpu().numpy()
    mr = (m + 63) // 64 * 64
    for b in range(batch):
        Ap = np.eye(ld)
        Ap[:m, :m] = mats[b]
        PA = _apply_ipiv(Ap, piv[b], mr)
        L = np.tril(LUh[b], -1) + np.eye(ld)
        U = np.triu(LUh[b])
        err = np.linalg.norm(PA - L @ U) / (np.linalg.norm(L) * np.linalg.norm(U))
        print(f"  matrix {b}: ||PA - LU|| / ||L|| ||U|| = {err:.2e}", flush=True)
        assert err < 1e-6
        # the identity padding is untouched
        np.testing.assert_array_equal(LUh[b, m:, m:], np.eye(ld - m))
        np.testing.assert_array_equal(LUh[b, :m, m:], 0.0)
    rhs = np.zeros((batch, 2, ld), np.float32)
    rhs[:, :, :m] = rng.standard_normal((batch, 2, m))
    x = torch.from_numpy(rhs.copy()).cuda()
    lu_f32_solve(ctx, lu, ipiv, m, x)
    xh = x.cpu().numpy().astype(np.float64)
    for b in range(batch):
        A = mats[b]
        for r in range(2):
            res = A @ xh[b, r, :m] - rhs[b, r, :m]
            be = np.max(np.abs(res)) / (np.max(np.abs(A).sum(axis=1)) * np.max(np.abs(xh[b, r, :m])))
            print(f"  matrix {b} rhs {r}: backward error {be:.2e}", flush=True)
            assert be < 1e-5
            assert np.all(xh[b, r, m:] == 0.0)


def test_lu_zero_pivot_reported():
    rng = np.random.default_rng(7)
    m = 100
    good = rng.standard_normal((m, m))
    bad = rng.standard_normal((m, m))
    bad[:, 5] = 0.0
    ctx = Context(0)
    lu = _padded([good, bad], 128)
    _, info = lu_f32_factor(ctx, lu, m)
    assert info.cpu().numpy().tolist() == [0, 6]


def test_hb_transforms_match_oracle():
    rng = np.random.default_rng(11)
    N, n, T = 129, 301, 1e-9
    ctx = Context(0)
    plan = HBPlan(ctx, N, n, T)
    xt = rng.standard_normal((N, n))
    dxt = torch.from_numpy(xt).cuda()
    xf = torch.zeros((n, 2 * N), dtype=torch.float64, device="cuda")
    plan.dft(dxt, xf)
    ref = ohb.dft_permuted(xt)
    np.testing.assert_allclose(xf.cpu().numpy(), ref, rtol=0, atol=1e-13 * np.abs(ref).max())
    back = torch.zeros_like(dxt)
    plan.ift(xf, back)
    np.testing.assert_allclose(back.cpu().numpy(), xt, rtol=0, atol=1e-12 * np.abs(xt).max())
    qt = rng.standard_normal((N, n))
    qf = torch.from_numpy(ohb.dft_permuted(qt)).cuda()
    ff = xf.clone()
    plan.add_jwq(ff, qf)
    want = ohb.add_jwq(xf.cpu().numpy(), qf.cpu().numpy(), T)
    np.testing.assert_allclose(ff.cpu().numpy(), want, rtol=0, atol=1e-14 * np.abs(want).max())


def _hb_pair(samples, n_transistors):
    cards, mi, params, lids, n_nodes, volts = netlists.hb_diffamp(samples, n_transistors)
    cp = CircuitPair([Part("bsim4", cards, mi, params, lids)], n_nodes, noise_runs=4)
    assert cp.n_nodes == n_nodes and n_nodes % samples == 0
    return cp, n_nodes // samples, volts


def _diag_slots(row_ptr, col_idx):
    n = len(row_ptr) - 1
    d = np.full(n, -1)
    for i in range(n):
        cols = col_idx[row_ptr[i]:row_ptr[i + 1]]
        hit = np.nonzero(cols == i)[0]
        if len(hit):
            d[i] = row_ptr[i] + hit[0]
    assert (d >= 0).all(), "every row of the pattern holds its diagonal"
    return d


def test_hb_load_and_block_solve_small():
    N, T = 33, 1e-9
    cp, n, volts = _hb_pair(N, 40)
    print(f"HB load: {N} samples x {n} unknowns, {cp.parts[0].n} BSIM4 instances", flush=True)
    rng = np.random.default_rng(21)
    cp.set_inputs(currSto=rng.uniform(0.0, 1.0, cp.n_sto), nextSto=rng.uniform(0.0, 1.0, cp.n_sto))
    xf_h = ohb.dft_permuted(volts.reshape(N, n))
    xf = torch.from_numpy(xf_h).cuda()
    plan = HBPlan(cp.ctx, N, n, T)
    flags = dict(newtonIter=1, transientFlag=1, pdt=1e12)
    outs = {k: torch.zeros((n, 2 * N), dtype=torch.float64, device="cuda")
            for k in ("ff", "qf", "dfp", "dqp")}
    plan.load(cp.gpu[0], cp._ss(xyce_amd, flags), xf, make_vectors(**cp.d), cp.nnz,
              outs["ff"], outs["qf"], outs["dfp"], outs["dqp"])
    torch.cuda.synchronize()
    keys = ("f", "q", "dFdxdVp", "dQdxdVp", "dFdx", "dQdx")
    tdom = {k: cp.d[k][:len(cp.h[k])].cpu().numpy().copy() for k in keys}
    xt = cp.d["nextSol"][:N * n].cpu().numpy()
    np.testing.assert_allclose(xt, ohb.ift_permuted(xf_h).reshape(-1), rtol=0, atol=1e-13)
    # the oracle at the same samples and inputs (the HB load wrote the GPU's
    # store and state vectors: put the inputs back)
    cp.set_inputs(nextSol=xt, **{k: cp.h[k].copy() for k in
                                 ("nextSto", "currSto", "nextSta", "currSta", "lastSta")})
    ro, rg = cp.load(**flags)
    for k in keys:
        st = assert_parity(tdom[k], ro[k], what=f"time-domain {k}", scale=cp.mag.get(k),
                           noise=cp.noise.get(k))
        print(f"  {k:8s} bit-exact {st['bitexact']:.4f} max rel {st['max_rel']:.2e}", flush=True)
    # the frequency-domain vectors are the transforms of those outputs
    for name, fk, qk in (("ff", "f", "q"), ("dfp", "dFdxdVp", "dQdxdVp")):
        want = ohb.add_jwq(ohb.dft_permuted(tdom[fk].reshape(N, n)),
                           ohb.dft_permuted(tdom[qk].reshape(N, n)), T)
        got = outs[name].cpu().numpy()
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-13 * max(np.abs(want).max(), 1e-300))
    for name, k in (("qf", "q"), ("dqp", "dQdxdVp")):
        want = ohb.dft_permuted(tdom[k].reshape(N, n))
        np.testing.assert_allclose(outs[name].cpu().numpy(), want, rtol=0,
                                   atol=1e-13 * max(np.abs(want).max(), 1e-300))
    # block-Jacobi solve from the per-sample Jacobians (a shunt on every
    # diagonal stands for the sources that hold the input and bias nodes)
    nnz1 = cp.row_ptr[n]
    rp, ci = cp.row_ptr[:n + 1], cp.col_idx[:nnz1]
    assert len(cp.col_idx) == N * nnz1
    diag = _diag_slots(rp, ci)
    dfdx = cp.d["dFdx"][:N * nnz1].clone()
    dfdx.view(N, nnz1)[:, torch.from_numpy(diag).cuda()] += SHUNT
    dqdx = cp.d["dQdx"][:N * nnz1].clone()
    pc = HBPrecond(plan, rp, ci)
    info = pc.factor(dfdx, dqdx)
    assert (info == 0).all(), info
    bf_h = rng.standard_normal((n, 2 * N))
    bf = torch.from_numpy(bf_h).cuda()
    xs = torch.zeros_like(bf)
    r32, rfin, its = pc.solve(bf, xs, max_refine=30, tol=1e-13)
    x = xs.cpu().numpy()
    Gd = np.zeros((n, n))
    Cd = np.zeros((n, n))
    rows = np.repeat(np.arange(n), np.diff(rp))
    G = dfdx.view(N, nnz1).cpu().numpy().sum(axis=0) * (1.0 / N)
    Cv = dqdx.view(N, nnz1).cpu().numpy().sum(axis=0) * (1.0 / N)
    np.add.at(Gd, (rows, ci), G)
    np.add.at(Cd, (rows, ci), Cv)
    w = 2 * np.pi / T
    worst = 0.0
    for p in range(N):
        k = p if p <= (N - 1) // 2 else p - N
        z = np.linalg.solve(Gd + 1j * k * w * Cd, bf_h[:, 2 * p] + 1j * bf_h[:, 2 * p + 1])
        got = x[:, 2 * p] + 1j * x[:, 2 * p + 1]
        fe = np.max(np.abs(got - z)) / np.max(np.abs(z))
        worst = max(worst, fe)
        print(f"  harmonic {k:4d}: FP32 solve rel residual {r32[p]:.2e}, refined {rfin[p]:.2e} "
              f"after {its[p]} steps, forward error vs FP64 {fe:.2e}", flush=True)
        assert rfin[p] <= 1e-13
    assert worst < 1e-9


def test_hb_block_solve_full_size():
    """129 samples of the 2000-transistor amplifier: 65 FP32 blocks of 2408
    rows (ld 2432), refined against torch's FP64 solve."""
    from xyce_amd.workloads import Workload
    N, T = 129, 1e-9
    cards, mi, params, lids, n_nodes, volts = netlists.hb_diffamp(N, 2000)
    n = n_nodes // N
    wl = Workload("bsim4", len(mi), cards=cards, model_index=mi, params=params,
                  netlist=(lids, n_nodes, volts))
    nnz1 = int(wl.row_ptr[n])
    assert wl.nnz == N * nnz1
    print(f"HB full size: {N} samples x {n} unknowns, {len(mi)} instances, nnz/sample {nnz1}",
          flush=True)
    plan = HBPlan(wl.ctx, N, n, T)
    xf = torch.from_numpy(ohb.dft_permuted(volts.reshape(N, n))).cuda()
    ff = torch.zeros((n, 2 * N), dtype=torch.float64, device="cuda")
    plan.load(wl.master, default_solver_state(transientFlag=1, newtonIter=1, pdt=1e12), xf,
              wl.vectors, wl.nnz, ff)
    rp, ci = wl.row_ptr[:n + 1], wl.col_idx[:nnz1]
    diag = _diag_slots(rp, ci)
    dfdx = wl.vec["dFdx"][:N * nnz1].clone()
    dfdx.view(N, nnz1)[:, torch.from_numpy(diag).cuda()] += SHUNT
    dqdx = wl.vec["dQdx"][:N * nnz1].clone()
    pc = HBPrecond(plan, rp, ci)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    info = pc.factor(dfdx, dqdx)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    assert (info == 0).all(), info
    rng = np.random.default_rng(31)
    bf = torch.from_numpy(rng.standard_normal((n, 2 * N))).cuda()
    xs = torch.zeros_like(bf)
    r32, rfin, its = pc.solve(bf, xs, max_refine=30, tol=1e-13)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    print(f"  factor {1e3 * (t1 - t0):.1f} ms ({(N + 1) // 2} blocks), solve with refinement "
          f"{1e3 * (t2 - t1):.1f} ms", flush=True)
    # FP64 reference: every harmonic's complex system on the GPU
    rows = torch.from_numpy(np.repeat(np.arange(n), np.diff(rp))).cuda()
    cols = torch.from_numpy(ci.astype(np.int64)).cuda()
    G = dfdx.view(N, nnz1).sum(dim=0) * (1.0 / N)
    Cv = dqdx.view(N, nnz1).sum(dim=0) * (1.0 / N)
    Gd = torch.zeros((n, n), dtype=torch.float64, device="cuda").index_put_((rows, cols), G,
                                                                            accumulate=True)
    Cd = torch.zeros((n, n), dtype=torch.float64, device="cuda").index_put_((rows, cols), Cv,
                                                                            accumulate=True)
    w = 2 * np.pi / T
    ks = [p if p <= (N - 1) // 2 else p - N for p in range(N)]
    x = xs.cpu().numpy()
    b = bf.cpu().numpy()
    worst = 0.0
    for p0 in range(0, N, 16):
        ps = list(range(p0, min(N, p0 + 16)))
        K = torch.stack([Gd.to(torch.complex128) + 1j * ks[p] * w * Cd.to(torch.complex128)
                         for p in ps])
        rhs = torch.stack([torch.from_numpy(b[:, 2 * p] + 1j * b[:, 2 * p + 1]) for p in ps]).cuda()
        z = torch.linalg.solve(K, rhs.unsqueeze(-1)).squeeze(-1).cpu().numpy()
        for q, p in enumerate(ps):
            got = x[:, 2 * p] + 1j * x[:, 2 * p + 1]
            fe = np.max(np.abs(got - z[q])) / np.max(np.abs(z[q]))
            worst = max(worst, fe)
            print(f"  harmonic {ks[p]:4d}: FP32 solve rel residual {r32[p]:.2e}, refined "
                  f"{rfin[p]:.2e} after {its[p]} steps, forward error vs FP64 {fe:.2e}", flush=True)
            assert rfin[p] <= 1e-12
    assert worst < 1e-9
