"""xa_batch_newton_step on hand-made systems: the banded solve against a
dense numpy solve, and the failure reports (a singular circuit gives NaN and
leaves its unknowns alone; the host driver then stops).
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _step(blocks, f, dFdxdVp, x0, kl=2, ku=2):
    """One DC step (rhs = -f + dFdxdVp, J = dFdx) of circuits given as dense
    blocks; returns (x, dx_max) from the device."""
    import torch
    import xyce_amd
    from xyce_amd._abi import check, lib
    sizes = [b.shape[0] for b in blocks]
    c_start = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    n = int(c_start[-1])
    rows, cols, vals = [], [], []
    for c, b in enumerate(blocks):
        r, k = np.nonzero(b)
        rows.append(r + c_start[c])
        cols.append(k + c_start[c])
        vals.append(b[r, k])
    rows, cols, vals = (np.concatenate(a) for a in (rows, cols, vals))
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    row_ptr = np.zeros(n + 1, np.int64)
    np.add.at(row_ptr, rows + 1, 1)
    row_ptr = np.cumsum(row_ptr).astype(np.int32)
    ctx = xyce_amd.Context(0)
    dev = torch.device("cuda", 0)
    t = lambda a, dt=torch.float64: torch.from_numpy(np.ascontiguousarray(a)).to(dev, dt)
    d = dict(c_start=t(c_start, torch.int32), row_ptr=t(row_ptr, torch.int32),
             col_idx=t(cols.astype(np.int32), torch.int32), dFdx=t(vals),
             dQdx=t(np.zeros(len(vals))), f=t(f), q=t(np.zeros(n)), dFdxdVp=t(dFdxdVp),
             dQdxdVp=t(np.zeros(n)), q_prev=t(np.zeros(n)), c_lin=t(np.zeros(n)),
             held=t(np.zeros(n, np.uint8), torch.uint8), x=t(x0), q_now=t(np.zeros(n)),
             dx_max=t(np.zeros(len(blocks))))
    P = lambda k: C.c_void_p(d[k].data_ptr())
    check(lib().xa_batch_newton_step(
        ctx.h, len(blocks), max(sizes), kl, ku, P("c_start"), None, P("row_ptr"), P("col_idx"),
        P("dFdx"), P("dQdx"), P("f"), P("q"), P("dFdxdVp"), P("dQdxdVp"), P("q_prev"),
        P("c_lin"), P("held"), C.c_double(1.0), 1, P("x"), P("q_now"), P("dx_max")),
        "xa_batch_newton_step")
    torch.cuda.synchronize()
    return d["x"].cpu().numpy(), d["dx_max"].cpu().numpy(), c_start


def test_banded_solve_matches_dense():
    rng = np.random.default_rng(3)
    blocks = []
    for n in (5, 9, 1, 17):
        a = np.zeros((n, n))
        for i in range(n):
            for j in range(max(0, i - 2), min(n, i + 3)):
                a[i, j] = rng.uniform(-1, 1)
        a[np.arange(n), np.arange(n)] += 0.2   # pivoting still happens on most columns
        blocks.append(a)
    nt = sum(b.shape[0] for b in blocks)
    f = rng.uniform(-1, 1, nt)
    lim = rng.uniform(-0.1, 0.1, nt)
    x0 = rng.uniform(-1, 1, nt)
    x, dx_max, cs = _step(blocks, f, lim, x0)
    for c, b in enumerate(blocks):
        s = slice(cs[c], cs[c + 1])
        dx = np.linalg.solve(b, -f[s] + lim[s])
        np.testing.assert_allclose(x[s] - x0[s], dx, rtol=1e-10, atol=1e-12)
        assert abs(dx_max[c] - np.max(np.abs(dx))) <= 1e-10 * np.max(np.abs(dx))


def test_singular_circuit_reports_nan():
    """A floating node (an empty row and column) makes the circuit's Jacobian
    singular: its dx_max is NaN and its unknowns keep their values; the other
    circuit of the batch is solved as usual."""
    good = np.array([[2.0, -1.0, 0.0], [-1.0, 2.0, -1.0], [0.0, -1.0, 2.0]])
    floating = np.array([[2.0, 0.0, -1.0], [0.0, 0.0, 0.0], [-1.0, 0.0, 2.0]])
    f = np.array([1.0, 0.5, -0.25, 1.0, 0.0, -1.0])
    x0 = np.arange(6, dtype=np.float64)
    x, dx_max, _ = _step([good, floating], f, np.zeros(6), x0)
    assert np.isfinite(dx_max[0])
    np.testing.assert_allclose(x[:3] - x0[:3], np.linalg.solve(good, -f[:3]), rtol=1e-12)
    assert np.isnan(dx_max[1])
    np.testing.assert_array_equal(x[3:], x0[3:])
