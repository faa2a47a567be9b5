"""Harmonic-balance Newton on the oracle -- TEST INFRASTRUCTURE: the CPU
restatement the device HB solve is checked against.

Each time sample is the netlist's circuit at that sample's `.STEP` point,
loaded by the oracle (one CircuitPair per sample, the reader's single-step
layout, which is the batched build's copy layout).  The residual and the
Jacobian follow N_LOA_HBLoader.C (:540-775 the load, :350-456 the product):
R = DFT F + j w k DFT Q, J = DFT dF/dx IFT + j w k DFT dQ/dx IFT, assembled
dense over the independent half of the spectrum (Re_0, then Re_k, Im_k for
k = 1..H) and solved with LAPACK -- no Krylov solve, no preconditioner, so
the device solve's fixed point is checked against an exact Newton's.
"""
import numpy as np

import oracle.hb as ohb
from circuit import CircuitPair, Part

FLAGS = dict(transientFlag=1, newtonIter=1)


def parts_of(nl, step):
    return [Part(p["type"], p["models"], p["model_index"], p["params"], p["ext_nodes"])
            for p in nl.parts(step)]


def pack(xf):
    N = xf.shape[1] // 2
    return np.concatenate([xf[:, 0:1], xf[:, 2:N + 1]], axis=1)


def unpack(y):
    n, N = y.shape
    H = (N - 1) // 2
    xf = np.zeros((n, 2 * N))
    xf[:, 0] = y[:, 0]
    xf[:, 2:N + 1] = y[:, 1:]
    for k in range(1, H + 1):
        xf[:, 2 * (N - k)] = y[:, 2 * k - 1]
        xf[:, 2 * (N - k) + 1] = -y[:, 2 * k]
    return xf


def dense(rp, ci, vals, n):
    A = np.zeros((n, n))
    for r in range(n):
        for k in range(rp[r], rp[r + 1]):
            A[r, ci[k]] += vals[k]
    return A


class OracleHB:
    def __init__(self, nl, period, options):
        self.N = nl.n_steps
        if self.N < 1 or self.N % 2 == 0:
            raise ValueError("odd sample count expected")
        self.period = float(period)
        self.cps = [CircuitPair(parts_of(nl, s), nl.n_nodes, gpu=False, options=options)
                    for s in range(self.N)]
        self.n = self.cps[0].n_nodes

    def load(self, xt):
        F, Q, G, C = [], [], [], []
        for s, cp in enumerate(self.cps):
            cp.set_inputs(nextSol=xt[s], currSol=xt[s])
            ro, _ = cp.load(**FLAGS)
            F.append(ro["f"])
            Q.append(ro["q"])
            G.append(dense(cp.row_ptr, cp.col_idx, ro["dFdx"], self.n))
            C.append(dense(cp.row_ptr, cp.col_idx, ro["dQdx"], self.n))
        return np.array(F), np.array(Q), G, C

    def residual(self, xf):
        F, Q, G, C = self.load(ohb.ift_permuted(xf))
        R = ohb.add_jwq(ohb.dft_permuted(F), ohb.dft_permuted(Q), self.period)
        return R, G, C

    def jacobian(self, G, C):
        n, N = self.n, self.N
        J = np.zeros((n * N, n * N))
        for col in range(n * N):
            y = np.zeros(n * N)
            y[col] = 1.0
            vt = ohb.ift_permuted(unpack(y.reshape(n, N)))
            g = np.array([G[s] @ vt[s] for s in range(N)])
            c = np.array([C[s] @ vt[s] for s in range(N)])
            out = ohb.add_jwq(ohb.dft_permuted(g), ohb.dft_permuted(c), self.period)
            J[:, col] = pack(out).reshape(-1)
        return J

    def solve(self, xf0, vtol=1e-12, max_newton=60, max_step=0.5):
        xf = np.array(xf0, dtype=np.float64)
        for it in range(max_newton):
            R, G, C = self.residual(xf)
            J = self.jacobian(G, C)
            dy = np.linalg.solve(J, -pack(R).reshape(-1)).reshape(self.n, self.N)
            dxf = unpack(dy)
            step = np.max(np.abs(ohb.ift_permuted(dxf)))
            scale = min(1.0, max_step / step) if step > 0 else 1.0
            xf = xf + scale * dxf
            if scale == 1.0 and step <= vtol:
                return xf, it + 1
        raise RuntimeError("oracle HB Newton did not converge")
