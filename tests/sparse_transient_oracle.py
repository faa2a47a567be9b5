"""Fixed-step transient of a netlist circuit on the oracle -- TEST
INFRASTRUCTURE: the CPU restatement the device sparse transient
(xyce_amd.sparse_transient) is checked against.  Same integrator formulas
(DC operating point with junction initialisation; backward Euler for the
first two steps, then the one-step trapezoid rule), the oracle's loads
(CircuitPair without a GPU, the reader's single-copy layout) and a direct
sparse LU solve (SuperLU) of every Newton system instead of GMRES + ILU(0)."""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from circuit import CircuitPair, Part


def parts_of(nl, step=-1):
    return [Part(p["type"], p["models"], p["model_index"], p["params"], p["ext_nodes"])
            for p in nl.parts(step)]


class OracleTransient:
    def __init__(self, nl, options=None, vtol=1e-9):
        self.cp = CircuitPair(parts_of(nl), nl.n_nodes, gpu=False, options=options)
        self.n = self.cp.n_nodes
        self.nv = nl.n_nodes
        self.vtol = vtol
        rp, ci = self.cp.row_ptr, self.cp.col_idx
        self.rows = np.repeat(np.arange(self.n), np.diff(rp))
        self.cols = ci

    def _mat(self, vals):
        return sp.csc_matrix((vals, (self.rows, self.cols)), shape=(self.n, self.n))

    def _newton(self, x, flags_of, system, max_iter=100):
        cp = self.cp
        for it in range(max_iter):
            cp.set_inputs(nextSol=x)
            ro, _ = cp.load(**flags_of(it))
            self.last = ro
            rhs, a, b = system(ro)
            J = self._mat(a * ro["dFdx"] + b * ro["dQdx"])
            dx = spla.spsolve(J, rhs)
            x_load = x
            x = x + dx
            if it > 0 and np.max(np.abs(dx[:self.nv])) <= self.vtol:
                return x, x_load, it + 1
        raise RuntimeError("oracle Newton did not converge")

    def _accept(self):
        self.cp.set_inputs(currSta=self.last["nextSta"].copy(),
                           currSto=self.last["nextSto"].copy())

    def run(self, x0, t_stop, h, method="trap", record=()):
        times = dict(startingTimeStep=h, finalTime=t_stop, bpTol=1e-18)

        def dc_flags(it):
            return dict(newtonIter=it, dcopFlag=1, tranopFlag=1,
                        initJctFlag=1 if it == 0 else 0, currTime=0.0, **times)

        x, _, its = self._newton(np.array(x0, dtype=np.float64), dc_flags,
                                 lambda ro: (-ro["f"] + ro["dFdxdVp"], 1.0, 0.0))
        self._accept()
        newton = [its]
        q_n, f_n = self.last["q"].copy(), self.last["f"].copy()
        rec = np.asarray(record, np.int64)
        out_v = [x[rec].copy()]
        out_t = [0.0]
        for k in range(int(round(t_stop / h))):
            tk = (k + 1) * h
            first = k == 0

            def flags(it, first=first, tk=tk):
                return dict(newtonIter=it, transientFlag=1, pdt=1.0 / h, currTimeStep=h,
                            lastTimeStep=h, currTime=tk, initTranFlag=1 if first else 0,
                            beginIntegrationFlag=1 if first else 0, **times)

            trap = method == "trap" and k >= 2

            def system(ro, q_n=q_n, f_n=f_n, trap=trap):
                if trap:
                    return (-((ro["q"] - q_n) / h + 0.5 * ro["f"] + 0.5 * f_n)
                            + 0.5 * ro["dFdxdVp"] + ro["dQdxdVp"] / h), 0.5, 1.0 / h
                return (-((ro["q"] - q_n) / h + ro["f"]) + ro["dFdxdVp"] + ro["dQdxdVp"] / h,
                        1.0, 1.0 / h)

            x, _, its = self._newton(x, flags, system)
            self._accept()
            newton.append(its)
            q_n, f_n = self.last["q"].copy(), self.last["f"].copy()
            out_t.append(tk)
            out_v.append(x[rec].copy())
        return np.asarray(out_t), np.asarray(out_v), np.asarray(newton)
