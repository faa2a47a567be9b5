"""Fixed-step transient of a netlist circuit on one GPU, with a sparse
Newton solve: every iteration loads all of the circuit's device groups
(``netlist_reader.Circuit``) and solves J dx = rhs by GMRES preconditioned
with ILU(0), on the device.

The Newton residual and Jacobian are formed the way the reference's
integrators form them (backward Euler N_TIA_BackwardDifferentiation15.C:
215-265; the one-step integrator's trapezoid rule N_TIA_OneStep.C:272-335,
463-500, order 1 for the first two steps as N_TIA_OneStep.C:2130,2211
raise it):

* DC operating point: rhs = -F + dFdx*dVp, J = dF/dx (junction
  initialisation on the first iteration, and tranopFlag set, as the
  reference's operating point of a transient, N_ANP_AnalysisManager.C:
  2749-2753, so the loads compute the charges the first step starts from);
* backward Euler: rhs = -((Q - q_n)/h + F) + dFdx*dVp + dQdx*dVp/h,
  J = dF/dx + dQ/dx / h;
* trapezoid: rhs = -((Q - q_n)/h + (F + F_n)/2) + dFdx*dVp/2 + dQdx*dVp/h,
  J = dF/dx / 2 + dQ/dx / h.

F includes the independent sources at the step's time (the sources are
devices of the circuit).  Each voltage source's branch row and its node's
row are exchanged in the solved system, so the incomplete factorisation
finds a pivot on every diagonal.  Newton stops when the largest update of a node
voltage is below vtol volts.  J lives in the circuit's CSR pattern plus
every diagonal (a voltage source's branch row has none of its own); the
reference solves these systems with KLU or preconditioned GMRES
(N_LAS_AztecOOSolver.C)."""
import sys
import time

import numpy as np

from .krylov import CSRMatrix, ILU0, fgmres


class NewtonFailure(RuntimeError):
    pass


class SparseTransient:
    def __init__(self, circuit, n_voltages=None, vtol=1e-9, krylov_tol=1e-10, restart=60,
                 max_krylov=600, order="topo", verbose=False, refactor_each_step_iter=False):
        """n_voltages: the unknowns [0, n_voltages) that are node voltages
        (the convergence test reads them; default: the netlist's nodes).
        order: the unknowns' (and rows') order in the solved system --
        "topo" (default) the block-triangular order of the first system
        solved (the DC operating point's Jacobian, see _topo_order), "rcm"
        reverse Cuthill-McKee of the Jacobian's graph, "natural" the LID
        order.  refactor_each_step_iter: factor the ILU(0) at every Newton
        iteration of a time step (default: at the first, reused after; the
        operating point refactors every iteration)."""
        import torch
        from . import default_solver_state
        self.torch = torch
        self._ss = default_solver_state
        self.c = circuit
        self.ctx = circuit.ctx
        n = circuit.n_unknowns
        self.n = n
        self.nv = circuit.netlist.n_nodes if n_voltages is None else int(n_voltages)
        self.vtol, self.krylov_tol = vtol, krylov_tol
        self.refactor_each_step_iter = refactor_each_step_iter
        self.restart, self.max_krylov = restart, max_krylov
        # rows of the system: a voltage source's branch row and its node's
        # row trade places, so that neither diagonal is structurally zero
        # (the node's KCL row has the branch current at the branch column,
        # the branch row the node voltage at the node column)
        src = np.arange(n, dtype=np.int64)
        for k in range(circuit.n_groups):
            if circuit.group(k)[1] != "vsource":
                continue
            for pos, neg, bra in circuit.group_lids(k).reshape(-1, 3):
                node = pos if pos >= 0 else neg
                if node < 0:
                    continue
                if src[node] != node or src[bra] != bra:
                    raise ValueError("two voltage sources on one node: rows cannot be paired")
                src[node], src[bra] = bra, node
        rp, ci = circuit.csr()
        self._rows = src[np.repeat(np.arange(n, dtype=np.int64), np.diff(rp))]   # src: own inverse
        self._cols = ci.astype(np.int64)
        self._src = src
        if order not in ("topo", "rcm", "natural"):
            raise ValueError("order must be 'topo', 'rcm' or 'natural'")
        self.order = order
        self.verbose = verbose
        self.J = None
        if order != "topo":
            self._build(self._static_order(order))
        f64 = dict(dtype=torch.float64, device="cuda")
        ns, nt, nz = circuit.n_sta, circuit.n_sto, circuit.nnz
        sizes = dict(nextSol=n, currSol=n, f=n, q=n, b=n, dFdxdVp=n, dQdxdVp=n,
                     nextSta=ns, currSta=ns, lastSta=ns, nextStaDeriv=ns,
                     nextSto=nt, currSto=nt, storeLeadF=nt, storeLeadQ=nt, dFdx=nz, dQdx=nz)
        self.v = {k: torch.zeros(max(m, 1), **f64) for k, m in sizes.items()}
        self.outs = ("f", "q", "b", "dFdxdVp", "dQdxdVp", "dFdx", "dQdx", "storeLeadF",
                     "storeLeadQ")
        self.stats = {"loads": 0, "newton": 0, "krylov": 0, "load_s": 0.0, "factor_s": 0.0,
                      "krylov_s": 0.0}

    def _static_order(self, order):
        n = self.n
        if order == "natural":
            return np.arange(n, dtype=np.int64)
        import scipy.sparse as sps
        from scipy.sparse.csgraph import reverse_cuthill_mckee
        g = sps.csr_matrix((np.ones(len(self._rows)), (self._rows, self._cols)), shape=(n, n))
        return np.asarray(reverse_cuthill_mckee((g + g.T).tocsr(), symmetric_mode=True), np.int64)

    def _topo_order(self, vals):
        """Unknown order from the numerically nonzero entries of a Jacobian:
        the strongly connected components of its graph in topological order
        (every entry below the diagonal blocks), and inside a component the
        unknowns that other components read last -- so that the incomplete
        factorisation drops no fill outside the diagonal blocks.  For the DC
        Jacobian of a feed-forward logic netlist the blocks are a cell's
        output net (driver, far end, stack nodes) and ILU(0) in this order is
        nearly the exact LU (the topological order sparse direct solvers
        reach by their block-triangular form)."""
        import scipy.sparse as sps
        from scipy.sparse.csgraph import connected_components
        n = self.n
        nz = vals != 0
        r, c = self._rows[nz], self._cols[nz]
        g = sps.csr_matrix((np.ones(len(r)), (r, c)), shape=(n, n))
        ncomp, lab = connected_components(g, directed=True, connection="strong")
        cr, cc = lab[r], lab[c]
        m = cr != cc
        E = sps.csr_matrix((np.ones(int(m.sum())), (cc[m], cr[m])), shape=(ncomp, ncomp))
        E.sum_duplicates()
        E.data[:] = 1.0
        indeg = np.asarray(E.sum(axis=0)).ravel().astype(np.int64)
        level = np.full(ncomp, -1, np.int64)
        frontier = np.nonzero(indeg == 0)[0]
        depth = 0
        while len(frontier):
            level[frontier] = depth
            succ = E[frontier].indices
            np.add.at(indeg, succ, -1)
            cand = np.unique(succ)
            frontier = cand[indeg[cand] == 0]
            depth += 1
        if (level < 0).any():
            raise RuntimeError("component graph is not acyclic")
        ext = np.zeros(n, np.int64)
        np.add.at(ext, c[m], 1)
        self.order_info = {"components": int(ncomp), "levels": int(depth),
                           "largest_component": int(np.bincount(lab).max())}
        return np.lexsort((np.arange(n), ext > 0, lab, level[lab])).astype(np.int64)

    def _build(self, q):
        """The solved system in order q (position p holds unknown q[p] and
        the row paired with it): its CSR with every diagonal, the map from
        the circuit's CSR slots and the ILU(0)."""
        torch = self.torch
        n = self.n
        qinv = np.empty(n, np.int64)
        qinv[q] = np.arange(n)
        self.row_of = torch.from_numpy(self._src[q]).cuda()
        self.q = torch.from_numpy(q).cuda()
        ckeys = qinv[self._rows] * n + qinv[self._cols]
        keys = np.unique(np.concatenate([ckeys, np.arange(n, dtype=np.int64) * (n + 1)]))
        jrp = np.zeros(n + 1, np.int64)
        np.add.at(jrp, keys // n + 1, 1)
        jrp = np.cumsum(jrp).astype(np.int64)
        jr, jc = keys // n, keys % n
        self.J = CSRMatrix(self.ctx, jrp.astype(np.int32), jc.astype(np.int32))
        self.map = torch.from_numpy(np.searchsorted(keys, ckeys)).cuda()
        # terminal rows: an unknown that no other row reads (a voltage
        # source's branch current, whose row is its node's KCL -- for a supply
        # one entry per transistor on the rail).  The incomplete factorisation
        # keeps only their diagonal; the preconditioner then solves them
        # exactly after the rest (block forward substitution), so the long
        # rows never enter the factorisation.
        readers = np.bincount(jc[jr != jc], minlength=n)
        term = np.nonzero(readers == 0)[0]
        has_off = np.zeros(n, bool)
        has_off[jr[jr != jc]] = True
        term = term[has_off[term]]
        keep = ~(np.isin(jr, term) & (jr != jc))
        krp = np.zeros(n + 1, np.int64)
        np.add.at(krp, jr[keep] + 1, 1)
        self.M = CSRMatrix(self.ctx, np.cumsum(krp).astype(np.int32), jc[keep].astype(np.int32))
        self.m_src = torch.from_numpy(np.nonzero(keep)[0]).cuda()
        self.term = torch.from_numpy(term).cuda()
        if len(term):
            sel = np.isin(jr, term) & (jr != jc)
            pos = np.searchsorted(term, jr[sel])
            trp = np.zeros(len(term) + 1, np.int64)
            np.add.at(trp, pos + 1, 1)
            self.T = CSRMatrix(self.ctx, np.cumsum(trp).astype(np.int32), jc[sel].astype(np.int32))
            self.t_src = torch.from_numpy(np.nonzero(sel)[0]).cuda()
            self.t_diag = torch.from_numpy(jrp[term] + np.array(
                [np.searchsorted(jc[jrp[t]:jrp[t + 1]], t) for t in term], np.int64)).cuda()
        self.ilu = ILU0(self.M)
        self.n_terminal = len(term)

    def _precond(self, b):
        y = self.ilu.solve(b)
        if self.n_terminal:
            t = self.T.matvec(y)
            y[self.term] = (b[self.term] - t) / self.J.values[self.t_diag]
        return y

    def load(self, x, flags):
        torch = self.torch
        v = self.v
        v["nextSol"][:self.n].copy_(x)
        for k in self.outs:
            v[k].zero_()
        ss = self._ss(gainScale=1.0, nltermScale=1.0, **flags)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        self.c.load(ss, **v)
        torch.cuda.synchronize()
        self.stats["load_s"] += time.perf_counter() - t0
        self.stats["loads"] += 1

    def accept(self):
        self.v["currSta"].copy_(self.v["nextSta"])
        self.v["currSto"].copy_(self.v["nextSto"])

    def _newton(self, x, flags_of, system, max_iter, refactor_each=True):
        """refactor_each False: the ILU(0) of the first iteration
        preconditions every iteration of this solve (GMRES still solves the
        current Jacobian exactly)."""
        torch = self.torch
        v = self.v
        for it in range(max_iter):
            self.load(x, flags_of(it))
            rhs, a, b = system()
            if self.J is None:
                k = len(self._rows)
                self._build(self._topo_order(
                    (a * v["dFdx"][:k] + b * v["dQdx"][:k]).cpu().numpy()))
            rhs = rhs[self.row_of]
            t0 = time.perf_counter()
            Jv = self.J.values
            Jv.zero_()
            Jv[self.map] = a * v["dFdx"][:len(self.map)] + b * v["dQdx"][:len(self.map)]
            self.J.sync()
            if self.n_terminal:
                self.T.values[:self.T.nnz] = Jv[self.t_src]
                self.T.sync()
            if refactor_each or it == 0:
                self.M.values[:self.M.nnz] = Jv[self.m_src]
                zp = self.ilu.factor()
                if zp >= 0:
                    raise NewtonFailure(f"ILU(0): zero pivot in row {zp}")
                self.stats["factors"] = self.stats.get("factors", 0) + 1
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            # the junction-initialisation iteration's step is discarded by
            # the limiter anyway: a looser solve of it is enough
            tol = self.krylov_tol if not flags_of(it).get("initJctFlag") else 1e-6
            dy, rel, its = fgmres(self.J.matvec, self._precond, rhs, tol,
                                  self.restart, self.max_krylov)
            dx = torch.empty_like(dy)
            dx[self.q] = dy
            torch.cuda.synchronize()
            self.stats["factor_s"] += t1 - t0
            self.stats["krylov_s"] += time.perf_counter() - t1
            self.stats["krylov"] += its
            self.stats["newton"] += 1
            x_load = x
            x = x + dx
            dmax = float(torch.max(torch.abs(dx[:self.nv])))
            if self.verbose:
                print(f"    Newton {it}: |dx| {dmax:.3e}, GMRES {its} (rel {rel:.1e})", file=sys.stderr, flush=True)
            if rel > 1e3 * tol:
                raise NewtonFailure(f"GMRES stalled at relative residual {rel:.2e}")
            if not np.isfinite(dmax):
                raise NewtonFailure("Newton update is not finite")
            if it > 0 and dmax <= self.vtol:
                return x, x_load, it + 1
        raise NewtonFailure(f"Newton did not converge in {max_iter} iterations")

    def dcop(self, x0, max_iter=100, times=None):
        """DC operating point from x0 (a device tensor of n unknowns)."""
        v = self.v
        extra = dict(times or {})

        def flags(it):
            return dict(newtonIter=it, dcopFlag=1, tranopFlag=1,
                        initJctFlag=1 if it == 0 else 0, currTime=0.0, **extra)

        def system():
            n = self.n
            return -v["f"][:n] + v["dFdxdVp"][:n], 1.0, 0.0

        x, _, its = self._newton(x0, flags, system, max_iter)
        self.accept()
        return x, its

    def run(self, x0, t_stop, h, method="trap", record=None, max_iter=50, progress=None):
        """DC operating point, then fixed steps of h to t_stop.  record:
        unknown indices whose values are kept at every step.  Returns
        (times, recorded[steps+1, len(record)], per-step Newton counts)."""
        if method not in ("be", "trap"):
            raise ValueError("method must be 'be' or 'trap'")
        torch = self.torch
        v = self.v
        n = self.n
        times = dict(startingTimeStep=h, finalTime=t_stop, bpTol=1e-18)
        rec = torch.as_tensor(np.asarray(record if record is not None else [], np.int64)).cuda()
        x, its = self.dcop(x0, times=times)
        newton = [its]
        q_n = v["q"][:n].clone()
        f_n = v["f"][:n].clone()
        out_t = [0.0]
        out_v = [x[rec].cpu().numpy()]
        steps = int(round(t_stop / h))
        for k in range(steps):
            first = k == 0
            tk = (k + 1) * h

            def flags(it, first=first, tk=tk):
                return dict(newtonIter=it, transientFlag=1, pdt=1.0 / h, currTimeStep=h,
                            lastTimeStep=h, currTime=tk, initTranFlag=1 if first else 0,
                            beginIntegrationFlag=1 if first else 0, **times)

            trap = method == "trap" and k >= 2

            def system(q_n=q_n, f_n=f_n, trap=trap):
                f, q = v["f"][:n], v["q"][:n]
                if trap:
                    rhs = (-((q - q_n) / h + 0.5 * f + 0.5 * f_n) + 0.5 * v["dFdxdVp"][:n]
                           + v["dQdxdVp"][:n] / h)
                    return rhs, 0.5, 1.0 / h
                rhs = -((q - q_n) / h + f) + v["dFdxdVp"][:n] + v["dQdxdVp"][:n] / h
                return rhs, 1.0, 1.0 / h

            x, _, its = self._newton(x, flags, system, max_iter,
                                     refactor_each=self.refactor_each_step_iter)
            self.accept()
            newton.append(its)
            q_n = v["q"][:n].clone()
            f_n = v["f"][:n].clone()
            out_t.append(tk)
            out_v.append(x[rec].cpu().numpy())
            if progress and (k + 1) % progress == 0:
                print(f"  step {k + 1}/{steps}: t = {tk:.3e} s, Newton {its}, "
                      f"GMRES total {self.stats['krylov']}", file=sys.stderr, flush=True)
        return np.asarray(out_t), np.asarray(out_v), np.asarray(newton)
