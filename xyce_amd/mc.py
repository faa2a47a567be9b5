"""Batched Monte Carlo ring-oscillator transient on one GPU.

Thousands of ring oscillators (odd stage counts, per-ring VTH0 and TOXE
samples) run as one circuit batch: every ring's BSIM4 instances are one
group, loaded together by the device loads, and every ring's Newton system
is formed and solved on the device by xa_batch_newton_step (one wavefront
per ring, banded LU in the ring's reverse Cuthill-McKee order).  The host
only launches and checks convergence, once per Newton iteration.  Each ring is the circuit of xyce_amd.ring.Ring with its sample's
TOXE and |VTH0| scale, so the plain drop-in (that module, one ring at a
time) runs exactly the same circuits.

Variation (one sample per ring): |VTH0| ~ N(1, 0.03) x 0.7 V applied as
every transistor's DELVTO; TOXE ~ N(1, 0.02) x 3 nm on the ring's NMOS and
PMOS cards.  Stage counts are uniform over the odd numbers 11..101.

Integration is the ring module's: DC operating point with node 0 of each
ring held at its initial condition, then fixed-step backward Euler, Newton
until the largest update of every ring is below vtol.
"""
import ctypes as C
import time

import numpy as np

from . import Context, Master, SolverState, make_vectors
from ._abi import XA_STEP_BE, XA_STEP_DCOP, XA_STEP_TRAP, check, lib
from .ring import Ring

SEED = 20261203


class RingBatch:
    """The sampled rings and their place in the batched system."""

    def __init__(self, n_rings=4096, seed=SEED, stages=None, sigma_vth=0.03, sigma_tox=0.02,
                 toxe=3.0e-9, **ring_kw):
        rng = np.random.default_rng(np.random.SeedSequence(seed))
        if stages is None:
            stages = rng.choice(np.arange(11, 102, 2), size=n_rings)
        self.stages = np.asarray(stages, np.int64)
        n_rings = len(self.stages)
        self.vth_scale = 1.0 + sigma_vth * rng.standard_normal(n_rings)
        self.toxe = toxe * (1.0 + sigma_tox * rng.standard_normal(n_rings))
        self.ring_kw = ring_kw
        self.n_rings = n_rings

    def ring(self, r):
        """Ring r as a stand-alone circuit (the plain drop-in's netlist)."""
        return Ring(stages=int(self.stages[r]), toxe=float(self.toxe[r]),
                    vth_scale=float(self.vth_scale[r]), **self.ring_kw)


class BatchTopology:
    """All rings in one BSIM4 group: ring r's nodes are [base_r, base_r +
    N_r] (outputs, then its supply); two cards per ring."""

    def __init__(self, master, batch):
        rings = [batch.ring(r) for r in range(batch.n_rings)]
        self.rings = rings
        for ring in rings:
            for level, card in ring.models:
                master.add_model(card, level=level)
        mi, cols, lids = [], {}, []
        base = 0
        self.base = np.zeros(batch.n_rings + 1, np.int64)
        for r, ring in enumerate(rings):
            self.base[r] = base
            mi.append(ring.model_index + 2 * r)
            for k, v in ring.params.items():
                cols.setdefault(k, []).append(v)
            for i in range(ring.n_inst):
                lids.extend(base + t if t >= 0 else -1 for t in ring.terminals(i))
            base += ring.N + 1
        self.base[-1] = base
        n_inst = sum(ring.n_inst for ring in rings)
        params = {k: np.concatenate(v) for k, v in cols.items()}
        if "DELVTO" not in params and any("DELVTO" in ring.params for ring in rings):
            raise ValueError("DELVTO must be given for every ring or none")
        master.add_instances(n_inst, np.concatenate(mi).astype(np.int32), params)
        master.setup()
        n_ext, n_int, n_sta, n_sto = master.counts()
        if n_int.any():
            raise ValueError("the batched rings assume cards without internal nodes")
        self.n_nodes = base
        self.n_sta = int(n_sta.sum())
        self.n_sto = int(n_sto.sum())
        master.register_lids(np.asarray(lids, np.int32), np.arange(self.n_sta, dtype=np.int32),
                             np.arange(self.n_sto, dtype=np.int32))
        r_, c_ = master.jac_pairs()
        outs = np.concatenate([np.arange(self.base[r], self.base[r] + ring.N)
                               for r, ring in enumerate(rings)])
        r_ = np.concatenate([r_, outs]).astype(np.int64)
        c_ = np.concatenate([c_, outs]).astype(np.int64)
        key = np.unique(r_ * self.n_nodes + c_)
        rows = key // self.n_nodes
        rp = np.zeros(self.n_nodes + 1, np.int64)
        np.add.at(rp, rows + 1, 1)
        self.row_ptr = np.cumsum(rp).astype(np.int32)
        self.col_idx = (key % self.n_nodes).astype(np.int32)
        self.nnz = len(key)
        master.register_csr(self.row_ptr, self.col_idx)
        master.finalize()
        self.max_unknowns = int(np.max(np.diff(self.base)))
        self._order(rings)

    def _order(self, rings):
        """Elimination order of each ring: reverse Cuthill-McKee over the
        couplings of its free nodes (the supply is held), which puts a
        ring's cycle in two bands either side of the diagonal."""
        from scipy.sparse import csr_matrix
        from scipy.sparse.csgraph import reverse_cuthill_mckee
        self.pos = np.zeros(self.n_nodes, np.int32)
        kl = ku = 0
        rows = np.repeat(np.arange(self.n_nodes), np.diff(self.row_ptr))
        for r, ring in enumerate(rings):
            b0, n = int(self.base[r]), ring.N + 1
            sel = (rows >= b0) & (rows < b0 + ring.N) & (self.col_idx >= b0) & \
                  (self.col_idx < b0 + ring.N)
            i, j = rows[sel] - b0, self.col_idx[sel] - b0
            g = csr_matrix((np.ones(len(i)), (i, j)), shape=(n, n))
            perm = reverse_cuthill_mckee(g + g.T, symmetric_mode=True)
            p = np.empty(n, np.int32)
            p[perm] = np.arange(n, dtype=np.int32)
            self.pos[b0:b0 + n] = p
            d = p[i] - p[j]
            kl = max(kl, int(d.max(initial=0)))
            ku = max(ku, int((-d).max(initial=0)))
        self.kl, self.ku = kl, ku


class BatchTransient:
    """The whole batch on one GPU."""

    def __init__(self, batch, device=0, options=None, use_graph=False, persistent=False):
        import torch
        self.use_graph = use_graph
        self.graphs = {}
        self.torch = torch
        self.batch = batch
        self.ctx = Context(device)
        self.master = Master(self.ctx, "bsim4", options)
        t0 = time.perf_counter()
        self.topo = t = BatchTopology(self.master, batch)
        self.setup_seconds = time.perf_counter() - t0
        if persistent:   # one resident grid loops over every ring's instances
            self.master.set_persistent(-1)
        dev = torch.device("cuda", device)
        self.dev = dev
        z = lambda m: torch.zeros(max(m, 1), dtype=torch.float64, device=dev)
        self.d = {"nextSol": z(t.n_nodes), "currSol": z(t.n_nodes), "nextSta": z(t.n_sta),
                  "currSta": z(t.n_sta), "lastSta": z(t.n_sta), "nextStaDeriv": z(t.n_sta),
                  "nextSto": z(t.n_sto), "currSto": z(t.n_sto), "b": z(t.n_nodes),
                  "storeLeadF": z(t.n_sto), "storeLeadQ": z(t.n_sto)}
        # the outputs zeroed before every load share one buffer: one fill
        sizes = {"f": t.n_nodes, "q": t.n_nodes, "dFdxdVp": t.n_nodes, "dQdxdVp": t.n_nodes,
                 "dFdx": t.nnz, "dQdx": t.nnz}
        self.outputs = z(sum(max(m, 1) for m in sizes.values()))
        off = 0
        for k, m in sizes.items():
            self.d[k] = self.outputs[off:off + max(m, 1)]
            off += max(m, 1)
        self.vectors = make_vectors(**self.d)
        c_lin = np.zeros(t.n_nodes)
        for r, ring in enumerate(t.rings):
            c_lin[t.base[r]:t.base[r] + ring.N] = ring.c_load
        i32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int32)).to(dev)
        self.c_start = i32(t.base)
        self.pos = i32(t.pos)
        self.row_ptr = i32(t.row_ptr)
        self.col_idx = i32(t.col_idx)
        self.c_lin = torch.from_numpy(c_lin).to(dev)
        self.q_prev = z(t.n_nodes)
        self.f_prev = z(t.n_nodes)
        self.q_now = z(t.n_nodes)
        self.dx_max = z(batch.n_rings)
        supply = t.base[1:] - 1
        first = t.base[:-1]
        held_dc = np.zeros(t.n_nodes, np.uint8)
        held_dc[supply] = 1
        held_dc[first] = 1
        held_tr = np.zeros(t.n_nodes, np.uint8)
        held_tr[supply] = 1
        self.held_dc = torch.from_numpy(held_dc).to(dev)
        self.held_tr = torch.from_numpy(held_tr).to(dev)
        x = np.zeros(t.n_nodes)
        for r, ring in enumerate(t.rings):
            k = np.arange(ring.N)
            x[t.base[r]:t.base[r] + ring.N] = np.where((k - ring.ic_node) % 2 == 0, ring.ic_value,
                                                       ring.vdd - ring.ic_value)
            x[t.base[r] + ring.N] = ring.vdd
        self.x = self.d["nextSol"]            # the loads read the iterate in place
        self.x.copy_(torch.from_numpy(x))
        self.loads = 0
        self.newton_iterations = 0

    def _ptr(self, t):
        return C.c_void_p(t.data_ptr())

    def _load(self, **flags):
        ss = SolverState()
        ss.gainScale = 1.0
        ss.nltermScale = 1.0
        for k, v in flags.items():
            setattr(ss, k, v)
        self.outputs.zero_()
        self.master.loadAll(ss, self.vectors)   # the four calls' results, fewer launches
        self.loads += 1

    def _solve(self, h, method):
        d, t = self.d, self.topo
        P = self._ptr
        dc = method == XA_STEP_DCOP
        check(lib().xa_batch_newton_step_ex(
            self.ctx.h, self.batch.n_rings, t.max_unknowns, t.kl, t.ku, P(self.c_start),
            P(self.pos), P(self.row_ptr), P(self.col_idx), P(d["dFdx"]), P(d["dQdx"]), P(d["f"]),
            P(d["q"]), P(d["dFdxdVp"]), P(d["dQdxdVp"]), P(self.q_prev), P(self.f_prev),
            P(self.c_lin), P(self.held_dc if dc else self.held_tr), C.c_double(h), method,
            P(self.x), P(self.q_now), P(self.dx_max)), "xa_batch_newton_step_ex")

    def _iteration(self, flags, h, method):
        """One Newton iteration's device work: zero, load, solve, reduce."""
        self._load(**flags)
        self._solve(h, method)
        return self.dx_max.max()

    def _graph(self, key, flags, h, method):
        """The iteration captured as a HIP graph.  The solver-state flags are
        baked into the captured kernel arguments, so one graph serves every
        iteration with the same flags; the loads test newtonIter only for
        == 0, so one graph stands for every later iteration."""
        torch = self.torch
        g = self.graphs.get(key)
        if g is None:
            g = torch.cuda.CUDAGraph()
            main = torch.cuda.current_stream()
            with torch.cuda.graph(g):
                self.ctx.set_stream(torch.cuda.current_stream())
                try:
                    out = self._iteration(flags, h, method)
                finally:
                    self.ctx.set_stream(main)
            self.graphs[key] = g = (g, out)
        return g

    def _newton(self, flags_of, h, method, vtol, max_iter):
        for it in range(max_iter):
            flags = flags_of(it)
            if self.use_graph:
                key = (method, h, it == 0, tuple(sorted((k, v) for k, v in flags.items()
                                                        if k != "newtonIter")))
                g, worst_t = self._graph(key, dict(flags, newtonIter=min(it, 1)), h, method)
                g.replay()
                self.loads += 1
            else:
                worst_t = self._iteration(flags, h, method)
            self.newton_iterations += 1
            worst = float(worst_t.item())
            if not np.isfinite(worst):
                raise RuntimeError("batched Newton: a circuit's update is not finite")
            if it > 0 and worst <= vtol:
                return it + 1
        raise RuntimeError(f"batched Newton did not converge in {max_iter} iterations")

    def _accept(self):
        d = self.d
        d["currSto"].copy_(d["nextSto"])
        d["currSta"].copy_(d["nextSta"])
        self.q_prev.copy_(self.q_now)
        self.f_prev.copy_(d["f"])   # the last load's F: the trapezoid's history

    def run(self, t_stop=2e-9, h=1e-12, vtol=1e-9, max_iter=50, method="be"):
        """DC operating point then fixed steps of backward Euler (method "be")
        or of the one-step integrator's trapezoid rule ("trap": two backward
        Euler steps, then order 2, as N_TIA_OneStep.C:2130,2211 raise the
        order after the second step); returns (times, V0) with V0[step, r]
        the voltage of ring r's node 0."""
        if method not in ("be", "trap"):
            raise ValueError("method must be 'be' or 'trap'")
        torch = self.torch
        t = self.topo
        first = torch.from_numpy(t.base[:-1].astype(np.int64)).to(self.dev)
        # the transient's operating point (tranopFlag: the loads compute the
        # charges the first step differences against, N_ANP_AnalysisManager.C:
        # 2749-2753)
        self._newton(lambda it: dict(newtonIter=it, dcopFlag=1, tranopFlag=1,
                                     initJctFlag=1 if it == 0 else 0),
                     1.0, XA_STEP_DCOP, vtol, max_iter)
        self._accept()
        steps = int(round(t_stop / h))
        V0 = torch.zeros((steps + 1, self.batch.n_rings), dtype=torch.float64, device=self.dev)
        V0[0] = self.x[first]
        for n in range(steps):
            f0 = n == 0

            def flags(it, f0=f0):
                return dict(newtonIter=it, transientFlag=1, pdt=1.0 / h, currTimeStep=h,
                            lastTimeStep=h, initTranFlag=1 if f0 else 0,
                            beginIntegrationFlag=1 if f0 else 0)
            order2 = method == "trap" and n >= 2
            self._newton(flags, h, XA_STEP_TRAP if order2 else XA_STEP_BE, vtol, max_iter)
            self._accept()
            V0[n + 1] = self.x[first]
        torch.cuda.synchronize()
        return np.arange(steps + 1) * h, V0.cpu().numpy()
