"""CMOS ring oscillator transient through the device loads.

This is the one-circuit ("plain") form of the ring-oscillator workload: N
inverter stages of BSIM4 level-54 devices (default NMOS/PMOS cards), a load
capacitor from every stage output to ground, a fixed supply node, and an
initial condition on the first stage output that starts the oscillation.

* Stage k drives node k from node (k-1) mod N.  NMOS: D = k, G = k-1,
  S = B = ground; PMOS: D = k, G = k-1, S = B = supply.
* The device loads are the reference's four Master calls
  (N_DEV_DeviceMaster.h:643-725) on any loader object with the interface of
  :class:`GpuLoader`; the capacitors and the supply are linear and are
  assembled here.
* DC operating point: Newton on F(x) = 0 with the supply and the initial-
  condition node held (initJct on the first iteration and tranopFlag set, as
  the reference's operating point of a transient); the solution starts the
  transient with the condition released.
* Transient: fixed-step backward Euler.  The Newton residual and right-hand
  side are formed the way the reference's BDF integrator forms them
  (N_TIA_BackwardDifferentiation15.C:215-265): RHS = -((Q(x) - q_n)/h + F(x))
  + dFdxdVp + dQdxdVp/h, Jacobian dF/dx + dQ/dx/h, with Q and dQ/dx including
  the load capacitors.
* Newton stops when the largest update is below `vtol` volts.

The dense linear solve is numpy's (LAPACK); it is the same for every loader,
so two runs differ only through their device loads.
"""
import time

import numpy as np

# Default ring: 51 stages, L = 45 nm, Wn = 90 nm, Wp = 180 nm, 1 fF per stage,
# VDD = 1 V, 1 ps steps to 2 ns.
DEFAULT = dict(stages=51, vdd=1.0, c_load=1e-15, l=45e-9, wn=90e-9, wp=180e-9)


class Ring:
    """Netlist of the ring: models, instance parameters and terminal nodes."""

    def __init__(self, stages=51, vdd=1.0, c_load=1e-15, l=45e-9, wn=90e-9, wp=180e-9,
                 ic_node=0, ic_value=0.0, toxe=None, vth_scale=1.0):
        """toxe: oxide thickness of both cards (default: the card default);
        vth_scale: |VTH0| scaled by this factor through each transistor's
        DELVTO (VTH0 defaults +-0.7 V)."""
        if stages < 3 or stages % 2 == 0:
            raise ValueError("a ring oscillator needs an odd number of stages >= 3")
        self.N = stages
        self.vdd = vdd
        self.c_load = c_load
        self.ic_node = ic_node
        self.ic_value = ic_value
        self.supply = stages                     # node index of the supply
        cards = [{"TYPE": 1.0}, {"TYPE": -1.0}]
        if toxe is not None:
            for c in cards:
                c["TOXE"] = toxe
        self.models = [(54, c) for c in cards]
        n = 2 * stages
        self.n_inst = n
        self.model_index = np.tile(np.array([0, 1], np.int32), stages)
        self.params = {"W": np.tile(np.array([wn, wp]), stages), "L": np.full(n, l)}
        if vth_scale != 1.0:
            self.params["DELVTO"] = np.tile(np.array([0.7, -0.7]) * (vth_scale - 1.0), stages)

    def terminals(self, inst):
        """External nodes (D, G, S, B) of instance `inst`; ground is -1."""
        k, pmos = divmod(inst, 2)
        d, g = k, (k - 1) % self.N
        if pmos:
            return [d, g, self.supply, self.supply]
        return [d, g, -1, -1]

    def netlist(self, t_step=1e-12, t_stop=2e-9):
        """The ring as SPICE text for the netlist reader: stage outputs
        S0..S{N-1}, the supply VDD, the load capacitors, the supply source,
        the initial condition and the transient.  Values print exactly (a
        bare number with no scale suffix reads back bit for bit)."""
        cards = [".model nch nmos " + " ".join(f"{k}={v!r}" for k, v in
                                                 [("level", 54)] + list(self.models[0][1].items())
                                                 if k != "TYPE"),
                 ".model pch pmos " + " ".join(f"{k}={v!r}" for k, v in
                                                 [("level", 54)] + list(self.models[1][1].items())
                                                 if k != "TYPE")]
        name = lambda x: "0" if x < 0 else ("VDD" if x == self.supply else f"S{x}")
        lines = [f"CMOS ring oscillator, {self.N} stages"] + cards
        for i in range(self.n_inst):
            nodes = " ".join(name(x) for x in self.terminals(i))
            card = "pch" if self.model_index[i] else "nch"
            extra = " ".join(f"{k}={float(v[i])!r}" for k, v in self.params.items())
            lines.append(f"M{i} {nodes} {card} {extra}")
        for k in range(self.N):
            lines.append(f"C{k} S{k} 0 {float(self.c_load)!r}")
        lines.append(f"VDD VDD 0 DC {float(self.vdd)!r}")
        lines.append(f".ic V(S{self.ic_node})={float(self.ic_value)!r}")
        lines.append(f".tran {float(t_step)!r} {float(t_stop)!r}")
        lines.append(".end")
        return "\n".join(lines) + "\n"


class Topology:
    """LIDs and CSR structure of a ring built into one BSIM4 group."""

    def __init__(self, master, ring):
        for level, card in ring.models:
            master.add_model(card, level=level)
        master.add_instances(ring.n_inst, ring.model_index, ring.params)
        master.setup()
        n_ext, n_int, n_sta, n_sto = master.counts()
        lids = []
        nxt = ring.N + 1                         # internal nodes follow the supply
        for i in range(ring.n_inst):
            lids.extend(ring.terminals(i)[:n_ext[i]])
            lids.extend(range(nxt, nxt + n_int[i]))
            nxt += int(n_int[i])
        self.n_nodes = nxt
        self.n_sta = int(n_sta.sum())
        self.n_sto = int(n_sto.sum())
        master.register_lids(np.asarray(lids, np.int32), np.arange(self.n_sta, dtype=np.int32),
                             np.arange(self.n_sto, dtype=np.int32))
        r, c = master.jac_pairs()
        # the load capacitors add the output diagonals
        r = np.concatenate([r, np.arange(ring.N, dtype=np.int32)])
        c = np.concatenate([c, np.arange(ring.N, dtype=np.int32)])
        key = np.unique(r.astype(np.int64) * self.n_nodes + c)
        self.rows = (key // self.n_nodes).astype(np.int64)
        self.cols = (key % self.n_nodes).astype(np.int64)
        rp = np.zeros(self.n_nodes + 1, np.int64)
        np.add.at(rp, self.rows + 1, 1)
        self.row_ptr = np.cumsum(rp).astype(np.int32)
        self.col_idx = self.cols.astype(np.int32)
        self.nnz = len(key)
        master.register_csr(self.row_ptr, self.col_idx)
        master.finalize()

    def dense(self, vals):
        a = np.zeros((self.n_nodes, self.n_nodes))
        a[self.rows, self.cols] = vals
        return a


class GpuLoader:
    """The ring's BSIM4 group on one GPU: one load = the four Master calls on
    device vectors (or, fused, the one-call xa_load_all with the same
    results), outputs copied back to the host."""

    def __init__(self, ring, device=0, options=None, fused=False, graph=False):
        """graph: each load (output zeroing and the device calls) is captured
        once per set of solver flags as a HIP graph and replayed, which
        removes the per-kernel launch cost of this launch-bound size; the
        flags are baked into the captured arguments, and newtonIter is only
        tested for == 0 by the loads, so one graph serves every later
        iteration of a flag set."""
        self.fused = fused
        self.graph = graph
        self.graphs = {}
        import torch
        from . import Context, Master, SolverState, make_vectors
        self.torch = torch
        self._ss_cls = SolverState
        self.ctx = Context(device)
        self.master = Master(self.ctx, "bsim4", options)
        self.topo = Topology(self.master, ring)
        t = self.topo
        dev = torch.device("cuda", device)
        z = lambda m: torch.zeros(max(m, 1), dtype=torch.float64, device=dev)
        self.d = {"nextSol": z(t.n_nodes), "currSol": z(t.n_nodes), "nextSta": z(t.n_sta),
                  "currSta": z(t.n_sta), "lastSta": z(t.n_sta), "nextStaDeriv": z(t.n_sta),
                  "nextSto": z(t.n_sto), "currSto": z(t.n_sto), "b": z(t.n_nodes),
                  "storeLeadF": z(t.n_sto), "storeLeadQ": z(t.n_sto)}
        # the outputs zeroed before every load share one buffer: one fill
        self.sizes = sizes = {"f": t.n_nodes, "q": t.n_nodes, "dFdxdVp": t.n_nodes,
                              "dQdxdVp": t.n_nodes, "dFdx": t.nnz, "dQdx": t.nnz}
        self.outputs = z(sum(max(m, 1) for m in sizes.values()))
        off = 0
        for k, m in sizes.items():
            self.d[k] = self.outputs[off:off + max(m, 1)]
            off += max(m, 1)
        self.vectors = make_vectors(**self.d)
        self.load_seconds = 0.0
        self.loads = 0

    def load(self, x, **flags):
        torch = self.torch
        ss = self._ss_cls()
        ss.gainScale = 1.0
        ss.nltermScale = 1.0
        for k, v in flags.items():
            setattr(ss, k, v)
        d = self.d
        d["nextSol"].copy_(torch.from_numpy(np.ascontiguousarray(x)))
        if self.graph:
            key = tuple(sorted((k, v) for k, v in flags.items() if k != "newtonIter")) + \
                (flags.get("newtonIter", 0) == 0,)
            g = self.graphs.get(key)
            if g is None:
                g = torch.cuda.CUDAGraph()
                main = torch.cuda.current_stream()
                with torch.cuda.graph(g):
                    self.ctx.set_stream(torch.cuda.current_stream())
                    try:
                        self._calls(ss)
                    finally:
                        self.ctx.set_stream(main)
                self.graphs[key] = g
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            g.replay()
        else:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            self._calls(ss)
        torch.cuda.synchronize()
        self.load_seconds += time.perf_counter() - t0
        self.loads += 1
        # one device-to-host copy of the output buffer, cut into the vectors
        host = self.outputs.cpu().numpy()
        out, off = {}, 0
        for k, m in self.sizes.items():
            out[k] = host[off:off + m]
            off += max(m, 1)
        return out

    def _calls(self, ss):
        """Zero the outputs and load (the four Master calls, or xa_load_all)."""
        self.outputs.zero_()
        m = self.master
        if self.fused:
            m.loadAll(ss, self.vectors)
        else:
            m.updateState(ss, self.vectors)
            m.updateSecondaryState(ss, self.vectors)
            m.loadDAEVectors(ss, self.vectors)
            m.loadDAEMatrices(ss, self.vectors)

    def accept(self):
        """End of a time step: the converged store/state become the previous
        step's (currSto/currSta)."""
        self.d["currSto"].copy_(self.d["nextSto"])
        self.d["currSta"].copy_(self.d["nextSta"])


class NewtonFailure(RuntimeError):
    pass


def _newton(loader, ring, x, fixed, flags_of, residual, vtol, max_iter):
    """Newton on the unknowns not in `fixed`; residual(load) -> (rhs, J)."""
    t = loader.topo
    free = np.setdiff1d(np.arange(t.n_nodes), np.asarray(sorted(fixed), np.int64))
    for it in range(max_iter):
        out = loader.load(x, **flags_of(it))
        rhs, jac = residual(out, x)
        dx = np.linalg.solve(jac[np.ix_(free, free)], rhs[free])
        x_load = x
        x = x.copy()
        x[free] += dx
        if it > 0 and np.max(np.abs(dx)) <= vtol:
            return x, out, x_load, it + 1
    raise NewtonFailure(f"Newton did not converge in {max_iter} iterations")


def transient(loader, ring, t_stop=2e-9, h=1e-12, vtol=1e-9, max_iter=50, record=None,
              method="be"):
    """DC operating point then fixed steps to t_stop: backward Euler (method
    "be") or the one-step integrator's trapezoid rule ("trap": two backward
    Euler steps, then order 2 from the third step on, as N_TIA_OneStep.C:
    2130,2211 raise the order; residual and Jacobian of N_TIA_OneStep.C:
    272-335,463-500, with the previous step's F as qHistory[2]).  Returns
    (times, V[steps+1, n_nodes], stats)."""
    if method not in ("be", "trap"):
        raise ValueError("method must be 'be' or 'trap'")
    t = loader.topo
    N = ring.N
    C = np.zeros(t.n_nodes)
    C[:N] = ring.c_load
    # initial guess: the alternating pattern the initial condition implies
    x = np.zeros(t.n_nodes)
    x[:N] = np.where((np.arange(N) - ring.ic_node) % 2 == 0, ring.ic_value,
                     ring.vdd - ring.ic_value)
    x[ring.supply] = ring.vdd
    x[ring.ic_node] = ring.ic_value

    def dc_residual(out, xv):
        return -out["f"] + out["dFdxdVp"], t.dense(out["dFdx"])

    def dc_flags(it):
        # the operating point of a transient: tranopFlag set, so the loads
        # compute the charges the first step differences against
        # (N_ANP_AnalysisManager.C:2749-2753)
        return dict(newtonIter=it, dcopFlag=1, tranopFlag=1, initJctFlag=1 if it == 0 else 0)

    newton_total = 0
    x, out, x_load, its = _newton(loader, ring, x, {ring.supply, ring.ic_node}, dc_flags,
                                  dc_residual, vtol, max_iter)
    newton_total += its
    loader.accept()
    # the charge history is the last load's: device Q and capacitor charge at
    # the voltages that load saw
    q_n = out["q"] + C * x_load
    f_n = out["f"].copy()
    times = [0.0]
    V = [x.copy()]
    steps = int(round(t_stop / h))
    for n in range(steps):
        x_n = x.copy()
        first = n == 0

        def tr_flags(it, first=first):
            return dict(newtonIter=it, transientFlag=1, pdt=1.0 / h, currTimeStep=h,
                        lastTimeStep=h, initTranFlag=1 if first else 0,
                        beginIntegrationFlag=1 if first else 0)

        def tr_residual(out, xv, q_n=q_n):
            q = out["q"] + C * xv
            rhs = -((q - q_n) / h + out["f"]) + out["dFdxdVp"] + out["dQdxdVp"] / h
            jac = t.dense(out["dFdx"]) + t.dense(out["dQdx"]) / h + np.diag(C / h)
            return rhs, jac

        def trap_residual(out, xv, q_n=q_n, f_n=f_n):
            q = out["q"] + C * xv
            rhs = (-((q - q_n) / h + 0.5 * out["f"] + 0.5 * f_n) + 0.5 * out["dFdxdVp"]
                   + out["dQdxdVp"] / h)
            jac = 0.5 * t.dense(out["dFdx"]) + t.dense(out["dQdx"]) / h + np.diag(C / h)
            return rhs, jac

        res = trap_residual if method == "trap" and n >= 2 else tr_residual
        x, out, x_load, its = _newton(loader, ring, x_n, {ring.supply}, tr_flags, res,
                                      vtol, max_iter)
        newton_total += its
        loader.accept()
        q_n = out["q"] + C * x_load
        f_n = out["f"].copy()
        times.append((n + 1) * h)
        V.append(x.copy())
    stats = {"steps": steps, "newton_iterations": newton_total, "loads": loader.loads,
             "load_seconds": loader.load_seconds}
    return np.asarray(times), np.asarray(V), stats


def crossings(times, v, level, rising=True):
    """Times at which v crosses `level` (linear interpolation)."""
    s = v - level
    idx = np.nonzero((s[:-1] < 0) & (s[1:] >= 0) if rising else (s[:-1] > 0) & (s[1:] <= 0))[0]
    return times[idx] + (times[idx + 1] - times[idx]) * (-s[idx]) / (s[idx + 1] - s[idx])


def period(times, v, level):
    """Mean oscillation period from successive rising crossings (None if
    fewer than two)."""
    c = crossings(times, v, level, rising=True)
    if len(c) < 2:
        return None
    return float(np.mean(np.diff(c)))
