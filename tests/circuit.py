"""Multi-group circuit harness: several device types on one set of circuit
vectors and one CSR structure, loaded on the oracle and on the GPU in the
reference's sequence.

DeviceMgr loads every device type in turn onto the same vectors
(N_DEV_DeviceMgr.C:2053-2502: updateState over all types, then
updateSecondaryState, loadDAEVectors and loadDAEMatrices over all types, in
creation order), so an entry shared by two types holds type A's terms
followed by type B's.  The GPU groups add onto the same device vectors in
that order; both back ends see the same LIDs, CSR and inputs.
"""
import numpy as np

from harness import OUT_VECS, STATE_VECS, _copy_opts, splitmix64


class Part:
    """One device type of a circuit: models, instances and their external
    node LIDs (instance-major; ground -1).  Internal nodes get fresh LIDs."""

    def __init__(self, model, models, model_index, params, ext_lids, lead=None):
        self.model = model
        self.models = models
        self.model_index = np.asarray(model_index, np.int32)
        self.params = params
        self.ext_lids = np.asarray(ext_lids, np.int32)
        self.lead = lead
        self.n = len(self.model_index)


def _build(master, part):
    for level, params in part.models:
        master.add_model(params, level=level)
    master.add_instances(part.n, part.model_index, part.params)
    if part.lead is not None:
        master.request_lead_current(part.lead)
    master.setup()
    return master.counts()


class CircuitPair:
    """The same multi-type circuit on the oracle and on the GPU."""

    def __init__(self, parts, n_nodes, noise_runs=0, noise_ulps=2, gpu=True, options=None):
        import oracle
        self.parts = parts
        self.noise_runs = noise_runs
        self.noise_ulps = noise_ulps
        self._loads = 0
        self.orc = [oracle.Master(p.model, _copy_opts(options, oracle.DeviceOptions)
                                  if options else None) for p in parts]
        counts = [_build(m, p) for m, p in zip(self.orc, parts)]
        nxt = n_nodes
        sta0 = sto0 = 0
        self.lids = []
        for p, (n_ext, n_int, n_sta, n_sto) in zip(parts, counts):
            if len(p.ext_lids) != int(n_ext.sum()):
                raise ValueError(f"{p.model}: {len(p.ext_lids)} external LIDs for "
                                 f"{int(n_ext.sum())} terminals")
            per_ext = np.split(p.ext_lids, np.cumsum(n_ext)[:-1]) if p.n else []
            rows = []
            for i in range(p.n):
                rows.append(per_ext[i])
                rows.append(np.arange(nxt, nxt + n_int[i], dtype=np.int32))
                nxt += int(n_int[i])
            node = np.concatenate(rows).astype(np.int32) if rows else np.zeros(0, np.int32)
            sta = np.arange(sta0, sta0 + int(n_sta.sum()), dtype=np.int32)
            sto = np.arange(sto0, sto0 + int(n_sto.sum()), dtype=np.int32)
            sta0 += len(sta)
            sto0 += len(sto)
            self.lids.append((node, sta, sto))
        self.n_nodes, self.n_sta, self.n_sto = nxt, sta0, sto0
        rows, cols = [], []
        for m, lids in zip(self.orc, self.lids):
            m.register_lids(*lids)
            r, c = m.jac_pairs()
            rows.append(r)
            cols.append(c)
        r = np.concatenate(rows).astype(np.int64)
        c = np.concatenate(cols).astype(np.int64)
        keys = np.unique(r * (self.n_nodes + 1) + c)
        kr = (keys // (self.n_nodes + 1)).astype(np.int64)
        rp = np.zeros(self.n_nodes + 1, np.int64)
        np.add.at(rp, kr + 1, 1)
        self.row_ptr = np.cumsum(rp).astype(np.int32)
        self.col_idx = (keys % (self.n_nodes + 1)).astype(np.int32)
        self.nnz = len(self.col_idx)
        for m in self.orc:
            m.register_csr(self.row_ptr, self.col_idx)
            m.finalize()
        self.gpu = None
        if gpu:
            import torch
            import xyce_amd
            self.torch = torch
            self.ctx = xyce_amd.Context(0)
            self.gpu = [xyce_amd.Master(self.ctx, p.model,
                                        _copy_opts(options, xyce_amd.DeviceOptions)
                                        if options else None) for p in parts]
            for m, p, cnt, lids in zip(self.gpu, parts, counts, self.lids):
                for a, b in zip(cnt, _build(m, p)):
                    np.testing.assert_array_equal(a, b)
                m.register_lids(*lids)
                m.register_csr(self.row_ptr, self.col_idx)
                m.finalize()
        n, ns, nt = self.n_nodes, self.n_sta, self.n_sto
        self.h = {"nextSol": np.zeros(n), "currSol": np.zeros(n), "flagSol": np.zeros(n),
                  "nextSta": np.zeros(ns), "currSta": np.zeros(ns), "lastSta": np.zeros(ns),
                  "nextStaDeriv": np.zeros(ns), "nextSto": np.zeros(nt), "currSto": np.zeros(nt),
                  "f": np.zeros(n), "q": np.zeros(n), "b": np.zeros(n), "dFdxdVp": np.zeros(n),
                  "dQdxdVp": np.zeros(n), "storeLeadF": np.zeros(nt), "storeLeadQ": np.zeros(nt),
                  "dFdx": np.zeros(self.nnz), "dQdx": np.zeros(self.nnz)}
        if self.gpu is not None:
            dev = self.torch.device("cuda", 0)
            self.d = {k: self.torch.zeros(max(len(v), 1), dtype=self.torch.float64, device=dev)
                      for k, v in self.h.items()}

    def unknowns(self, volts, branch=0.0):
        """A solution vector: the given node voltages followed by `branch`
        for the unknowns the device types added (internal nodes, branch
        currents)."""
        x = np.full(self.n_nodes, float(branch))
        x[:len(volts)] = volts
        return x

    def set_inputs(self, **arrays):
        for k, a in arrays.items():
            self.h[k][:] = a
            if self.gpu is not None and len(a):
                self.d[k][:len(a)] = self.torch.from_numpy(
                    np.ascontiguousarray(a, np.float64)).cuda()

    @staticmethod
    def _ss(mod, flags):
        ss = mod.SolverState()
        ss.gainScale = 1.0
        ss.nltermScale = 1.0
        for k, v in flags.items():
            setattr(ss, k, v)
        return ss

    def _sequence(self, masters, ss, v):
        """DeviceMgr's order: each call over every type before the next call."""
        for m in masters:
            m.updateState(ss, v)
        for m in masters:
            m.updateSecondaryState(ss, v)
        for m in masters:
            m.loadDAEVectors(ss, v)
        for m in masters:
            m.loadDAEMatrices(ss, v)

    def _oracle_pass(self, ss):
        import oracle
        for k in OUT_VECS + ("dFdx", "dQdx"):
            self.h[k][:] = 0.0
        v = oracle.make_vectors(**{k: a for k, a in self.h.items() if len(a)})
        self._sequence(self.orc, ss, v)
        return {k: self.h[k].copy() for k in OUT_VECS + STATE_VECS + ("dFdx", "dQdx")}

    def load(self, **flags):
        """One Newton load on both back ends: (oracle outputs, GPU outputs);
        sets self.mag (sum of |contributions| per entry) and self.noise (the
        entries' spread under libm perturbation of the oracle)."""
        import oracle
        so = self._ss(oracle, flags)
        self._loads += 1
        runs = []
        if self.noise_runs:
            inputs = {k: v.copy() for k, v in self.h.items()}
            for m in self.orc:
                m.save_state()
            for r in range(self.noise_runs):
                for m in self.orc:
                    m.restore_state()
                for k, v in inputs.items():
                    self.h[k][:] = v
                oracle.lib().or_set_libm_perturbation(self.noise_ulps,
                                                      1000003 * self._loads + r + 1)
                try:
                    runs.append(self._oracle_pass(so))
                finally:
                    oracle.lib().or_set_libm_perturbation(0, 0)
            for m in self.orc:
                m.restore_state()
            for k, v in inputs.items():
                self.h[k][:] = v
        res_o = self._oracle_pass(so)
        self.noise = ({k: np.max([np.abs(r[k] - res_o[k]) for r in runs], axis=0)
                       for k in res_o} if runs else {})
        mag = {k: np.zeros_like(self.h[k]) for k in OUT_VECS + ("dFdx", "dQdx")}
        vm = oracle.make_vectors(**{k: (mag[k] if k in mag else v)
                                    for k, v in self.h.items() if len(v)})
        oracle.lib().or_set_magnitude_mode(1)
        try:
            for m in self.orc:
                m.loadDAEVectors(so, vm)
            for m in self.orc:
                m.loadDAEMatrices(so, vm)
        finally:
            oracle.lib().or_set_magnitude_mode(0)
        self.mag = mag
        if self.gpu is None:
            return res_o, None
        import xyce_amd
        for k in OUT_VECS + ("dFdx", "dQdx"):
            self.d[k].zero_()
        self._sequence(self.gpu, self._ss(xyce_amd, flags), xyce_amd.make_vectors(**self.d))
        self.torch.cuda.synchronize()
        res_g = {k: self.d[k][:len(self.h[k])].cpu().numpy().copy()
                 for k in OUT_VECS + STATE_VECS + ("dFdx", "dQdx")}
        return res_o, res_g


def two_terminal(model, nodes, values, name, card=None):
    """A Part of resistors or capacitors: nodes[k] = (pos, neg)."""
    nodes = np.asarray(nodes, np.int32).reshape(-1, 2)
    n = len(nodes)
    return Part(model, [(1, dict(card or {}))], np.zeros(n, np.int32),
                {name: np.asarray(values, np.float64)}, nodes.reshape(-1))


def rng_for(seed):
    return splitmix64(seed)
