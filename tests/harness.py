"""Test harness: one synthetic circuit driven through the oracle (host numpy)
and through the GPU library (device tensors) with identical setup, LIDs,
CSR structure and load sequence, so that every output can be compared.

The sequence mirrors the reference's Newton load (N_LOA_NonlinearEquationLoader.C
:120-249): zero F/Q/limiter vectors -> updateState -> updateSecondaryState ->
loadDAEVectors; zero dFdx/dQdx -> loadDAEMatrices.
"""
import numpy as np

OUT_VECS = ("f", "q", "dFdxdVp", "dQdxdVp", "storeLeadF", "storeLeadQ")
STATE_VECS = ("nextSto", "nextSta", "currSta")


def splitmix64(seed):
    """SplitMix64 stream as a numpy Generator seed (seeded inputs)."""
    return np.random.default_rng(np.random.SeedSequence(seed))


class Topology:
    """LIDs and CSR structure shared by both back ends."""

    def __init__(self, counts, mode, rng, n_shared_nodes=None, ground_frac=0.0):
        n_ext, n_int, n_sta, n_sto = counts
        n = len(n_ext)
        nodes = []
        next_lid = 0
        if mode == "private":
            for i in range(n):
                k = n_ext[i] + n_int[i]
                nodes.append(np.arange(next_lid, next_lid + k, dtype=np.int32))
                next_lid += k
        elif mode == "shared":
            pool = n_shared_nodes or max(2, n // 2)
            next_lid = pool
            for i in range(n):
                ext = rng.integers(0, pool, size=n_ext[i]).astype(np.int32)
                if ground_frac > 0:
                    ext[rng.random(n_ext[i]) < ground_frac] = -1
                it = np.arange(next_lid, next_lid + n_int[i], dtype=np.int32)
                next_lid += n_int[i]
                nodes.append(np.concatenate([ext, it]))
        else:
            raise ValueError(mode)
        self.n_nodes = next_lid
        self.node_lids = np.concatenate(nodes) if nodes else np.zeros(0, np.int32)
        self.n_sta = int(np.sum(n_sta))
        self.n_sto = int(np.sum(n_sto))
        self.sta_lids = np.arange(self.n_sta, dtype=np.int32)
        self.sto_lids = np.arange(self.n_sto, dtype=np.int32)

    def build_csr(self, rows, cols):
        keys = np.unique(rows.astype(np.int64) * (self.n_nodes + 1) + cols)
        r = (keys // (self.n_nodes + 1)).astype(np.int32)
        c = (keys % (self.n_nodes + 1)).astype(np.int32)
        row_ptr = np.zeros(self.n_nodes + 1, np.int64)
        np.add.at(row_ptr, r + 1, 1)
        self.row_ptr = np.cumsum(row_ptr).astype(np.int32)
        self.col_idx = c
        self.nnz = len(c)


def build_group(master, spec, lead=None):
    """Feed a circuit spec (models + instance parameter columns) to a Master."""
    midx = []
    for level, params in spec["models"]:
        midx.append(master.add_model(params, level=level))
    master.add_instances(spec["n"], np.asarray(spec["model_index"], np.int32),
                         spec.get("inst_params", {}))
    if lead is not None:
        master.request_lead_current(lead)
    master.setup()
    return master.counts()


def finish_group(master, topo):
    master.register_lids(topo.node_lids, topo.sta_lids, topo.sto_lids)
    return master.jac_pairs()


class Pair:
    """The same circuit on the oracle and on the GPU."""

    def __init__(self, model_type, spec, options=None, topology="private", seed=1,
                 lead=None, gpu=True, ground_frac=0.0, n_shared_nodes=None):
        import oracle
        self.rng = splitmix64(seed)
        self.spec = spec
        self.orc = oracle.Master(model_type, _copy_opts(options, oracle.DeviceOptions)
                                 if options else None)
        counts = build_group(self.orc, spec, lead)
        self.counts = counts
        self.topo = Topology(counts, topology, self.rng, n_shared_nodes, ground_frac)
        r, c = finish_group(self.orc, self.topo)
        self.topo.build_csr(r, c)
        self.orc.register_csr(self.topo.row_ptr, self.topo.col_idx)
        self.orc.finalize()
        self.gpu = None
        if gpu:
            import torch
            import xyce_amd
            self.torch = torch
            self.ctx = xyce_amd.Context(0)
            self.gpu = xyce_amd.Master(self.ctx, model_type,
                                       _copy_opts(options, xyce_amd.DeviceOptions) if options else None)
            gcounts = build_group(self.gpu, spec, lead)
            for a, b in zip(counts, gcounts):
                np.testing.assert_array_equal(a, b)
            gr, gc = finish_group(self.gpu, self.topo)
            got = set(zip(gr.tolist(), gc.tolist()))
            want = set(zip(r.tolist(), c.tolist()))
            assert got == want, "GPU and oracle Jacobian patterns differ"
            self.gpu.register_csr(self.topo.row_ptr, self.topo.col_idx)
            self.gpu.finalize()
        t = self.topo
        self.h = {
            "nextSol": np.zeros(t.n_nodes), "currSol": np.zeros(t.n_nodes),
            "nextSta": np.zeros(t.n_sta), "currSta": np.zeros(t.n_sta), "lastSta": np.zeros(t.n_sta),
            "nextStaDeriv": np.zeros(t.n_sta),
            "nextSto": np.zeros(t.n_sto), "currSto": np.zeros(t.n_sto),
            "f": np.zeros(t.n_nodes), "q": np.zeros(t.n_nodes), "b": np.zeros(t.n_nodes),
            "dFdxdVp": np.zeros(t.n_nodes), "dQdxdVp": np.zeros(t.n_nodes),
            "storeLeadF": np.zeros(t.n_sto), "storeLeadQ": np.zeros(t.n_sto),
            "dFdx": np.zeros(t.nnz), "dQdx": np.zeros(t.nnz),
        }
        if self.gpu is not None:
            dev = torch.device("cuda", 0)
            self.d = {k: torch.zeros(max(len(v), 1), dtype=torch.float64, device=dev)
                      for k, v in self.h.items()}

    # ---- inputs ----
    def set_inputs(self, **arrays):
        """Set named host arrays and mirror them to the device."""
        for k, a in arrays.items():
            self.h[k][:] = a
            if self.gpu is not None and len(a):
                self.d[k][:len(a)] = self.torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()

    def _vecs(self, which):
        import oracle
        if which == "orc":
            return oracle.make_vectors(**{k: v for k, v in self.h.items() if len(v)})
        import xyce_amd
        return xyce_amd.make_vectors(**self.d)

    def _ss(self, which, **flags):
        if which == "orc":
            import oracle
            ss = oracle.SolverState()
        else:
            import xyce_amd
            ss = xyce_amd.SolverState()
        ss.gainScale = 1.0
        ss.nltermScale = 1.0
        for k, v in flags.items():
            setattr(ss, k, v)
        return ss

    def load(self, fused=False, **flags):
        """One full load on both back ends; returns (oracle outs, gpu outs)."""
        for k in OUT_VECS + ("dFdx", "dQdx"):
            self.h[k][:] = 0.0
        vo = self._vecs("orc")
        so = self._ss("orc", **flags)
        self.orc.updateState(so, vo)
        self.orc.updateSecondaryState(so, vo)
        self.orc.loadDAEVectors(so, vo)
        self.orc.loadDAEMatrices(so, vo)
        res_o = {k: self.h[k].copy() for k in OUT_VECS + STATE_VECS + ("dFdx", "dQdx")}
        # error scale of each assembled entry: sum of |contributions|
        import oracle
        mag = {k: np.zeros_like(self.h[k]) for k in OUT_VECS + ("dFdx", "dQdx")}
        vm = oracle.make_vectors(**{k: (mag[k] if k in mag else v)
                                    for k, v in self.h.items() if len(v)})
        oracle.lib().or_set_magnitude_mode(1)
        try:
            self.orc.loadDAEVectors(so, vm)
            self.orc.loadDAEMatrices(so, vm)
        finally:
            oracle.lib().or_set_magnitude_mode(0)
        self.mag = mag
        if self.gpu is None:
            return res_o, None
        for k in OUT_VECS + ("dFdx", "dQdx"):
            self.d[k].zero_()
        vg = self._vecs("gpu")
        sg = self._ss("gpu", **flags)
        if fused:
            self.gpu.loadAll(sg, vg)
        else:
            self.gpu.updateState(sg, vg)
            self.gpu.updateSecondaryState(sg, vg)
            self.gpu.loadDAEVectors(sg, vg)
            self.gpu.loadDAEMatrices(sg, vg)
        self.torch.cuda.synchronize()
        res_g = {}
        for k in OUT_VECS + STATE_VECS + ("dFdx", "dQdx"):
            n = len(self.h[k])
            res_g[k] = self.d[k][:n].cpu().numpy().copy()
        return res_o, res_g


def _copy_opts(src, cls):
    dst = cls()
    for name, _ in src._fields_:
        setattr(dst, name, getattr(src, name))
    return dst


def max_rel_err(a, b, floor=1e-300):
    """Largest |a-b| / max(|a|,|b|) over entries (0 where both are 0)."""
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.maximum(np.maximum(np.abs(a), np.abs(b)), floor)
    return float(np.max(np.abs(a - b) / den)) if a.size else 0.0


def assert_close(a, b, rtol, atol=0.0, what="", scale=None):
    """|a-b| <= rtol * max(|a|, |b|, scale) + atol, entrywise.  `scale` is
    the sum of |contributions| of an assembled entry (harness Pair.mag)."""
    a = np.asarray(a)
    b = np.asarray(b)
    ref = np.maximum(np.abs(a), np.abs(b))
    if scale is not None:
        ref = np.maximum(ref, np.asarray(scale))
    bad = np.abs(a - b) > rtol * ref + atol
    if np.any(bad):
        idx = np.nonzero(bad)[0][:5]
        raise AssertionError(
            f"{what}: {int(bad.sum())} of {a.size} entries differ; first {idx.tolist()}: "
            f"{a[idx].tolist()} vs {b[idx].tolist()}; max rel {max_rel_err(a, b):.3e}")
