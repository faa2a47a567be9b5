"""Test harness: one synthetic circuit driven through the oracle (host numpy)
and through the GPU library (device tensors) with identical setup, LIDs,
CSR structure and load sequence, so that every output can be compared.

The sequence mirrors the reference's Newton load (N_LOA_NonlinearEquationLoader.C
:120-249): zero F/Q/limiter vectors -> updateState -> updateSecondaryState ->
loadDAEVectors; zero dFdx/dQdx -> loadDAEMatrices.
"""
import numpy as np

OUT_VECS = ("f", "q", "dFdxdVp", "dQdxdVp", "storeLeadF", "storeLeadQ")
STATE_VECS = ("nextSto", "nextSta", "currSta", "lastSta")


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
        elif mode == "blocked":
            # nodes shared only within groups of 64 consecutive instances
            # (so every target is written by one block of the fused load)
            pool = n_shared_nodes or 24
            groups = (n + 63) // 64
            next_lid = groups * pool
            for i in range(n):
                g = i // 64
                ext = (g * pool + rng.integers(0, pool, size=n_ext[i])).astype(np.int32)
                if ground_frac > 0:
                    ext[rng.random(n_ext[i]) < ground_frac] = -1
                it = np.arange(next_lid, next_lid + n_int[i], dtype=np.int32)
                next_lid += n_int[i]
                nodes.append(np.concatenate([ext, it]))
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
                 lead=None, gpu=True, ground_frac=0.0, n_shared_nodes=None,
                 noise_runs=0, noise_ulps=2, variant=""):
        """variant "pm": both back ends in their parity builds (shared
        portable exp/log/pow/tan), where every output must match bit for bit."""
        import oracle
        self.variant = variant
        self.noise_runs = noise_runs
        self.noise_ulps = noise_ulps
        self.noise = {}
        self._loads = 0
        self.rng = splitmix64(seed)
        self.spec = spec
        self.orc = oracle.Master(model_type, _copy_opts(options, oracle.DeviceOptions)
                                 if options else None, variant=variant)
        self.olib = self.orc.L
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
            self.ctx = xyce_amd.Context(0, variant=variant)
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
            "flagSol": np.zeros(t.n_nodes),
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

    def _oracle_pass(self, so):
        for k in OUT_VECS + ("dFdx", "dQdx"):
            self.h[k][:] = 0.0
        vo = self._vecs("orc")
        self.orc.updateState(so, vo)
        self.orc.updateSecondaryState(so, vo)
        self.orc.loadDAEVectors(so, vo)
        self.orc.loadDAEMatrices(so, vo)
        return {k: self.h[k].copy() for k in OUT_VECS + STATE_VECS + ("dFdx", "dQdx")}

    def load(self, fused=False, **flags):
        """One full load on both back ends; returns (oracle outs, gpu outs).

        With noise_runs > 0 the oracle first repeats the load from the same
        inputs and instance state with its exp/log/pow results moved by up to
        noise_ulps ulps (different draws per run); self.noise[k] is the
        largest departure of those runs from the unperturbed load: the
        precision to which libm-level rounding determines each entry."""
        import oracle
        so = self._ss("orc", **flags)
        self._loads += 1
        runs = []
        if self.noise_runs:
            inputs = {k: v.copy() for k, v in self.h.items()}
            self.orc.save_state()
            for r in range(self.noise_runs):
                self.orc.restore_state()
                for k, v in inputs.items():
                    self.h[k][:] = v
                self.olib.or_set_libm_perturbation(self.noise_ulps, 1000003 * self._loads + r + 1)
                try:
                    runs.append(self._oracle_pass(so))
                finally:
                    self.olib.or_set_libm_perturbation(0, 0)
            self.orc.restore_state()
            for k, v in inputs.items():
                self.h[k][:] = v
        res_o = self._oracle_pass(so)
        self.noise = {k: np.max([np.abs(r[k] - res_o[k]) for r in runs], axis=0)
                      for k in res_o} if runs else {}
        # error scale of each assembled entry: sum of |contributions|
        import oracle
        mag = {k: np.zeros_like(self.h[k]) for k in OUT_VECS + ("dFdx", "dQdx")}
        vm = oracle.make_vectors(**{k: (mag[k] if k in mag else v)
                                    for k, v in self.h.items() if len(v)})
        self.olib.or_set_magnitude_mode(1)
        try:
            self.orc.loadDAEVectors(so, vm)
            self.orc.loadDAEMatrices(so, vm)
        finally:
            self.olib.or_set_magnitude_mode(0)
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


NOISE_FACTOR = 2.0


def assert_close(a, b, rtol, atol=0.0, what="", scale=None, noise=None):
    """|a-b| <= rtol * max(|a|, |b|, scale) + NOISE_FACTOR * noise + atol,
    entrywise.  `scale` is the sum of |contributions| of an assembled entry
    (harness Pair.mag); `noise` its libm-rounding spread (Pair.noise)."""
    a = np.asarray(a)
    b = np.asarray(b)
    ref = np.maximum(np.abs(a), np.abs(b))
    if scale is not None:
        ref = np.maximum(ref, np.asarray(scale))
    tol = rtol * ref + atol
    if noise is not None:
        tol = tol + NOISE_FACTOR * np.asarray(noise)
    bad = np.abs(a - b) > tol
    if np.any(bad):
        idx = np.nonzero(bad)[0][:5]
        raise AssertionError(
            f"{what}: {int(bad.sum())} of {a.size} entries differ; first {idx.tolist()}: "
            f"{a[idx].tolist()} vs {b[idx].tolist()}; max rel {max_rel_err(a, b):.3e}")


# Parity bar of a GPU load against the oracle.  Every entry must lie within
# RTOL of its error scale S = max(|gpu|, |oracle|, sum of |contributions|) or
# within NOISE_FACTOR times its libm noise (how far the entry moves when the
# oracle's exp/log/pow results move by a few ulps, Pair.noise), or within
# NORM_FLOOR of the output's largest entry: an entry that is a cancellation
# residue many orders below its neighbours (a 1e-28 F capacitance in a block
# of 1e-17 F) carries the last-bit differences of its large intermediates,
# amplified past any entrywise bar, while a difference below 1e-15 of the
# output's norm is below the rounding of every product a Newton step forms
# with it.  (The code paths themselves are shown identical bit for bit by the
# parity builds, tests/test_gpu_bitexact.py.)  A discrete
# event -- a branch taken one way or the other by a last-bit difference in a
# transcendental -- can still put a few entries outside that envelope (seen:
# about 1 in 10^4 entries, mostly ~1e-11 of their scale, the largest 3e-7 on
# capacitances of 1e-25 F); at most OUTLIER_FRAC of the entries (and at least
# 2) may do so, and none may be further than OUTLIER_CAP * S.  A wrong term,
# sign or stamp slot moves entries by far more than either bound.
RTOL = 1e-12
NORM_FLOOR = 1e-15
OUTLIER_FRAC = 5e-4
OUTLIER_CAP = 1e-6


def parity_stats(a, b, scale=None, noise=None):
    a = np.asarray(a)
    b = np.asarray(b)
    S = np.maximum(np.abs(a), np.abs(b))
    if scale is not None:
        S = np.maximum(S, np.asarray(scale))
    d = np.abs(a - b)
    bar = RTOL * S + (NOISE_FACTOR * np.asarray(noise) if noise is not None else 0.0)
    if b.size:
        bar = np.maximum(bar, NORM_FLOOR * float(np.max(np.abs(b))))
    out = d > bar
    rel = d / np.maximum(S, 1e-300)
    return {
        "n": int(a.size),
        "bitexact": float(np.mean(a.view(np.int64) == b.view(np.int64))) if a.size else 1.0,
        "strict_fail": int(np.sum(d > RTOL * S)),
        "outliers": int(np.sum(out)),
        "max_rel": float(rel.max()) if a.size else 0.0,
        "max_rel_outlier": float(rel[out].max()) if out.any() else 0.0,
        "out_idx": np.nonzero(out)[0][:5].tolist(),
    }


def assert_parity(a, b, what="", scale=None, noise=None):
    st = parity_stats(a, b, scale, noise)
    allowed = max(2, int(OUTLIER_FRAC * st["n"]))
    if st["outliers"] > allowed or st["max_rel_outlier"] > OUTLIER_CAP:
        idx = st["out_idx"]
        raise AssertionError(
            f"{what}: {st['outliers']} of {st['n']} entries outside the bar (allowed {allowed}), "
            f"worst outlier rel {st['max_rel_outlier']:.3e}; first {idx}: "
            f"{np.asarray(a)[idx].tolist()} vs {np.asarray(b)[idx].tolist()}")
    return st
