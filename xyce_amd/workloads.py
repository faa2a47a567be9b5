"""Synthetic device-load workloads built through the product ABI.

``microbench(model, n)`` is the device-evaluation microbenchmark: ``n``
instances of one model type, each with private nodes (disjoint CSR blocks),
terminal voltages i.i.d. U[-0.5, 1.5] V from a seed, previous-iteration store
values from a second draw, and per-instance geometry spread lognormally
(sigma 0.2) -- W/L around 90 nm / 45 nm for MOSFETs, AREA around 1 for the
diode and BJT, R around 1 kOhm and C around 1 fF for the linear devices.
Solver state: transient, newtonIter = 1, limiter on.
"""
import numpy as np
import torch

from . import Context, Master, default_options, default_solver_state, make_vectors

SEED = 20261203

# Model cards used by the microbenchmark (the BSIM4 card is the default
# level-54 card with only the type given).
CARDS = {
    "diode": [(1, {"IS": 1e-14, "N": 1.0, "CJO": 1e-12, "VJ": 0.8, "M": 0.5, "TT": 1e-9})],
    "mosfet1": [(1, {"TYPE": 1.0, "VTO": 0.45, "KP": 2e-4, "GAMMA": 0.4, "PHI": 0.7, "LAMBDA": 0.05,
                     "TOX": 2e-9, "CGSO": 3e-10, "CGDO": 3e-10, "CJ": 1e-3, "MJ": 0.5,
                     "CJSW": 3e-10, "MJSW": 0.33}),
                 (1, {"TYPE": -1.0, "VTO": -0.45, "KP": 1e-4, "GAMMA": 0.4, "PHI": 0.7, "LAMBDA": 0.05,
                      "TOX": 2e-9, "CGSO": 3e-10, "CGDO": 3e-10, "CJ": 1e-3, "MJ": 0.5,
                      "CJSW": 3e-10, "MJSW": 0.33})],
    "bjt": [(1, {"TYPE": 1.0, "IS": 1e-16, "BF": 100.0, "BR": 1.0, "VAF": 50.0, "IKF": 0.01,
                 "ISE": 1e-14, "NE": 1.5, "RB": 10.0, "RC": 1.0, "RE": 0.5, "CJE": 1e-12,
                 "CJC": 5e-13, "TF": 1e-10, "TR": 1e-8})],
    "bsim4": [(54, {"TYPE": 1.0}), (54, {"TYPE": -1.0})],
    "resistor": [(1, {})],
    "capacitor": [(1, {})],
}


def lognormal(rng, median, sigma, n):
    return median * np.exp(sigma * rng.standard_normal(n))


def instance_params(model, n, rng):
    if model in ("diode", "bjt"):
        return {"AREA": lognormal(rng, 1.0, 0.2, n)}
    if model == "resistor":
        return {"R": lognormal(rng, 1000.0, 0.2, n)}
    if model == "capacitor":
        return {"C": lognormal(rng, 1e-15, 0.2, n)}
    return {"W": lognormal(rng, 90e-9, 0.2, n), "L": lognormal(rng, 45e-9, 0.2, n)}


class Workload:
    """A built group plus its device vectors."""

    def __init__(self, model, n, seed=SEED, device=0, options=None, cards=None,
                 model_index=None, params=None, netlist=None):
        """netlist: (node_lids, n_nodes, voltages) of a shared-node netlist
        (xyce_amd.netlists); default: private nodes, V ~ U[-0.5, 1.5]."""
        self.model = model
        self.n = n
        rng = np.random.default_rng(np.random.SeedSequence(seed))
        self.ctx = Context(device)
        self.master = Master(self.ctx, model, options or default_options())
        cards = cards or CARDS[model]
        for level, card in cards:
            self.master.add_model(card, level=level)
        if model_index is None:
            model_index = rng.integers(0, len(cards), size=n).astype(np.int32)
        self.master.add_instances(n, model_index, params or instance_params(model, n, rng))
        self.master.setup()
        n_ext, n_int, n_sta, n_sto = self.master.counts()
        per = (n_ext + n_int).astype(np.int64)
        if netlist is None:
            self.n_nodes = int(per.sum())
            node_lids = np.arange(self.n_nodes, dtype=np.int32)      # private nodes
            volts = rng.uniform(-0.5, 1.5, self.n_nodes)
        else:
            ext_lids, n_nodes, volts = netlist
            if n_int.any():
                raise ValueError("netlist workloads assume cards without internal nodes")
            if len(ext_lids) != int(n_ext.sum()):
                raise ValueError("netlist gives the wrong number of terminal nodes")
            node_lids = np.ascontiguousarray(ext_lids, np.int32)
            self.n_nodes = int(n_nodes)
        self.n_sta = int(n_sta.sum())
        self.n_sto = int(n_sto.sum())
        self.master.register_lids(node_lids, np.arange(self.n_sta, dtype=np.int32),
                                  np.arange(self.n_sto, dtype=np.int32))
        rows, cols = self.master.jac_pairs()
        key = np.unique(rows.astype(np.int64) * self.n_nodes + cols)
        r = (key // self.n_nodes).astype(np.int64)
        c = (key % self.n_nodes).astype(np.int32)
        row_ptr = np.zeros(self.n_nodes + 1, np.int64)
        np.add.at(row_ptr, r + 1, 1)
        self.row_ptr = np.cumsum(row_ptr).astype(np.int32)
        self.col_idx = c
        self.nnz = len(c)
        self.master.register_csr(self.row_ptr, self.col_idx)
        self.master.finalize()
        dev = torch.device("cuda", device)
        z = lambda m: torch.zeros(max(m, 1), dtype=torch.float64, device=dev)
        self.vec = {
            "nextSol": torch.from_numpy(np.ascontiguousarray(volts, np.float64)).to(dev),
            "currSol": z(self.n_nodes),
            "nextSta": z(self.n_sta), "currSta": z(self.n_sta), "lastSta": z(self.n_sta),
            "nextStaDeriv": z(self.n_sta),
            "nextSto": torch.from_numpy(rng.uniform(-0.5, 1.5, max(self.n_sto, 1))).to(dev),
            "currSto": torch.from_numpy(rng.uniform(-0.5, 1.5, max(self.n_sto, 1))).to(dev),
            "f": z(self.n_nodes), "q": z(self.n_nodes), "b": z(self.n_nodes),
            "dFdxdVp": z(self.n_nodes), "dQdxdVp": z(self.n_nodes),
            "storeLeadF": z(self.n_sto), "storeLeadQ": z(self.n_sto),
            "dFdx": z(self.nnz), "dQdx": z(self.nnz),
        }
        self.vectors = make_vectors(**self.vec)
        self.ss = default_solver_state(transientFlag=1, newtonIter=1)

    def zero_outputs(self):
        for k in ("f", "q", "dFdxdVp", "dQdxdVp", "dFdx", "dQdx"):
            self.vec[k].zero_()

    def load_pass(self):
        """One load: updateState, updateSecondaryState, loadDAEVectors,
        loadDAEMatrices (N_DEV_DeviceMgr.C:2053-2502 order)."""
        m = self.master
        m.updateState(self.ss, self.vectors)
        m.updateSecondaryState(self.ss, self.vectors)
        m.loadDAEVectors(self.ss, self.vectors)
        m.loadDAEMatrices(self.ss, self.vectors)

    def load_fused(self):
        self.master.loadAll(self.ss, self.vectors)
