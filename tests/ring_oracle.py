"""The ring oscillator's BSIM4 group on the CPU oracle, with the loader
interface of xyce_amd.ring.GpuLoader, so one transient driver runs both."""
import numpy as np

import oracle
from xyce_amd.ring import Topology


class OracleLoader:
    def __init__(self, ring, options=None):
        self.master = oracle.Master("bsim4", options)
        self.topo = Topology(self.master, ring)
        t = self.topo
        z = lambda m: np.zeros(max(m, 1))
        self.h = {"nextSol": z(t.n_nodes), "currSol": z(t.n_nodes), "nextSta": z(t.n_sta),
                  "currSta": z(t.n_sta), "lastSta": z(t.n_sta), "nextStaDeriv": z(t.n_sta),
                  "nextSto": z(t.n_sto), "currSto": z(t.n_sto), "f": z(t.n_nodes),
                  "q": z(t.n_nodes), "b": z(t.n_nodes), "dFdxdVp": z(t.n_nodes),
                  "dQdxdVp": z(t.n_nodes), "storeLeadF": z(t.n_sto), "storeLeadQ": z(t.n_sto),
                  "dFdx": z(t.nnz), "dQdx": z(t.nnz)}
        self.vectors = oracle.make_vectors(**self.h)
        self.loads = 0
        self.load_seconds = 0.0

    def load(self, x, **flags):
        import time
        ss = oracle.SolverState()
        ss.gainScale = 1.0
        ss.nltermScale = 1.0
        for k, v in flags.items():
            setattr(ss, k, v)
        h = self.h
        h["nextSol"][:] = x
        for k in ("f", "q", "dFdxdVp", "dQdxdVp", "dFdx", "dQdx"):
            h[k][:] = 0.0
        t0 = time.perf_counter()
        m = self.master
        m.updateState(ss, self.vectors)
        m.updateSecondaryState(ss, self.vectors)
        m.loadDAEVectors(ss, self.vectors)
        m.loadDAEMatrices(ss, self.vectors)
        self.load_seconds += time.perf_counter() - t0
        self.loads += 1
        n, nnz = self.topo.n_nodes, self.topo.nnz
        return {k: h[k][:(nnz if k in ("dFdx", "dQdx") else n)].copy()
                for k in ("f", "q", "dFdxdVp", "dQdxdVp", "dFdx", "dQdx")}

    def accept(self):
        self.h["currSto"][:] = self.h["nextSto"]
        self.h["currSta"][:] = self.h["nextSta"]
