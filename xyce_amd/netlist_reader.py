"""SPICE-netlist subset reader and circuit builder over the C ABI
(xa_netlist_* / xa_circuit_* in include/xyce_amd.h).

The parser and the builder are the library's host code: :class:`Netlist`
reads the netlist (no GPU needed), :class:`Circuit` builds its device groups,
LIDs and CSR structure on one GPU and runs the loads over every group in
creation order, as the reference's DeviceMgr does over its device types
(N_DEV_DeviceMgr.C:2053-2502).
"""
import ctypes as C

import numpy as np

from ._abi import MODEL_TYPES, SolverState, check, lib

TYPE_NAMES = {v: k for k, v in MODEL_TYPES.items()}


class Netlist:
    """A parsed netlist (xa_netlist)."""

    def __init__(self, text, variant=""):
        self.L = lib(variant)
        self.h = C.c_void_p()
        check(self.L.xa_netlist_parse(text.encode(), C.byref(self.h)), "xa_netlist_parse", self.L)
        a, b, c, d = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(self.L.xa_netlist_sizes(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)),
              "xa_netlist_sizes", self.L)
        self.n_nodes, self.n_parts, self.n_steps, self.n_ic = a.value, b.value, c.value, d.value

    def close(self):
        if self.h:
            self.L.xa_netlist_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def node_names(self):
        buf = C.create_string_buffer(256)
        out = []
        for k in range(self.n_nodes):
            check(self.L.xa_netlist_node_name(self.h, k, buf, 256), "node_name", self.L)
            out.append(buf.value.decode())
        return out

    def node(self, name):
        """Node id of `name` (case-insensitive; ground is -1)."""
        u = name.upper()
        if u in ("0", "GND"):
            return -1
        return self.node_names.index(u)

    @property
    def tran(self):
        """(tstep, tstop, tstart, tmax), or None without .TRAN."""
        out = (C.c_double * 4)()
        rc = self.L.xa_netlist_tran(self.h, out)
        return None if rc != 0 else tuple(out)

    @property
    def ic(self):
        """.IC pairs as (node ids, values)."""
        n = np.zeros(self.n_ic, np.int32)
        v = np.zeros(self.n_ic, np.float64)
        check(self.L.xa_netlist_ic(self.h, n.ctypes.data, v.ctypes.data), "ic", self.L)
        return n, v

    def part(self, k, step=-1):
        """Part k as a dict: type (model type name), models [(level,
        {name: value})], model_index, ext_nodes (instance-major node ids),
        params {name: column}."""
        t, nm, ni, ne, npar = (C.c_int() for _ in range(5))
        check(self.L.xa_netlist_part(self.h, k, C.byref(t), C.byref(nm), C.byref(ni),
                                     C.byref(ne), C.byref(npar)), "part", self.L)
        models = []
        for m in range(nm.value):
            lev, cnt = C.c_int(), C.c_int()
            check(self.L.xa_netlist_part_model(self.h, k, m, step, C.byref(lev), C.byref(cnt),
                                               None, None), "part_model", self.L)
            names = (C.c_char_p * max(cnt.value, 1))()
            vals = np.zeros(max(cnt.value, 1), np.float64)
            check(self.L.xa_netlist_part_model(self.h, k, m, step, C.byref(lev), C.byref(cnt),
                                               names, vals.ctypes.data), "part_model", self.L)
            models.append((lev.value, {names[j].decode(): float(vals[j])
                                       for j in range(cnt.value)}))
        n = ni.value
        mi = np.zeros(n, np.int32)
        ext = np.zeros(ne.value, np.int32)
        names = (C.c_char_p * max(npar.value, 1))()
        vals = np.zeros(max(npar.value * n, 1), np.float64)
        check(self.L.xa_netlist_part_instances(self.h, k, step, mi.ctypes.data, ext.ctypes.data,
                                               names, vals.ctypes.data), "part_instances", self.L)
        params = {names[p].decode(): vals[p * n:(p + 1) * n].copy() for p in range(npar.value)}
        return {"type": TYPE_NAMES[t.value], "models": models, "model_index": mi,
                "ext_nodes": ext, "params": params}

    def parts(self, step=-1):
        return [self.part(k, step) for k in range(self.n_parts)]


class Circuit:
    """A netlist built on one GPU (xa_circuit): its groups, LIDs and CSR.
    step: a .STEP point, -1 the netlist's values, -2 every point batched."""

    def __init__(self, ctx, netlist, step=-1, options=None):
        from . import default_options
        self.ctx = ctx
        self.L = ctx.L
        self.netlist = netlist
        self.h = C.c_void_p()
        opt = options if options is not None else default_options()
        check(self.L.xa_circuit_build(netlist.h, ctx.h, C.byref(opt), step, C.byref(self.h)),
              "xa_circuit_build", self.L)
        u, sa, so, ng, nc = (C.c_int() for _ in range(5))
        nnz = C.c_int64()
        check(self.L.xa_circuit_sizes(self.h, C.byref(u), C.byref(sa), C.byref(so),
                                      C.byref(nnz), C.byref(ng), C.byref(nc)), "sizes", self.L)
        self.n_unknowns, self.n_sta, self.n_sto = u.value, sa.value, so.value
        self.nnz, self.n_groups, self.n_copies = nnz.value, ng.value, nc.value

    def close(self):
        if self.h:
            self.L.xa_circuit_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def csr(self):
        rp = np.zeros(self.n_unknowns + 1, np.int32)
        ci = np.zeros(max(self.nnz, 1), np.int32)
        check(self.L.xa_circuit_csr(self.h, rp.ctypes.data, ci.ctypes.data), "csr", self.L)
        return rp, ci[:self.nnz]

    def group(self, k):
        """(xa_group handle, model type name) of group k."""
        g, t = C.c_void_p(), C.c_int()
        check(self.L.xa_circuit_group(self.h, k, C.byref(g), C.byref(t)), "group", self.L)
        return g, TYPE_NAMES[t.value]

    def group_lids(self, k):
        """Node LIDs of group k, instance-major (ground -1)."""
        cnt = C.c_int64()
        check(self.L.xa_circuit_group_lids(self.h, k, C.byref(cnt), None), "group_lids", self.L)
        out = np.zeros(max(cnt.value, 1), np.int32)
        check(self.L.xa_circuit_group_lids(self.h, k, C.byref(cnt), out.ctypes.data), "group_lids",
              self.L)
        return out[:cnt.value]

    def load(self, ss, **tensors):
        """One load of every group onto the given device vectors."""
        from . import make_vectors
        if not isinstance(ss, SolverState):
            raise TypeError("ss must be an xyce_amd.SolverState")
        v = make_vectors(**tensors)
        check(self.L.xa_circuit_load(self.h, C.byref(ss), C.byref(v)), "xa_circuit_load", self.L)
