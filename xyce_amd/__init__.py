"""xyce_amd -- MI355X device-load library: Python host mirror of Xyce's
device Master interface over the C ABI in include/xyce_amd.h.

A :class:`Master` holds every instance of one model type on the GPU and
offers the reference's four per-type calls (N_DEV_Device.h; default loops
N_DEV_DeviceMaster.h:643-725)::

    updateState(...)            # Master::updateState(solVec, staVec, stoVec)
    updateSecondaryState(...)   # Master::updateSecondaryState(staDerivVec, stoVec)
    loadDAEVectors(...)         # Master::loadDAEVectors(solVec, fVec, qVec, bVec, leadF, leadQ)
    loadDAEMatrices(...)        # Master::loadDAEMatrices(dFdx, dQdx)

Vectors and CSR value arrays are torch CUDA (HIP) float64 tensors; torch is
used only for device memory and streams.  The evaluation and stamping run in
the hand-written HIP kernels of libxyce_amd.so; there is no CPU path.
"""
import ctypes as C

import numpy as np

from . import _abi
from ._abi import (DeviceOptions, SolverState, Vectors, VECTOR_FIELDS, MODEL_TYPES,
                   XaError, check, lib)

__all__ = ["Context", "Master", "DeviceOptions", "SolverState", "Vectors",
           "default_options", "default_solver_state", "MODEL_TYPES", "XaError", "lib"]


def default_options(**kw):
    o = DeviceOptions()
    lib().xa_default_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def default_solver_state(**kw):
    s = SolverState()
    lib().xa_default_solver_state(C.byref(s))
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def _ptr(t):
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        if not t.is_cuda or t.dtype.itemsize != 8 or not t.is_contiguous():
            raise ValueError("vectors must be contiguous 8-byte device tensors")
        return t.data_ptr()
    raise TypeError("expected a torch device tensor")


def make_vectors(**tensors):
    v = Vectors()
    for k, t in tensors.items():
        if k not in VECTOR_FIELDS:
            raise KeyError(k)
        setattr(v, k, _ptr(t))
    return v


class Context:
    """One GPU and one stream (xa_ctx).  By default the library is bound to
    torch's current stream so torch allocations and library kernels order."""

    def __init__(self, device=0, use_torch_stream=True, variant=""):
        """variant: "" the production library, "pm" its parity build (the
        evaluation's exp/log/pow/tan shared with the oracle's parity build)."""
        self.L = lib(variant)
        self.variant = variant
        self.h = C.c_void_p()
        check(self.L.xa_ctx_create(device, C.byref(self.h)), "xa_ctx_create", self.L)
        self.device = device
        if use_torch_stream:
            import torch
            s = torch.cuda.current_stream(device)
            check(self.L.xa_ctx_set_stream(self.h, C.c_void_p(s.cuda_stream)), "set_stream", self.L)

    def set_stream(self, stream):
        """Order the library's work on `stream` (a torch.cuda.Stream), e.g. the
        stream a HIP graph is being captured on."""
        check(self.L.xa_ctx_set_stream(self.h, C.c_void_p(stream.cuda_stream)), "set_stream", self.L)

    def sync(self):
        check(self.L.xa_ctx_sync(self.h), "xa_ctx_sync", self.L)

    def close(self):
        if self.h:
            self.L.xa_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _cstrs(names):
    arr = (C.c_char_p * max(len(names), 1))()
    for i, n in enumerate(names):
        arr[i] = n.encode()
    return arr


class Master:
    """All instances of one model type on one GPU (an xa_group)."""

    def __init__(self, ctx, model_type, options=None):
        self.ctx = ctx
        self.L = ctx.L
        self.type = MODEL_TYPES[model_type] if isinstance(model_type, str) else model_type
        self.h = C.c_void_p()
        opt = options if options is not None else default_options()
        check(self.L.xa_group_create(ctx.h, self.type, C.byref(opt), C.byref(self.h)),
              "xa_group_create", self.L)
        self._keep = []

    # ---- setup, mirroring Model/Instance construction and topology ----
    def add_model(self, params, level=1):
        names = list(params.keys())
        vals = (C.c_double * max(len(names), 1))(*[float(params[k]) for k in names])
        idx = C.c_int()
        check(self.L.xa_group_add_model(self.h, level, len(names), _cstrs(names), vals,
                                       C.byref(idx)), "add_model", self.L)
        return idx.value

    def add_instances(self, n, model_index=None, params=None):
        params = params or {}
        names = list(params.keys())
        cols = [np.ascontiguousarray(np.broadcast_to(np.asarray(params[k], np.float64), (n,)))
                for k in names]
        self._keep.extend(cols)
        ptrs = (C.POINTER(C.c_double) * max(len(cols), 1))(
            *[c.ctypes.data_as(C.POINTER(C.c_double)) for c in cols])
        mi = None
        if model_index is not None:
            mi_arr = np.ascontiguousarray(np.broadcast_to(np.asarray(model_index, np.int32), (n,)))
            self._keep.append(mi_arr)
            mi = mi_arr.ctypes.data_as(C.POINTER(C.c_int))
        check(self.L.xa_group_add_instances(self.h, n, mi, len(names), _cstrs(names), ptrs),
              "add_instances", self.L)

    def request_lead_current(self, flags):
        f = np.ascontiguousarray(flags, np.int32)
        check(self.L.xa_group_request_lead_current(self.h, f.ctypes.data_as(C.POINTER(C.c_int32))),
              "lead", self.L)

    def setup(self):
        check(self.L.xa_group_setup(self.h), "setup", self.L)

    @property
    def n(self):
        return self.L.xa_group_num_instances(self.h)

    def counts(self):
        n = self.n
        out = [np.zeros(n, np.int32) for _ in range(4)]
        check(self.L.xa_group_counts(self.h, *[o.ctypes.data_as(C.POINTER(C.c_int)) for o in out]),
              "counts", self.L)
        return out

    def register_lids(self, node_lids, sta_lids, sto_lids):
        arrs = [np.ascontiguousarray(a, np.int32) for a in (node_lids, sta_lids, sto_lids)]
        check(self.L.xa_group_register_lids(
            self.h, *[a.ctypes.data_as(C.POINTER(C.c_int32)) for a in arrs]), "register_lids", self.L)

    def jac_pairs(self):
        cnt = C.c_int64()
        check(self.L.xa_group_jac_pairs(self.h, C.byref(cnt), None, None), "jac_pairs", self.L)
        r = np.zeros(cnt.value, np.int32)
        c = np.zeros(cnt.value, np.int32)
        check(self.L.xa_group_jac_pairs(self.h, C.byref(cnt), r.ctypes.data_as(C.POINTER(C.c_int32)),
                                       c.ctypes.data_as(C.POINTER(C.c_int32))), "jac_pairs", self.L)
        return r, c

    def register_csr(self, row_ptr, col_idx):
        rp = np.ascontiguousarray(row_ptr, np.int32)
        ci = np.ascontiguousarray(col_idx, np.int32)
        check(self.L.xa_group_register_csr(self.h, len(rp) - 1, rp.ctypes.data_as(C.POINTER(C.c_int32)),
                                          ci.ctypes.data_as(C.POINTER(C.c_int32))), "register_csr", self.L)

    def finalize(self):
        check(self.L.xa_group_finalize(self.h), "finalize", self.L)

    # ---- the four per-type loads ----
    def updateState(self, ss, vectors):
        check(self.L.xa_update_state(self.h, C.byref(ss), C.byref(vectors)), "updateState", self.L)

    def updateSecondaryState(self, ss, vectors):
        check(self.L.xa_update_secondary_state(self.h, C.byref(ss), C.byref(vectors)),
              "updateSecondaryState", self.L)

    def loadDAEVectors(self, ss, vectors):
        check(self.L.xa_load_vectors(self.h, C.byref(ss), C.byref(vectors)), "loadDAEVectors", self.L)

    def loadDAEMatrices(self, ss, vectors):
        check(self.L.xa_load_matrices(self.h, C.byref(ss), C.byref(vectors)), "loadDAEMatrices", self.L)

    def loadAll(self, ss, vectors):
        check(self.L.xa_load_all(self.h, C.byref(ss), C.byref(vectors)), "loadAll", self.L)

    def loadAllNoAlloc(self, ss, vectors):
        """xa_load_all_noalloc: loadAll with its scratch in the library's
        static device array (no allocation; groups up to its size)."""
        check(self.L.xa_load_all_noalloc(self.h, C.byref(ss), C.byref(vectors)),
              "loadAllNoAlloc", self.L)

    def plan_info(self, kind):
        """(targets, contributions, pattern slots per instance or 0) of the
        scatter plan of output kind `kind` (0 F, 1 Q, 2 dFdxdVp, 3 dQdxdVp,
        4 dFdx, 5 dQdx)."""
        nt, ns, ps = C.c_int64(), C.c_int64(), C.c_int()
        check(self.L.xa_group_plan_info(self.h, int(kind), C.byref(nt), C.byref(ns), C.byref(ps)),
              "plan_info", self.L)
        return nt.value, ns.value, ps.value

    def param_layout(self):
        """(per-instance columns read from HBM, columns served from the model
        cards) of the update kernel (xa_group_param_layout)."""
        a, b = C.c_int(), C.c_int()
        check(self.L.xa_group_param_layout(self.h, C.byref(a), C.byref(b)), "param_layout",
              self.L)
        return a.value, b.value

    PATHS = {"split": 1, "fused": 2, "pattern": 3}

    def set_path(self, path, block=256):
        """Pin xa_load_all's path ("split", "fused" or "pattern") and block
        size (xa_group_set_path); raises where the group has no such path."""
        check(self.L.xa_group_set_path(self.h, self.PATHS[path], int(block)), "set_path", self.L)

    def set_persistent(self, blocks=-1):
        """Run the stored-mode update as a persistent kernel of `blocks`
        resident workgroups looping over the instance tiles (-1: as many as
        fit the GPU; 0: one workgroup per tile) (xa_group_set_persistent)."""
        check(self.L.xa_group_set_persistent(self.h, int(blocks)), "set_persistent", self.L)

    def autotune(self, ss, vectors, reps=5):
        """Time the block shapes on these vectors and keep the fastest
        (xa_group_autotune); returns ("split" | "fused", block).  The tuning
        loads add into the output vectors: zero them afterwards."""
        route, block = C.c_int(), C.c_int()
        check(self.L.xa_group_autotune(self.h, C.byref(ss), C.byref(vectors), int(reps),
                                      C.byref(route), C.byref(block)), "autotune", self.L)
        return {1: "split", 2: "fused", 3: "pattern"}[route.value], block.value

    def failures(self):
        c = C.c_int64()
        check(self.L.xa_group_failures(self.h, C.byref(c)), "failures", self.L)
        return c.value

    def read(self, name):
        out = np.zeros(self.n, np.float64)
        check(self.L.xa_group_read_intermediate(self.h, name.encode(),
                                               out.ctypes.data_as(C.POINTER(C.c_double))), "read", self.L)
        return out

    def traffic_bytes(self):
        b = [C.c_double() for _ in range(3)]
        check(self.L.xa_group_traffic_bytes(self.h, *[C.byref(x) for x in b]), "traffic", self.L)
        return [x.value for x in b]

    def enable_timing(self, on=True):
        check(self.L.xa_group_enable_timing(self.h, 1 if on else 0), "timing", self.L)

    def last_times(self):
        t = [C.c_float() for _ in range(3)]
        check(self.L.xa_group_last_times(self.h, *[C.byref(x) for x in t]), "times", self.L)
        return [x.value for x in t]

    def close(self):
        if self.h:
            self.L.xa_group_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
