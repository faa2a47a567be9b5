"""CPU oracle for the device-load path -- TEST INFRASTRUCTURE ONLY.

Used by tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py as the checker; never imported by the product package.  Wraps
``oracle/liboracle.so`` (an FP64 C restatement of Xyce 6.2's device loads,
built by ``make -C oracle``) with the same call shape as the product ABI but
host numpy arrays.

Parity status per function is stated in the header of each C file and in the
tests that pin it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
# "pm": the parity build, whose evaluation takes exp/log/pow/tan from
# xyce_amd/csrc/xa_pmath.h (the functions of the GPU library's parity build)
VARIANT_PATHS = {"": LIB_PATH, "pm": os.path.join(_HERE, "liboracle_pm.so")}
MODEL_TYPES = {"diode": 1, "mosfet1": 2, "bjt": 3, "bsim4": 4, "resistor": 5, "capacitor": 6,
               "vsource": 7}


class SolverState(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "newtonIter", "dcopFlag", "initJctFlag", "initFixFlag", "inputOPFlag",
        "locaEnabledFlag", "initTranFlag", "beginIntegrationFlag", "tranopFlag",
        "acopFlag", "transientFlag", "artParameterFlag", "bjtArtParameterFlag", "_pad")] + \
        [(n, C.c_double) for n in (
            "pdt", "currTimeStep", "lastTimeStep", "currTime", "gainScale", "nltermScale",
            "bpTol", "finalTime", "startingTimeStep")] + \
        [(n, C.c_int32) for n in ("HBspecified", "_pad2")]


class DeviceOptions(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "gmin", "tnom", "temp", "defl", "defw", "defad", "defas",
        "excessPhaseScalar1", "excessPhaseScalar2")] + \
        [(n, C.c_int32) for n in (
            "voltageLimiterFlag", "newMeyerFlag", "newExcessPhase", "lambertWFlag")]


VECTOR_FIELDS = ("nextSol", "currSol", "flagSol", "nextSta", "currSta", "lastSta",
                 "nextStaDeriv", "nextSto", "currSto", "f", "q", "b", "dFdxdVp",
                 "dQdxdVp", "storeLeadF", "storeLeadQ", "dFdx", "dQdx")


class Vectors(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in VECTOR_FIELDS]


_libs = {}


def lib(variant=""):
    L = _libs.get(variant)
    if L is None:
        path = VARIANT_PATHS[variant]
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run make -C oracle")
        L = C.CDLL(path)
        vp = C.c_void_p
        P = C.POINTER
        for name, res, args in [
            ("or_group_create", C.c_int, [C.c_int, P(DeviceOptions), P(vp)]),
            ("or_group_destroy", C.c_int, [vp]),
            ("or_group_add_model", C.c_int, [vp, C.c_int, C.c_int, P(C.c_char_p), P(C.c_double), P(C.c_int)]),
            ("or_group_add_instances", C.c_int, [vp, C.c_int, P(C.c_int), C.c_int, P(C.c_char_p), P(P(C.c_double))]),
            ("or_group_request_lead_current", C.c_int, [vp, P(C.c_int32)]),
            ("or_group_setup", C.c_int, [vp]),
            ("or_group_counts", C.c_int, [vp, P(C.c_int), P(C.c_int), P(C.c_int), P(C.c_int)]),
            ("or_group_register_lids", C.c_int, [vp, P(C.c_int32), P(C.c_int32), P(C.c_int32)]),
            ("or_group_jac_pairs", C.c_int, [vp, P(C.c_int64), P(C.c_int32), P(C.c_int32)]),
            ("or_group_register_csr", C.c_int, [vp, C.c_int, P(C.c_int32), P(C.c_int32)]),
            ("or_group_finalize", C.c_int, [vp]),
            ("or_update_state", C.c_int, [vp, P(SolverState), P(Vectors)]),
            ("or_update_secondary_state", C.c_int, [vp, P(SolverState), P(Vectors)]),
            ("or_load_vectors", C.c_int, [vp, P(SolverState), P(Vectors)]),
            ("or_load_matrices", C.c_int, [vp, P(SolverState), P(Vectors)]),
            ("or_read_intermediate", C.c_int, [vp, C.c_char_p, P(C.c_double)]),
            ("or_last_error", C.c_char_p, []),
            ("or_set_threads", None, [C.c_int]),
            ("or_math", C.c_double, [C.c_int, C.c_double, C.c_double]),
            ("or_default_options", None, [P(DeviceOptions)]),
            ("or_set_magnitude_mode", None, [C.c_int]),
            ("or_set_libm_perturbation", None, [C.c_int, C.c_uint64]),
            ("or_group_save_state", C.c_int, [vp]),
            ("or_group_restore_state", C.c_int, [vp]),
        ]:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _libs[variant] = L
    return L


def set_threads(n):
    """OpenMP threads for the updateState loops (1: the serial order)."""
    lib().or_set_threads(int(n))


def check(rc, what="", L=None):
    if rc != 0:
        raise RuntimeError(f"oracle {what}: rc={rc}: {(L or lib()).or_last_error().decode()}")


def default_options(**kw):
    o = DeviceOptions()
    lib().or_default_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _cstrs(names):
    arr = (C.c_char_p * max(len(names), 1))()
    for i, n in enumerate(names):
        arr[i] = n.encode()
    return arr


def _p(a, t=C.c_double):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def make_vectors(**arrays):
    v = Vectors()
    for k, a in arrays.items():
        if a is not None:
            assert a.dtype == np.float64 and a.flags.c_contiguous
            setattr(v, k, a.ctypes.data)
    return v


class Master:
    """Oracle group with the product Master's call shape (host arrays)."""

    def __init__(self, model_type, options=None, variant=""):
        self.L = lib(variant)
        self.type = MODEL_TYPES[model_type] if isinstance(model_type, str) else model_type
        self.h = C.c_void_p()
        opt = options if options is not None else default_options()
        check(self.L.or_group_create(self.type, C.byref(opt), C.byref(self.h)), "create", self.L)
        self._keep = []
        self._n = 0

    def add_model(self, params, level=1):
        names = list(params.keys())
        vals = (C.c_double * max(len(names), 1))(*[float(params[k]) for k in names])
        idx = C.c_int()
        check(self.L.or_group_add_model(self.h, level, len(names), _cstrs(names), vals, C.byref(idx)),
              "add_model", self.L)
        return idx.value

    def add_instances(self, n, model_index=None, params=None):
        params = params or {}
        names = list(params.keys())
        cols = [np.ascontiguousarray(np.broadcast_to(np.asarray(params[k], np.float64), (n,)))
                for k in names]
        self._keep.extend(cols)
        ptrs = (C.POINTER(C.c_double) * max(len(cols), 1))(*[_p(c) for c in cols])
        mi = None
        if model_index is not None:
            mi_arr = np.ascontiguousarray(np.broadcast_to(np.asarray(model_index, np.int32), (n,)))
            self._keep.append(mi_arr)
            mi = _p(mi_arr, C.c_int)
        check(self.L.or_group_add_instances(self.h, n, mi, len(names), _cstrs(names), ptrs),
              "add_instances", self.L)
        self._n += n

    def request_lead_current(self, flags):
        f = np.ascontiguousarray(flags, np.int32)
        check(self.L.or_group_request_lead_current(self.h, _p(f, C.c_int32)), "lead", self.L)

    def setup(self):
        check(self.L.or_group_setup(self.h), "setup", self.L)

    @property
    def n(self):
        return self._n

    def counts(self):
        out = [np.zeros(self._n, np.int32) for _ in range(4)]
        check(self.L.or_group_counts(self.h, *[_p(o, C.c_int) for o in out]), "counts", self.L)
        return out

    def register_lids(self, node_lids, sta_lids, sto_lids):
        arrs = [np.ascontiguousarray(a, np.int32) for a in (node_lids, sta_lids, sto_lids)]
        check(self.L.or_group_register_lids(self.h, *[_p(a, C.c_int32) for a in arrs]), "lids", self.L)

    def jac_pairs(self):
        cnt = C.c_int64()
        check(self.L.or_group_jac_pairs(self.h, C.byref(cnt), None, None), "jac_pairs", self.L)
        r = np.zeros(cnt.value, np.int32)
        c = np.zeros(cnt.value, np.int32)
        check(self.L.or_group_jac_pairs(self.h, C.byref(cnt), _p(r, C.c_int32), _p(c, C.c_int32)),
              "jac_pairs", self.L)
        return r, c

    def register_csr(self, row_ptr, col_idx):
        self._rp = np.ascontiguousarray(row_ptr, np.int32)
        self._ci = np.ascontiguousarray(col_idx, np.int32)
        check(self.L.or_group_register_csr(self.h, len(self._rp) - 1, _p(self._rp, C.c_int32),
                                          _p(self._ci, C.c_int32)), "csr", self.L)

    def finalize(self):
        check(self.L.or_group_finalize(self.h), "finalize", self.L)

    def updateState(self, ss, v):
        check(self.L.or_update_state(self.h, C.byref(ss), C.byref(v)), "updateState", self.L)

    def updateSecondaryState(self, ss, v):
        check(self.L.or_update_secondary_state(self.h, C.byref(ss), C.byref(v)), "updateSecondary", self.L)

    def loadDAEVectors(self, ss, v):
        check(self.L.or_load_vectors(self.h, C.byref(ss), C.byref(v)), "loadDAEVectors", self.L)

    def loadDAEMatrices(self, ss, v):
        check(self.L.or_load_matrices(self.h, C.byref(ss), C.byref(v)), "loadDAEMatrices", self.L)

    def save_state(self):
        """Snapshot the instance records (values kept between calls)."""
        check(self.L.or_group_save_state(self.h), "save_state", self.L)

    def restore_state(self):
        check(self.L.or_group_restore_state(self.h), "restore_state", self.L)

    def read(self, name):
        out = np.zeros(self._n, np.float64)
        check(self.L.or_read_intermediate(self.h, name.encode(), _p(out)), "read", self.L)
        return out

    def close(self):
        if self.h:
            self.L.or_group_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
