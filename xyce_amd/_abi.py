"""ctypes declarations of the C ABI in include/xyce_amd.h.

The shared library ``libxyce_amd.so`` is built in-tree (``xyce_amd/``) by
``__graft_entry__.build()``; importing this module loads it and fails loudly
when it is missing -- there is no CPU fallback for the device loads.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libxyce_amd.so")

XA_OK = 0
XA_ERR_ARG, XA_ERR_STATE, XA_ERR_HIP, XA_ERR_UNSUPPORTED, XA_ERR_NOMEM, XA_ERR_DEVICE = range(-1, -7, -1)
XA_DIODE, XA_MOSFET1, XA_BJT, XA_BSIM4, XA_RESISTOR, XA_CAPACITOR, XA_VSOURCE = range(1, 8)
XA_STEP_DCOP, XA_STEP_BE, XA_STEP_TRAP = range(3)   # xa_batch_newton_step_ex methods
MODEL_TYPES = {"diode": XA_DIODE, "mosfet1": XA_MOSFET1, "bjt": XA_BJT, "bsim4": XA_BSIM4,
               "resistor": XA_RESISTOR, "capacitor": XA_CAPACITOR, "vsource": XA_VSOURCE}


class SolverState(C.Structure):
    """xa_solver_state: N_DEV_SolverState.h:93-190 subset."""
    _fields_ = [(n, C.c_int32) for n in (
        "newtonIter", "dcopFlag", "initJctFlag", "initFixFlag", "inputOPFlag",
        "locaEnabledFlag", "initTranFlag", "beginIntegrationFlag", "tranopFlag",
        "acopFlag", "transientFlag", "artParameterFlag", "bjtArtParameterFlag", "_pad")] + \
        [(n, C.c_double) for n in (
            "pdt", "currTimeStep", "lastTimeStep", "currTime", "gainScale", "nltermScale",
            "bpTol", "finalTime", "startingTimeStep")] + \
        [(n, C.c_int32) for n in ("HBspecified", "_pad2")]


class DeviceOptions(C.Structure):
    """xa_device_options: N_DEV_DeviceOptions.C:66-140 subset."""
    _fields_ = [(n, C.c_double) for n in (
        "gmin", "tnom", "temp", "defl", "defw", "defad", "defas",
        "excessPhaseScalar1", "excessPhaseScalar2")] + \
        [(n, C.c_int32) for n in (
            "voltageLimiterFlag", "newMeyerFlag", "newExcessPhase", "lambertWFlag")]


VECTOR_FIELDS = ("nextSol", "currSol", "flagSol", "nextSta", "currSta", "lastSta",
                 "nextStaDeriv", "nextSto", "currSto", "f", "q", "b", "dFdxdVp",
                 "dQdxdVp", "storeLeadF", "storeLeadQ", "dFdx", "dQdx")


class Vectors(C.Structure):
    """xa_vectors: device pointers, names after N_DEV_ExternData.h."""
    _fields_ = [(n, C.c_void_p) for n in VECTOR_FIELDS]


def _proto(lib):
    vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
    P = C.POINTER
    sig = {
        "xa_ctx_create": (C.c_int, [C.c_int, P(vp)]),
        "xa_ctx_destroy": (C.c_int, [vp]),
        "xa_ctx_sync": (C.c_int, [vp]),
        "xa_ctx_set_stream": (C.c_int, [vp, vp]),
        "xa_ctx_stream": (vp, [vp]),
        "xa_device_alloc": (C.c_int, [vp, C.c_size_t, P(vp)]),
        "xa_device_free": (C.c_int, [vp, vp]),
        "xa_copy_to_device": (C.c_int, [vp, vp, vp, C.c_size_t]),
        "xa_copy_to_host": (C.c_int, [vp, vp, vp, C.c_size_t]),
        "xa_last_error": (C.c_char_p, []),
        "xa_abi_version": (C.c_int, []),
        "xa_default_options": (None, [P(DeviceOptions)]),
        "xa_default_solver_state": (None, [P(SolverState)]),
        "xa_group_create": (C.c_int, [vp, C.c_int, P(DeviceOptions), P(vp)]),
        "xa_group_destroy": (C.c_int, [vp]),
        "xa_group_add_model": (C.c_int, [vp, C.c_int, C.c_int, P(C.c_char_p),
                                          P(C.c_double), P(C.c_int)]),
        "xa_group_add_instances": (C.c_int, [vp, C.c_int, P(C.c_int), C.c_int,
                                              P(C.c_char_p), P(P(C.c_double))]),
        "xa_group_request_lead_current": (C.c_int, [vp, P(i32)]),
        "xa_group_setup": (C.c_int, [vp]),
        "xa_group_counts": (C.c_int, [vp, P(C.c_int), P(C.c_int), P(C.c_int), P(C.c_int)]),
        "xa_group_num_instances": (C.c_int, [vp]),
        "xa_group_register_lids": (C.c_int, [vp, P(i32), P(i32), P(i32)]),
        "xa_group_jac_pairs": (C.c_int, [vp, P(i64), P(i32), P(i32)]),
        "xa_group_register_csr": (C.c_int, [vp, C.c_int, P(i32), P(i32)]),
        "xa_group_finalize": (C.c_int, [vp]),
        "xa_update_state": (C.c_int, [vp, P(SolverState), P(Vectors)]),
        "xa_update_secondary_state": (C.c_int, [vp, P(SolverState), P(Vectors)]),
        "xa_load_vectors": (C.c_int, [vp, P(SolverState), P(Vectors)]),
        "xa_load_matrices": (C.c_int, [vp, P(SolverState), P(Vectors)]),
        "xa_load_all": (C.c_int, [vp, P(SolverState), P(Vectors)]),
        "xa_load_all_noalloc": (C.c_int, [vp, P(SolverState), P(Vectors)]),
        "xa_group_failures": (C.c_int, [vp, P(i64)]),
        "xa_group_plan_info": (C.c_int, [vp, C.c_int, P(i64), P(i64), P(C.c_int)]),
        "xa_group_param_layout": (C.c_int, [vp, P(C.c_int), P(C.c_int)]),
        "xa_group_set_path": (C.c_int, [vp, C.c_int, C.c_int]),
        "xa_group_set_persistent": (C.c_int, [vp, C.c_int]),
        "xa_group_autotune": (C.c_int, [vp, P(SolverState), P(Vectors), C.c_int, P(C.c_int),
                                        P(C.c_int)]),
        "xa_group_read_intermediate": (C.c_int, [vp, C.c_char_p, P(C.c_double)]),
        "xa_group_traffic_bytes": (C.c_int, [vp, P(C.c_double), P(C.c_double), P(C.c_double)]),
        "xa_group_enable_timing": (C.c_int, [vp, C.c_int]),
        "xa_group_last_times": (C.c_int, [vp, P(C.c_float), P(C.c_float), P(C.c_float)]),
        "xa_batch_newton_step": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int] + [vp] * 13 +
                                 [C.c_double, C.c_int, vp, vp, vp]),
        "xa_batch_newton_step_ex": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int] +
                                    [vp] * 14 + [C.c_double, C.c_int, vp, vp, vp]),
        "xa_lu_f32_factor": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, vp, vp, vp]),
        "xa_lu_f32_solve": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, vp]),
        "xa_hb_plan_create": (C.c_int, [vp, C.c_int, C.c_int, C.c_double, P(vp)]),
        "xa_hb_plan_destroy": (C.c_int, [vp]),
        "xa_hb_ift": (C.c_int, [vp, vp, vp]),
        "xa_hb_dft": (C.c_int, [vp, vp, vp]),
        "xa_hb_add_jwq": (C.c_int, [vp, vp, vp]),
        "xa_hb_load": (C.c_int, [vp, vp, P(SolverState), vp, P(Vectors), C.c_int64, vp, vp, vp,
                                 vp]),
        "xa_hb_precond_create": (C.c_int, [vp, C.c_int, vp, vp, P(vp)]),
        "xa_hb_precond_destroy": (C.c_int, [vp]),
        "xa_hb_precond_factor": (C.c_int, [vp, vp, vp, vp]),
        "xa_hb_precond_solve": (C.c_int, [vp, vp, vp, C.c_int, C.c_double, vp, vp, vp]),
        "xa_csr_spmv": (C.c_int, [vp, C.c_int, vp, vp, vp, vp, vp]),
        "xa_csr_spmv_long": (C.c_int, [vp, C.c_int, vp, vp, vp, vp, vp]),
        "xa_ilu0_create": (C.c_int, [vp, C.c_int, C.c_int, vp, vp, P(vp)]),
        "xa_ilu0_destroy": (C.c_int, [vp]),
        "xa_ilu0_factor": (C.c_int, [vp, vp, P(C.c_int)]),
        "xa_ilu0_solve": (C.c_int, [vp, vp, vp]),
        "xa_ilu0_config": (C.c_int, [vp, C.c_int, P(C.c_int), P(C.c_int)]),
        "xa_netlist_parse": (C.c_int, [C.c_char_p, P(vp)]),
        "xa_netlist_destroy": (C.c_int, [vp]),
        "xa_netlist_sizes": (C.c_int, [vp, vp, vp, vp, vp]),
        "xa_netlist_node_name": (C.c_int, [vp, C.c_int, C.c_char_p, C.c_int]),
        "xa_netlist_tran": (C.c_int, [vp, vp]),
        "xa_netlist_ic": (C.c_int, [vp, vp, vp]),
        "xa_netlist_part": (C.c_int, [vp, C.c_int, vp, vp, vp, vp, vp]),
        "xa_netlist_part_model": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, vp, vp,
                                            P(C.c_char_p), vp]),
        "xa_netlist_part_instances": (C.c_int, [vp, C.c_int, C.c_int, vp, vp, P(C.c_char_p),
                                                vp]),
        "xa_circuit_build": (C.c_int, [vp, vp, P(DeviceOptions), C.c_int, P(vp)]),
        "xa_circuit_destroy": (C.c_int, [vp]),
        "xa_circuit_sizes": (C.c_int, [vp, vp, vp, vp, vp, vp, vp]),
        "xa_circuit_csr": (C.c_int, [vp, vp, vp]),
        "xa_circuit_group": (C.c_int, [vp, C.c_int, P(vp), vp]),
        "xa_circuit_load": (C.c_int, [vp, P(SolverState), P(Vectors)]),
        "xa_circuit_group_lids": (C.c_int, [vp, C.c_int, P(C.c_int64), vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


# Library variants: "" the production build; "pm" the parity build, whose
# evaluation takes exp/log/pow/tan from xa_pmath.h (the functions of the
# oracle's parity build) instead of the GPU math library.  Both may be loaded
# in one process; each has its own kernels, state and error message.
VARIANT_PATHS = {"": LIB_PATH, "pm": os.path.join(_HERE, "libxyce_amd_pm.so")}
_libs = {}


def lib(variant=""):
    """Load libxyce_amd.so (or a variant; raises if it has not been built)."""
    L = _libs.get(variant)
    if L is None:
        path = VARIANT_PATHS[variant]
        if not os.path.exists(path):
            raise RuntimeError(
                f"{path} is missing: build it with __graft_entry__.build(); "
                "the device loads have no CPU fallback")
        L = _libs[variant] = _proto(C.CDLL(path))
    return L


class XaError(RuntimeError):
    pass


def check(rc, what="", L=None):
    if rc != XA_OK:
        msg = (L or lib()).xa_last_error().decode()
        raise XaError(f"{what}: rc={rc}: {msg}")
    return rc
