"""CPU checks of the C ABI: the library loads and exports every entry point
declared in include/xyce_amd.h; no GPU compute is called."""
import ctypes as C
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "xyce_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(xa_[a-z_0-9]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    import xyce_amd
    lib = xyce_amd.lib()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert len(_declared()) >= 30


def test_abi_version_and_defaults():
    import xyce_amd
    assert xyce_amd.lib().xa_abi_version() == 2
    o = xyce_amd.default_options()
    assert o.gmin == 1e-12 and o.voltageLimiterFlag == 1 and abs(o.tnom - 300.15) < 1e-12
    s = xyce_amd.default_solver_state()
    assert s.newtonIter == 0 and s.gainScale == 1.0


def test_struct_layouts_match_header():
    import xyce_amd
    assert C.sizeof(xyce_amd.SolverState) == 14 * 4 + 9 * 8 + 2 * 4
    assert C.sizeof(xyce_amd.DeviceOptions) == 9 * 8 + 4 * 4
    assert C.sizeof(xyce_amd.Vectors) == 18 * 8


def test_errors_without_gpu_are_reported_not_raised_in_c():
    """Context creation without a visible GPU returns an error code."""
    import torch
    import xyce_amd
    if torch.cuda.is_available():
        return
    h = C.c_void_p()
    rc = xyce_amd.lib().xa_ctx_create(0, C.byref(h))
    assert rc != 0
    assert xyce_amd.lib().xa_last_error()
