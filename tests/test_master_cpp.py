"""The C++ drop-in (include/xyce_amd_master.hpp): a host-vector Master with
the reference's four Master signatures over the C ABI, run by the compiled
test program tests/cpp/master_parity against the oracle for every model
type (host vectors copied in and out, outputs pre-filled with other device
types' contributions; bar in the program's header)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "master_parity")


def _binary():
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is not built (run __graft_entry__.build())")
    return BIN


def _gpu_visible():
    import torch
    return torch.cuda.device_count() > 0


def test_master_without_gpu_fails_loudly():
    """No GPU: the Master's constructor throws with the library's message."""
    if _gpu_visible():
        pytest.skip("a GPU is visible")
    r = subprocess.run([_binary(), "diode"], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "xa_ctx_create" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("model", ["diode", "mosfet1", "bjt", "bsim4"])
def test_master_host_vectors_match_oracle(model):
    print(f"master_parity {model}", flush=True)
    r = subprocess.run([_binary(), model], capture_output=True, text=True, timeout=300)
    print(r.stdout[-4000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stdout[-2000:]
    assert f"{model}: PASS" in r.stdout
