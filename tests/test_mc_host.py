"""Host logic of the batched driver (no GPU): a non-finite update stops the
Newton loop instead of passing as converged."""
import numpy as np
import pytest
import torch

from xyce_amd.mc import BatchTransient


def _driver(worst):
    bt = BatchTransient.__new__(BatchTransient)
    bt.use_graph = False
    bt.newton_iterations = 0
    bt._iteration = lambda flags, h, dc: torch.tensor(worst, dtype=torch.float64)
    return bt


def test_nan_update_raises():
    with pytest.raises(RuntimeError, match="not finite"):
        _driver(float("nan"))._newton(lambda it: {}, 1.0, True, 1e-9, 5)


def test_converged_update_returns():
    assert _driver(0.0)._newton(lambda it: {}, 1.0, True, 1e-9, 5) == 2


def test_nonconvergence_raises():
    with pytest.raises(RuntimeError, match="did not converge"):
        _driver(1.0)._newton(lambda it: {}, 1.0, True, 1e-9, 3)
