"""Argument checks of the LibTorch operators (torch_ext/ops.cpp) that need no
GPU: tensors of the wrong kind are rejected before any context or kernel is
touched (a host tensor handed to a kernel would be a device memory fault)."""
import pytest
import torch


@pytest.fixture(scope="module")
def ops():
    from proxpde_amd import autograd as A
    A.load()
    return torch.ops.proxpde


def test_kappa_apply_rejects_host_tensors(ops):
    k = torch.ones(5, dtype=torch.float64)
    x = torch.ones(5, 2, dtype=torch.float64)
    with pytest.raises(RuntimeError, match="GPU tensor"):
        ops.kappa_apply(k, x, 0, 0, False)


def test_kappa_apply_rejects_float32(ops):
    k = torch.ones(5, dtype=torch.float32)
    with pytest.raises(RuntimeError, match="float64"):
        ops.kappa_apply(k, k, 0, 0, False)


def test_kappa_solve_rejects_host_tensors(ops):
    k = torch.ones(5, dtype=torch.float64)
    with pytest.raises(RuntimeError, match="GPU tensor"):
        ops.kappa_solve(k, k, [0], [], 0, 0, 1e-8, 0, 10, False, False)
