"""The torch.autograd harness over the forward/backward C-ABI pair: gradients
of log Z equal the oracle's posterior marginals and central finite
differences of log Z."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
import parallel_tree_hmm_amd as thmm
from conftest import caterpillar, random_recursive

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype_name", ["float64", "float32"])
def test_autograd_matches_marginals_and_finite_differences(require_gpu, dtype_name):
    import torch

    from parallel_tree_hmm_amd.autograd import log_partition, reserve_for

    dtype = getattr(torch, dtype_name)
    rng = np.random.default_rng(3)
    plan = thmm.Plan(caterpillar(31))
    k, batch = 5, 2
    log_nodes = torch.tensor(rng.normal(size=(batch, plan.num_nodes, k)), dtype=dtype, device="cuda", requires_grad=True)
    log_edges = torch.tensor(rng.normal(size=(plan.num_edges, k, k)), dtype=dtype, device="cuda", requires_grad=True)
    ws = reserve_for(plan, k, batch, dtype)
    weights = torch.tensor([[0.7], [-1.3]], dtype=torch.float64, device="cuda")
    loss = (log_partition(log_nodes, log_edges, ws) * weights).sum()
    loss.backward()

    nodes = np.exp(log_nodes.detach().double().cpu().numpy())
    edges = np.exp(log_edges.detach().double().cpu().numpy())
    _, nm, em = oracle.marginals_batched(plan, k, nodes, edges)
    w = weights.cpu().numpy()[:, 0]
    tol = 1e-11 if dtype_name == "float64" else 2e-5
    assert np.max(np.abs(log_nodes.grad.double().cpu().numpy() - w[:, None, None] * nm)) <= tol
    assert np.max(np.abs(log_edges.grad.double().cpu().numpy() - np.einsum("b,bekl->ekl", w, em))) <= 2 * tol

    if dtype_name == "float64":
        h = 1e-6

        def objective(ln, le):
            with torch.no_grad():
                return float((log_partition(ln, le, ws) * weights).sum())

        for _ in range(4):
            b, i, s = int(rng.integers(batch)), int(rng.integers(plan.num_nodes)), int(rng.integers(k))
            up, dn = log_nodes.detach().clone(), log_nodes.detach().clone()
            up[b, i, s] += h
            dn[b, i, s] -= h
            fd = (objective(up, log_edges.detach()) - objective(dn, log_edges.detach())) / (2 * h)
            assert abs(fd - float(log_nodes.grad[b, i, s])) < 1e-7
            e, p, c = int(rng.integers(plan.num_edges)), int(rng.integers(k)), int(rng.integers(k))
            up, dn = log_edges.detach().clone(), log_edges.detach().clone()
            up[e, p, c] += h
            dn[e, p, c] -= h
            fd = (objective(log_nodes.detach(), up) - objective(log_nodes.detach(), dn)) / (2 * h)
            assert abs(fd - float(log_edges.grad[e, p, c])) < 1e-7


def test_autograd_forest_roots(require_gpu):
    import torch

    from conftest import forest
    from parallel_tree_hmm_amd.autograd import log_partition, reserve_for

    rng = np.random.default_rng(4)
    plan = thmm.Plan(forest([random_recursive(20, rng), caterpillar(15), random_recursive(9, rng)]))
    k = 3
    log_nodes = torch.tensor(rng.normal(size=(1, plan.num_nodes, k)), device="cuda", requires_grad=True)
    log_edges = torch.tensor(rng.normal(size=(plan.num_edges, k, k)), device="cuda", requires_grad=True)
    ws = reserve_for(plan, k, 1, torch.float64)
    log_z = log_partition(log_nodes, log_edges, ws)
    assert log_z.shape == (1, 3)
    log_z.sum().backward()
    nodes = np.exp(log_nodes.detach().cpu().numpy()[0])
    edges = np.exp(log_edges.detach().cpu().numpy())
    olz, nm, em = oracle.marginals(plan, k, nodes, edges)
    assert np.allclose(log_z.detach().cpu().numpy()[0], olz, rtol=0, atol=1e-11)
    assert np.allclose(log_nodes.grad.cpu().numpy()[0], nm, atol=1e-11)
    assert np.allclose(log_edges.grad.cpu().numpy(), em, atol=1e-11)


def test_autograd_keeps_log_edges_for_backward(require_gpu):
    """K = 16: the backward pass reads the log edges again (the rake kernels
    take them in place), so the wrapper keeps the tensor it passed to the
    forward call alive, here the contiguous copy of a transposed view, and
    an in-place change of it before backward is refused."""
    import torch

    from conftest import forest
    from parallel_tree_hmm_amd.autograd import log_partition, reserve_for

    rng = np.random.default_rng(5)
    plan = thmm.Plan(forest([caterpillar(61), random_recursive(80, rng)]))
    k = 16
    log_nodes = torch.tensor(rng.normal(size=(1, plan.num_nodes, k)), dtype=torch.float32, device="cuda",
                             requires_grad=True)
    base = torch.tensor(rng.normal(size=(plan.num_edges, k, k)), dtype=torch.float32, device="cuda",
                        requires_grad=True)
    ws = reserve_for(plan, k, 1, torch.float32)
    view = base.transpose(1, 2)  # not contiguous: the wrapper passes a copy
    log_z = log_partition(log_nodes, view, ws)
    churn = [torch.full_like(base, 123.0) for _ in range(4)]  # reuses freed blocks, if any
    log_z.sum().backward()
    del churn
    edges = np.exp(view.detach().double().cpu().numpy())
    olz, nm, em = oracle.marginals(plan, k, np.exp(log_nodes.detach().double().cpu().numpy()[0]), edges)
    assert np.max(np.abs(log_z.detach().cpu().numpy()[0] - olz) / np.maximum(1.0, np.abs(olz))) <= 1e-5
    assert np.max(np.abs(log_nodes.grad.double().cpu().numpy()[0] - nm)) <= 2e-5
    assert np.max(np.abs(base.grad.transpose(1, 2).double().cpu().numpy() - em)) <= 4e-5

    shifted = base.detach().clone().requires_grad_()
    inner = shifted * 1.0
    again = log_partition(log_nodes.detach(), inner, ws)
    inner.add_(1.0)
    with pytest.raises(RuntimeError):
        again.sum().backward()
