"""thmm_log_partition_device_static: the workspace-free log partition whose
contraction state lives in the library's static per-GPU array (no device
allocation in the call), checked through the C-ABI against the workspace
entry point thmm_log_partition_device on the same device inputs:

* two calls in a row, the second on a smaller plan and batch (the array
  keeps the first call's contents), K = 16 FP32 within the FP32 tolerance
  (1e-5 relative on log Z) and against the FP64 oracle;
* K = 4 FP64, where both entry points run the same plan-level kernels:
  bit-equal;
* a shape whose state exceeds the array's maximum is a length_error, and
  the call after it still agrees.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
import parallel_tree_hmm_amd as thmm
from parallel_tree_hmm_amd import _abi
from conftest import forest, path, random_recursive, star

pytestmark = pytest.mark.gpu


def inputs(torch, plan, k, batch, seed, dtype):
    gen = torch.Generator(device="cuda").manual_seed(seed)
    nodes = torch.exp(torch.randn((batch, plan.num_nodes, k), device="cuda", generator=gen, dtype=torch.float64))
    rows = -torch.log1p(-torch.rand((plan.num_edges, k, k), device="cuda", generator=gen, dtype=torch.float64))
    edges = rows / rows.sum(2, keepdim=True)
    return nodes.to(dtype), edges.to(dtype)


def workspace_log_z(torch, plan, k, batch, precision, nodes, edges):
    ws = thmm.Workspace()
    ws.reserve(plan, k, batch, precision=precision, flags=_abi.THMM_RESERVE_DEVICE_RESIDENT)
    out = torch.empty(batch * plan.roots.size, dtype=torch.float64, device="cuda")
    ws.log_partition_device(nodes.data_ptr(), edges.data_ptr(), batch, out.data_ptr(), 0)
    torch.cuda.synchronize()
    ws.close()
    return out.cpu().numpy()


def static_log_z(torch, plan, k, batch, precision, nodes, edges):
    out = torch.full((batch * plan.roots.size,), float("nan"), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    thmm.log_partition_device_static(plan, k, batch, nodes.data_ptr(), edges.data_ptr(), out.data_ptr(), precision,
                                     stream)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_static_entry_point_matches_the_workspace_entry_point(require_gpu):
    import torch

    rng = np.random.default_rng(77)
    big = thmm.Plan(forest([random_recursive(1500, rng) for _ in range(3)] + [path(800), star(300)]))
    small = thmm.Plan(random_recursive(300, rng))
    k = 16
    for call, (plan, batch) in enumerate(((big, 6), (small, 2))):
        need = thmm.static_scratch_bytes(plan, k, batch, "f32")
        assert need <= _abi.THMM_STATIC_SCRATCH_BYTES
        nodes, edges = inputs(torch, plan, k, batch, 100 + call, torch.float32)
        got = static_log_z(torch, plan, k, batch, "f32", nodes, edges)
        want = workspace_log_z(torch, plan, k, batch, "f32", nodes, edges)
        assert np.all(np.isfinite(got))
        assert np.max(np.abs(got - want) / np.maximum(1.0, np.abs(want))) <= 1e-5, call
        olz, _, _ = oracle.marginals_batched(plan, k, nodes.double().cpu().numpy(), edges.double().cpu().numpy(),
                                             threads=4)
        assert np.max(np.abs(got - olz.reshape(-1)) / np.maximum(1.0, np.abs(olz.reshape(-1)))) <= 1e-5, call
        print(f"call {call}: {plan.num_nodes} nodes x {batch}, {need} bytes of state", flush=True)


def test_static_entry_point_is_bit_equal_on_the_plan_levels(require_gpu):
    import torch

    rng = np.random.default_rng(78)
    for call, (plan, batch) in enumerate(((thmm.Plan(random_recursive(1023, rng)), 16),
                                          (thmm.Plan(forest([path(200), random_recursive(100, rng)])), 3))):
        nodes, edges = inputs(torch, plan, 4, batch, 200 + call, torch.float64)
        got = static_log_z(torch, plan, 4, batch, "f64", nodes, edges)
        want = workspace_log_z(torch, plan, 4, batch, "f64", nodes, edges)
        assert np.array_equal(got, want), call


def test_static_entry_point_rejects_shapes_beyond_its_maximum(require_gpu):
    import torch

    plan = thmm.Plan(path(20000))
    k, batch = 64, 1
    assert thmm.static_scratch_bytes(plan, k, batch, "f32") > _abi.THMM_STATIC_SCRATCH_BYTES
    nodes = torch.ones((batch, plan.num_nodes, k), device="cuda")
    edges = torch.ones((plan.num_edges, k, k), device="cuda")
    out = torch.empty(batch, dtype=torch.float64, device="cuda")
    with pytest.raises(thmm.LengthError, match="static maximum"):
        thmm.log_partition_device_static(plan, k, batch, nodes.data_ptr(), edges.data_ptr(), out.data_ptr(), "f32")
    with pytest.raises(thmm.InvalidArgument):
        thmm.log_partition_device_static(plan, 0, batch, nodes.data_ptr(), edges.data_ptr(), out.data_ptr(), "f32")
    # the array still serves a call that fits
    small = thmm.Plan(path(50))
    n2, e2 = inputs(torch, small, 8, 2, 300, torch.float32)
    got = static_log_z(torch, small, 8, 2, "f32", n2, e2)
    want = workspace_log_z(torch, small, 8, 2, "f32", n2, e2)
    assert np.max(np.abs(got - want) / np.maximum(1.0, np.abs(want))) <= 1e-5
