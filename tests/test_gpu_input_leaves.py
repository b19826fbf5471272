"""Dense leaf rakes reading the caller's linear potentials (thmm_tuning
input_leaves = 1, the default): round 0's rakes on the fused forward program
(K = 8, 16, 32, 64) form their leaf vectors from the caller's node
potentials, round 0's absorptions (in the gather's direct chunks and in its
partial-sum levels) form their parents' vectors from them too, and the node
initialisation skips those leaves and parents. Every case
compares the default with input_leaves = 0 (every node initialised first)
on the same workspace and inputs, bit for bit: thmm_marginals_device (log Z,
node and edge marginals, eager and captured-graph replay) and the
host-staged marginal and partition calls. Forests mix
stars (many leaves of one parent), paths (no leaves but the last), recursive
trees and single nodes; zero potentials, and leaves without support (device
calls) and a round-0 parent without support; a star of 600 leaves whose
partials take two reduction levels; batch 1 and 3; FP32 and FP64. A node
array off 16-byte alignment takes the initialise-first path. MAP (K = 8,
16, 32) logs round 0's leaves and parents where it first reads them under
the same flag, and gives the same states and weights as logging every node
first (quad kernel for aligned inputs, one-value kernel otherwise)."""

from __future__ import annotations

import numpy as np
import pytest

import parallel_tree_hmm_amd as thmm
from conftest import forest, path, random_recursive, star

pytestmark = pytest.mark.gpu


def case(k, precision, batch, seed):
    rng = np.random.default_rng(seed)
    parents = forest([random_recursive(300, rng), star(70), path(40), np.array([-1]), random_recursive(90, rng),
                      np.array([-1, 0]), star(600)])
    plan = thmm.Plan(parents)
    dtype = np.float32 if precision == "f32" else np.float64
    nodes = np.exp(rng.normal(size=(batch, plan.num_nodes, k)) * 2.0)
    nodes[rng.random(nodes.shape) < 0.05] = 0.0
    rows = rng.exponential(size=(plan.num_edges, k, k))
    edges = rows / rows.sum(2, keepdims=True)
    return plan, nodes.astype(dtype), edges.astype(dtype)


def device_calls(torch, ws, plan, k, batch, nodes, edges, dtype):
    n, e, roots = plan.num_nodes, plan.num_edges, plan.roots.size
    dn = torch.from_numpy(nodes).cuda()
    de = torch.from_numpy(edges).cuda()
    lz = torch.empty(batch * roots, dtype=torch.float64, device="cuda")
    no = torch.empty((batch, n, k), dtype=dtype, device="cuda")
    eo = torch.empty((batch, e, k, k), dtype=dtype, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    ws.marginals_device(dn.data_ptr(), de.data_ptr(), batch, lz.data_ptr(), no.data_ptr(), eo.data_ptr(), stream)
    torch.cuda.synchronize()
    return lz.cpu().numpy(), no.cpu().numpy(), eo.cpu().numpy()


def same(a, b):
    return a.shape == b.shape and a.tobytes() == b.tobytes()


@pytest.mark.parametrize("precision", ["f32", "f64"])
@pytest.mark.parametrize("k", [8, 16, 32, 64])
@pytest.mark.parametrize("batch", [1, 3])
def test_input_leaves_equal_initialised_leaves(require_gpu, k, precision, batch):
    import torch

    plan, nodes, edges = case(k, precision, batch, 10 * k + batch)
    dtype = torch.float32 if precision == "f32" else torch.float64
    # A copy with leaves without support (log Z = -inf for their trees;
    # device calls only: the host-staged ones reject it).
    dead = nodes.copy()
    leaves = np.setdiff1d(plan.edge_children, plan.edge_parents)
    dead[0, leaves[:3], :] = 0.0
    dead[0, 300, :] = 0.0  # the centre of star(70), absorbed in round 0
    results = {}
    for flag in (1, 0):
        out = {}
        ws = thmm.Workspace()
        ws.reserve_marginals(plan, k, batch, precision=precision, device_resident=True)
        ws.set_tuning(input_leaves=flag)
        out["device"] = device_calls(torch, ws, plan, k, batch, nodes, edges, dtype)
        out["dead"] = device_calls(torch, ws, plan, k, batch, dead, edges, dtype)
        ws.close()
        wg = thmm.Workspace()
        wg.reserve_marginals(plan, k, batch, precision=precision, device_resident=True, graphs=True)
        wg.set_tuning(input_leaves=flag)
        out["graph"] = [device_calls(torch, wg, plan, k, batch, nodes, edges, dtype) for _ in range(2)]
        wg.close()
        wh = thmm.Workspace()
        wh.reserve_marginals(plan, k, batch, precision=precision)
        wh.set_tuning(input_leaves=flag)
        m = thmm.posterior_marginals_prepared(thmm.BatchedModelView(plan, k, batch, nodes, edges), wh)
        out["staged"] = (m.log_partitions.copy(), m.nodes.copy(), m.edges.copy())
        p = thmm.log_partition_function_prepared(thmm.BatchedModelView(plan, k, batch, nodes, edges), wh)
        out["partition"] = (p.values.copy(), wh.log_partitions_f64()[: batch * plan.roots.size].copy())
        wh.close()
        results[flag] = out
    assert np.isneginf(results[1]["dead"][0]).any()
    for name in ("device", "dead", "staged", "partition"):
        for a, b in zip(results[1][name], results[0][name]):
            assert same(a, b), f"K={k} {precision} batch {batch}: {name} differs"
    for replay in results[1]["graph"] + results[0]["graph"]:
        for a, b in zip(replay, results[1]["device"]):
            assert same(a, b), f"K={k} {precision} batch {batch}: graph replay differs"
    assert np.all(np.isfinite(results[1]["device"][0]))


def test_unaligned_inputs_fall_back(require_gpu):
    """A node array that is not 16-byte aligned takes the initialise-first
    path; its values equal the aligned call's."""
    import torch

    k, batch = 16, 1
    plan, nodes, edges = case(k, "f32", batch, 5)
    n, e, roots = plan.num_nodes, plan.num_edges, plan.roots.size
    flat = torch.empty(n * k + 1, dtype=torch.float32, device="cuda")
    flat[1:] = torch.from_numpy(nodes.reshape(-1)).cuda()
    shifted = flat[1:]  # 4-byte offset
    aligned = torch.from_numpy(nodes).cuda()
    de = torch.from_numpy(edges).cuda()
    stream = torch.cuda.current_stream().cuda_stream
    ws = thmm.Workspace()
    ws.reserve_marginals(plan, k, batch, device_resident=True)
    outs = []
    for dn in (aligned, shifted):
        lz = torch.empty(roots, dtype=torch.float64, device="cuda")
        no = torch.empty((1, n, k), device="cuda")
        eo = torch.empty((1, e, k, k), device="cuda")
        ws.marginals_device(dn.data_ptr(), de.data_ptr(), 1, lz.data_ptr(), no.data_ptr(), eo.data_ptr(), stream)
        torch.cuda.synchronize()
        outs.append((lz.cpu().numpy(), no.cpu().numpy(), eo.cpu().numpy()))
    ws.close()
    # The unaligned initialisation takes FP64 logs where the team transpose
    # takes FP32 ones, so the two agree to FP32 rounding, not bit for bit
    # (the centre of the 600-leaf star sums 600 such messages).
    assert np.allclose(outs[0][0], outs[1][0], rtol=1e-6, atol=1e-6)
    assert np.max(np.abs(outs[0][1] - outs[1][1])) <= 4e-6
    assert np.max(np.abs(outs[0][2] - outs[1][2])) <= 4e-6


@pytest.mark.parametrize("precision", ["f32", "f64"])
@pytest.mark.parametrize("k", [8, 16, 32])
@pytest.mark.parametrize("batch", [1, 3])
def test_map_log_transform_aligned_and_unaligned(require_gpu, k, precision, batch):
    """MAP's log transform of the node potentials: with 16-byte aligned
    inputs and input_leaves = 1 (the default) round 0's leaves are logged by
    the first rake level as it loads them and round 0's parents by their
    absorption level, and the quad kernel logs the other nodes from a list;
    with input_leaves = 0 the quad kernel logs every node, and unaligned
    inputs take the one-value kernel for every node. All log each value in
    FP64 the same way, so states and log weights are equal bit for bit, also
    with a leaf and a round-0 parent without support, and the states equal
    the oracle's."""
    import torch

    import oracle

    plan, nodes, edges = case(k, precision, batch, 7 * k + batch)
    n = plan.num_nodes
    dead = nodes.copy()
    leaves = np.setdiff1d(plan.edge_children, plan.edge_parents)
    dead[0, leaves[:2], :] = 0.0
    dead[0, 300, :] = 0.0  # the centre of star(70), absorbed in round 0
    dtype = torch.float32 if precision == "f32" else torch.float64
    de = torch.from_numpy(edges).cuda()
    for inputs in (nodes, dead):
        aligned = torch.from_numpy(inputs).cuda()
        storage = torch.empty(inputs.size + 1, dtype=dtype, device="cuda")
        shifted = storage[1:]
        shifted.copy_(aligned.reshape(-1))
        results = []
        for dn, flag in ((aligned, 1), (aligned, 0), (shifted, 1)):
            ws = thmm.Workspace()
            ws.reserve(plan, k, batch, precision=precision,
                       flags=thmm._abi.THMM_RESERVE_MAXIMUM | thmm._abi.THMM_RESERVE_DEVICE_RESIDENT)
            ws.set_tuning(input_leaves=flag)
            lw = torch.empty(batch * plan.roots.size, dtype=torch.float64, device="cuda")
            states = torch.empty(batch * n, dtype=torch.int32, device="cuda")
            ws.maximum_device(dn.data_ptr(), de.data_ptr(), batch, lw.data_ptr(), states.data_ptr(),
                              torch.cuda.current_stream().cuda_stream)
            torch.cuda.synchronize()
            results.append((lw.cpu().numpy(), states.cpu().numpy()))
            ws.close()
        for other in results[1:]:
            for a, b in zip(results[0], other):
                assert same(a, b), f"K={k} {precision} batch {batch}: MAP differs"
    assert np.isneginf(results[0][0]).any()
    lw0, expected = oracle.maximum(plan, k, nodes[0].astype(np.float64), edges.astype(np.float64))
    ws = thmm.Workspace()
    ws.reserve(plan, k, batch, precision=precision,
               flags=thmm._abi.THMM_RESERVE_MAXIMUM | thmm._abi.THMM_RESERVE_DEVICE_RESIDENT)
    lw = torch.empty(batch * plan.roots.size, dtype=torch.float64, device="cuda")
    states = torch.empty(batch * n, dtype=torch.int32, device="cuda")
    ws.maximum_device(torch.from_numpy(nodes).cuda().data_ptr(), de.data_ptr(), batch, lw.data_ptr(),
                      states.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    ws.close()
    assert np.array_equal(states.cpu().numpy().reshape(batch, n)[0].astype(np.uint32), expected)
