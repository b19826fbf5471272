"""hipGraph replay of the device-resident entry points (THMM_RESERVE_GRAPH):
every replay is compared bit for bit with the eager launch sequence on the
same inputs, with the program's own eager work and synchronisations between
replays, after the inputs change in place, after the buffers change (a new
capture), and on a non-default stream."""

from __future__ import annotations

import time

import numpy as np
import pytest

import parallel_tree_hmm_amd as thmm
from conftest import forest, path, random_recursive, star

pytestmark = pytest.mark.gpu


def inputs(torch, rng, plan, k, batch, dtype):
    nodes = torch.tensor(np.exp(rng.normal(size=(batch, plan.num_nodes, k))), dtype=dtype, device="cuda")
    rows = rng.exponential(1.0, size=(plan.num_edges, k, k))
    edges = torch.tensor(rows / rows.sum(axis=2, keepdims=True), dtype=dtype, device="cuda")
    return nodes, edges


def outputs(torch, plan, k, batch, dtype):
    return (
        torch.empty(batch * plan.roots.size, dtype=torch.float64, device="cuda"),
        torch.empty((batch, plan.num_nodes, k), dtype=dtype, device="cuda"),
        torch.empty((batch, plan.num_edges, k, k), dtype=dtype, device="cuda"),
    )


def marginals(ws, nodes, edges, batch, out, stream=None):
    log_z, nm, em = out
    ws.marginals_device(nodes.data_ptr(), edges.data_ptr(), batch, log_z.data_ptr(), nm.data_ptr(), em.data_ptr(),
                        stream)


@pytest.mark.parametrize("k,batch,precision", [(16, 1, "f32"), (8, 3, "f32"), (4, 5, "f64")])
def test_marginal_replays_match_eager(require_gpu, k, batch, precision):
    import torch

    rng = np.random.default_rng(100 + k)
    plan = thmm.Plan(forest([random_recursive(400, rng), path(300), star(200)]))
    dtype = torch.float32 if precision == "f32" else torch.float64
    eager, graph = thmm.Workspace(), thmm.Workspace()
    eager.reserve_marginals(plan, k, batch, precision=precision, device_resident=True)
    graph.reserve_marginals(plan, k, batch, precision=precision, device_resident=True, graphs=True)
    nodes, edges = inputs(torch, rng, plan, k, batch, dtype)
    ref, out = outputs(torch, plan, k, batch, dtype), outputs(torch, plan, k, batch, dtype)
    stream = torch.cuda.current_stream().cuda_stream

    for replay in range(4):
        if replay == 2:  # new values in the same buffers: the graph reads them
            fresh_nodes, fresh_edges = inputs(torch, rng, plan, k, batch, dtype)
            nodes.copy_(fresh_nodes)
            edges.copy_(fresh_edges)
        marginals(eager, nodes, edges, batch, ref, stream)
        torch.cuda.synchronize()
        for t in out:
            t.fill_(-7.0)  # eager work between replays
        marginals(graph, nodes, edges, batch, out, stream)
        torch.cuda.synchronize()
        for a, b in zip(ref, out):
            assert torch.equal(a, b), f"replay {replay} differs"

    # New output buffers: a new capture, on a side stream.
    side = torch.cuda.Stream()
    out2 = outputs(torch, plan, k, batch, dtype)
    with torch.cuda.stream(side):
        marginals(graph, nodes, edges, batch, out2, side.cuda_stream)
    side.synchronize()
    for a, b in zip(ref, out2):
        assert torch.equal(a, b)


def test_adjoint_pair_and_log_partition_replays(require_gpu):
    import torch

    rng = np.random.default_rng(7)
    plan = thmm.Plan(forest([random_recursive(300, rng), path(200)]))
    k, batch = 16, 2
    eager, graph = thmm.Workspace(), thmm.Workspace()
    eager.reserve_marginals(plan, k, batch, precision="f32", device_resident=True)
    graph.reserve_marginals(plan, k, batch, precision="f32", device_resident=True, graphs=True)
    log_nodes = torch.randn((batch, plan.num_nodes, k), device="cuda")
    log_edges = torch.randn((plan.num_edges, k, k), device="cuda")
    grad = torch.tensor(rng.normal(size=batch * 2), dtype=torch.float64, device="cuda")

    def run(ws):
        lz = torch.empty(batch * 2, dtype=torch.float64, device="cuda")
        gn = torch.empty((batch, plan.num_nodes, k), device="cuda")
        ge = torch.empty((batch, plan.num_edges, k, k), device="cuda")
        ws.logz_forward_device(log_nodes.data_ptr(), log_edges.data_ptr(), batch, lz.data_ptr(), 0)
        ws.logz_backward_device(grad.data_ptr(), gn.data_ptr(), ge.data_ptr(), False, 0)
        torch.cuda.synchronize()
        return lz, gn, ge

    expected = run(eager)
    for _ in range(3):
        got = run(graph)
        torch.cuda.synchronize()
        for a, b in zip(expected, got):
            assert torch.equal(a, b)

    nodes = log_nodes.exp()
    edges = log_edges.exp()
    a = torch.empty(batch * 2, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    eager.log_partition_device(nodes.data_ptr(), edges.data_ptr(), batch, a.data_ptr(), 0)
    for _ in range(3):
        graph.log_partition_device(nodes.data_ptr(), edges.data_ptr(), batch, b.data_ptr(), 0)
        torch.cuda.synchronize()
        assert torch.equal(a, b)


def test_small_workload_replay_timing(require_gpu):
    """Latency-bound shape (64 items on one 1,023-node tree, K = 4, FP64):
    reports eager and replayed per-call times (not asserted)."""
    import torch

    rng = np.random.default_rng(1)
    plan = thmm.Plan(random_recursive(1023, rng))
    k, batch = 4, 64
    nodes, edges = inputs(torch, rng, plan, k, batch, torch.float64)
    timings = {}
    for mode in ("eager", "graph"):
        ws = thmm.Workspace()
        ws.reserve_marginals(plan, k, batch, precision="f64", device_resident=True, graphs=(mode == "graph"))
        out = outputs(torch, plan, k, batch, torch.float64)
        stream = torch.cuda.current_stream().cuda_stream
        for _ in range(3):
            marginals(ws, nodes, edges, batch, out, stream)
        torch.cuda.synchronize()
        start = time.perf_counter()
        for _ in range(50):
            marginals(ws, nodes, edges, batch, out, stream)
        torch.cuda.synchronize()
        timings[mode] = (time.perf_counter() - start) / 50 * 1e3
    print(f"64 x 1023 nodes, K=4, FP64 marginals: eager {timings['eager']:.3f} ms, graph {timings['graph']:.3f} ms")
