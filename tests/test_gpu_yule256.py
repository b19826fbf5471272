"""The full K = 64 dense-state workload: 256 random full binary (Yule) trees
of 16,383 nodes (4,194,048 nodes, 4,193,792 edges), log node and edge
potentials N(0, 1) from the benchmark's SplitMix64 streams, FP32, through the
device-resident forward/backward pair that `bench.py --config 5` times
(about 172 GB of device memory: 68.7 GB each of log edges and edge
gradients, the rest workspace).

Checked over the whole forest: every log Z finite, every node's gradient
(its marginal) summing to one, every edge gradient's rows and columns
summing to the parent's and child's node gradients (1e-5). Checked against
the FP64 oracle on the first and last trees: log Z (1e-5 relative) and
every node and edge gradient (1e-5). The smaller shapes of the same
workload (4 trees of 1,023 nodes against the oracle, 16 trees against the
plain FP64 form) are in test_gpu_workloads.py.
"""

from __future__ import annotations

import numpy as np
import pytest

import bench
import oracle
import parallel_tree_hmm_amd as thmm

pytestmark = pytest.mark.gpu

TREES, LEAVES, K = 256, 8192, 64
NEEDED_BYTES = 190 * (1 << 30)


def test_full_dense_state_workload(require_gpu):
    import torch

    torch.cuda.empty_cache()
    free, _ = torch.cuda.mem_get_info()
    if free < NEEDED_BYTES:
        pytest.skip(f"needs {NEEDED_BYTES >> 30} GiB of free device memory, {free >> 30} GiB free")
    topology = bench.uniforms_host(bench.config_seed(5, bench.STREAM_TOPOLOGY), 0, TREES * (LEAVES - 1))
    topology = topology.reshape(TREES, LEAVES - 1)
    parts = [bench.yule_tree(LEAVES, topology[t]) for t in range(TREES)]
    parents, offsets = bench.concatenate(parts)
    plan = thmm.Plan(parents)
    n, e = plan.num_nodes, plan.num_edges
    assert n == TREES * (2 * LEAVES - 1)
    print(f"Yule forest: {TREES} trees, {n} nodes, {plan.num_compressions} compressions", flush=True)

    log_nodes = bench.fill_lognormal_device(torch, torch.empty((1, n, K), device="cuda"),
                                            bench.config_seed(5, bench.STREAM_NODES), log=True)
    log_edges = bench.fill_lognormal_device(torch, torch.empty((e, K, K), device="cuda"),
                                            bench.config_seed(5, bench.STREAM_EDGES), log=True)
    grad_nodes = torch.empty((1, n, K), device="cuda")
    grad_edges = torch.empty((e, K, K), device="cuda")
    log_z = torch.empty(plan.roots.size, dtype=torch.float64, device="cuda")
    ws = thmm.Workspace()
    ws.reserve_marginals(plan, K, 1, precision="f32", device_resident=True)
    stream = torch.cuda.current_stream().cuda_stream
    ws.logz_forward_device(log_nodes.data_ptr(), log_edges.data_ptr(), 1, log_z.data_ptr(), stream)
    ws.logz_backward_device(0, grad_nodes.data_ptr(), grad_edges.data_ptr(), False, stream)
    torch.cuda.synchronize()
    ws.close()
    print("forward/backward done", flush=True)

    assert bool(torch.isfinite(log_z).all())
    node_grad = grad_nodes[0]
    assert bool(torch.isfinite(node_grad).all())
    assert float((node_grad.double().sum(1) - 1.0).abs().max()) <= 1e-5
    edge_parents = torch.from_numpy(plan.edge_parents.astype(np.int64)).cuda()
    edge_children = torch.from_numpy(plan.edge_children.astype(np.int64)).cuda()
    rows = cols = 0.0
    for lo in range(0, e, 1 << 15):
        hi = min(e, lo + (1 << 15))
        block = grad_edges[lo:hi].double()
        assert bool(torch.isfinite(block).all())
        rows = max(rows, float((block.sum(2) - node_grad[edge_parents[lo:hi]].double()).abs().max()))
        cols = max(cols, float((block.sum(1) - node_grad[edge_children[lo:hi]].double()).abs().max()))
        del block
    print(f"edge rows {rows:.3g}, columns {cols:.3g} from the endpoint marginals", flush=True)
    assert rows <= 1e-5 and cols <= 1e-5

    log_z_host = log_z.cpu().numpy()
    for t in (0, TREES - 1):
        lo, hi, elo, ehi = bench.tree_slices(offsets, t)
        tree_plan = thmm.Plan(parts[t])
        nodes = torch.exp(log_nodes[0, lo:hi].double()).cpu().numpy()
        edges = torch.exp(log_edges[elo:ehi].double()).cpu().numpy()
        olz, onm, oem = oracle.marginals(tree_plan, K, nodes, edges)
        errors = (abs(log_z_host[t] - olz[0]) / max(1.0, abs(olz[0])),
                  float(np.max(np.abs(grad_nodes[0, lo:hi].double().cpu().numpy() - onm))),
                  float(np.max(np.abs(grad_edges[elo:ehi].double().cpu().numpy() - oem))))
        print(f"tree {t}: log Z rel {errors[0]:.3g}, node {errors[1]:.3g}, edge {errors[2]:.3g}", flush=True)
        assert errors[0] <= 1e-5 and errors[1] <= 1e-5 and errors[2] <= 1e-5
    del log_nodes, log_edges, grad_nodes, grad_edges
    torch.cuda.empty_cache()
