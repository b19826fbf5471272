"""Parity at the shapes of the benchmark workloads that fit an oracle run:

* 64 random recursive trees of 1,023 nodes, K = 4, plain FP64: log Z and
  all node and edge marginals against the oracle, both as one topology with
  64 batch items (shared edges) and as a 64-tree forest (per-tree edges);
* 16 paths of 100,000 nodes, K = 8, FP32, node potentials exp(N(0, 2)) with
  10 % of entries exactly 0, edges exp(N(0, 1)): scaled log Z and marginals
  within 1e-5, MAP states and posterior samples (one seeded U[0,1) variate
  per node) bit-exact against the oracle, variates within 1e-12 of a CDF
  boundary counted separately;
* random full binary (Yule) trees, K = 64, log potentials N(0, 1), FP32
  through the forward/backward adjoint pair: gradients against the oracle's
  marginals on 4 trees of 1,023 nodes, and on 16 trees of 16,383 nodes
  (1/16 of the 256-tree workload) through size-independent properties
  (marginals sum to one, edge rows and columns sum to the endpoint
  marginals) and against the plain FP64 form.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
import parallel_tree_hmm_amd as thmm
from conftest import forest, path, random_recursive, yule

pytestmark = pytest.mark.gpu


def dirichlet_rows(rng, e, k):
    rows = rng.exponential(1.0, size=(e, k, k))
    return rows / rows.sum(axis=2, keepdims=True)


def test_recursive_trees_fp64_batched(require_gpu):
    rng = np.random.default_rng(0x54484D4D00000001)
    plan = thmm.Plan(random_recursive(1023, rng))
    k, batch = 4, 64
    nodes = np.exp(rng.normal(size=(batch, 1023, k)))
    edges = dirichlet_rows(rng, 1022, k)
    ws = thmm.Workspace()
    ws.reserve_marginals(plan, k, batch, precision="f64")
    out = thmm.posterior_marginals_prepared(thmm.BatchedModelView(plan, k, batch, nodes, edges), ws)
    olz, onm, oem = oracle.marginals_batched(plan, k, nodes, edges, threads=8)
    assert np.max(np.abs(out.log_partitions.reshape(batch, 1) - olz) / np.maximum(1, np.abs(olz))) <= 1e-11
    assert np.max(np.abs(out.nodes.reshape(onm.shape) - onm)) <= 1e-11
    assert np.max(np.abs(out.edges.reshape(oem.shape) - oem)) <= 1e-11


def test_recursive_trees_fp64_forest(require_gpu):
    rng = np.random.default_rng(0x54484D4D00000001 + 1)
    trees = [random_recursive(1023, rng) for _ in range(64)]
    plan = thmm.Plan(forest(trees))
    k = 4
    nodes = np.exp(rng.normal(size=(1, plan.num_nodes, k)))
    edges = dirichlet_rows(rng, plan.num_edges, k)
    ws = thmm.Workspace()
    ws.reserve_marginals(plan, k, 1, precision="f64")
    out = thmm.posterior_marginals_prepared(thmm.BatchedModelView(plan, k, 1, nodes, edges), ws)
    olz, onm, oem = oracle.marginals(plan, k, nodes[0], edges)
    assert olz.shape == (64,)
    assert np.max(np.abs(ws.log_partitions_f64()[:64] - olz) / np.maximum(1, np.abs(olz))) <= 1e-11
    assert np.max(np.abs(out.nodes.reshape(onm.shape) - onm)) <= 1e-11
    assert np.max(np.abs(out.edges.reshape(oem.shape) - oem)) <= 1e-11


@pytest.fixture(scope="module")
def long_chains():
    rng = np.random.default_rng(0x54484D4D00000002)
    n, k, batch = 100_000, 8, 16
    plan = thmm.Plan(path(n))
    nodes = np.exp(rng.normal(0.0, 2.0, size=(batch, n, k))).astype(np.float32)
    nodes[rng.random(nodes.shape) < 0.1] = 0.0
    dead = np.all(nodes == 0.0, axis=2)
    nodes[dead, 0] = 1.0  # Z > 0 needs one supported state per node
    edges = np.exp(rng.normal(size=(n - 1, k, k))).astype(np.float32)
    u = rng.random((batch, n)).astype(np.float32)
    return plan, k, batch, nodes, edges, u


def test_long_chains_marginals(require_gpu, long_chains):
    plan, k, batch, nodes, edges, _ = long_chains
    ws = thmm.Workspace()
    ws.reserve_marginals(plan, k, batch, precision="f32")
    out = thmm.posterior_marginals_prepared(thmm.BatchedModelView(plan, k, batch, nodes, edges), ws)
    olz, onm, oem = oracle.marginals_batched(plan, k, nodes, edges, threads=16)
    log_z = ws.log_partitions_f64()[:batch].reshape(batch, 1)
    assert np.max(np.abs(log_z - olz) / np.maximum(1, np.abs(olz))) <= 1e-5
    assert np.max(np.abs(out.nodes.reshape(onm.shape) - onm)) <= 1e-5
    assert np.max(np.abs(out.edges.reshape(oem.shape) - oem)) <= 1e-5


def test_long_chains_map(require_gpu, long_chains):
    plan, k, batch, nodes, edges, _ = long_chains
    ws = thmm.Workspace()
    ws.reserve_maximum(plan, k, batch, precision="f32")
    out = thmm.maximum_a_posteriori_prepared(thmm.BatchedModelView(plan, k, batch, nodes, edges), ws)
    states = out.states.reshape(batch, -1)
    for b in range(batch):
        lw, expected = oracle.maximum(plan, k, nodes[b], edges)
        assert abs(float(out.log_weights[b]) - lw[0]) <= 1e-6 * max(1.0, abs(lw[0]))
        assert np.array_equal(states[b], expected), f"MAP states differ for chain {b}"


def test_long_chains_samples(require_gpu, long_chains):
    plan, k, batch, nodes, edges, u = long_chains
    ws = thmm.Workspace()
    ws.reserve_sampling(plan, k, batch, precision="f32")
    out = thmm.posterior_sample_prepared(thmm.BatchedModelView(plan, k, batch, nodes, edges), u, ws)
    states = out.states.reshape(batch, -1)
    near_boundary = 0
    for b in range(batch):
        expected, dist = oracle.sample(plan, k, nodes[b], edges, u[b])
        close = dist < 1e-12
        near_boundary += int(close.sum())
        # Ancestral sampling: a boundary variate may change its node and, through
        # it, the draws below; compare exactly when none is present.
        if close.any():
            continue
        assert np.array_equal(states[b], expected), f"sample differs for chain {b}"
    print(f"variates within 1e-12 of a CDF boundary: {near_boundary}")
    assert near_boundary <= 2


def yule_forest(rng, trees, leaves):
    return forest([yule(leaves, rng) for _ in range(trees)])


def test_dense_state_adjoint_against_oracle(require_gpu):
    import torch

    from parallel_tree_hmm_amd.autograd import log_partition, reserve_for

    rng = np.random.default_rng(0x54484D4D00000005)
    plan = thmm.Plan(yule_forest(rng, 4, 512))
    k = 64
    log_nodes = torch.tensor(rng.normal(size=(1, plan.num_nodes, k)), dtype=torch.float32, device="cuda",
                             requires_grad=True)
    log_edges = torch.tensor(rng.normal(size=(plan.num_edges, k, k)), dtype=torch.float32, device="cuda",
                             requires_grad=True)
    ws = reserve_for(plan, k, 1, torch.float32)
    log_z = log_partition(log_nodes, log_edges, ws)
    log_z.sum().backward()
    nodes = np.exp(log_nodes.detach().double().cpu().numpy()[0])
    edges = np.exp(log_edges.detach().double().cpu().numpy())
    olz, onm, oem = oracle.marginals(plan, k, nodes, edges)
    errors = {
        "log_z_rel": float(np.max(np.abs(log_z.detach().cpu().numpy()[0] - olz) / np.maximum(1, np.abs(olz)))),
        "node_grad_abs": float(np.max(np.abs(log_nodes.grad.double().cpu().numpy()[0] - onm))),
        "edge_grad_abs": float(np.max(np.abs(log_edges.grad.double().cpu().numpy() - oem))),
    }
    print(f"K=64 FP32 adjoint errors versus the FP64 oracle: {errors}")
    assert errors["log_z_rel"] <= 1e-5
    assert errors["node_grad_abs"] <= 1e-5
    assert errors["edge_grad_abs"] <= 1e-5


def test_dense_state_adjoint_properties_at_scale(require_gpu):
    import torch

    from parallel_tree_hmm_amd.autograd import log_partition, reserve_for

    rng = np.random.default_rng(0x54484D4D00000005 + 1)
    plan = thmm.Plan(yule_forest(rng, 16, 8192))
    assert plan.num_nodes == 16 * 16383
    k = 64
    gen = torch.Generator(device="cuda")
    gen.manual_seed(5)
    log_nodes = torch.randn((1, plan.num_nodes, k), device="cuda", generator=gen)
    log_edges = torch.randn((plan.num_edges, k, k), device="cuda", generator=gen)
    ws = reserve_for(plan, k, 1, torch.float32)
    x, y = log_nodes.clone().requires_grad_(), log_edges.clone().requires_grad_()
    log_z32 = log_partition(x, y, ws)
    log_z32.sum().backward()
    node_grad, edge_grad = x.grad[0].double(), y.grad
    del ws
    assert torch.isfinite(node_grad).all() and torch.isfinite(edge_grad).all()
    assert float((node_grad.sum(1) - 1).abs().max()) <= 1e-5
    parents = torch.from_numpy(plan.edge_parents.astype(np.int64)).cuda()
    children = torch.from_numpy(plan.edge_children.astype(np.int64)).cuda()
    for lo in range(0, plan.num_edges, 1 << 16):
        hi = min(plan.num_edges, lo + (1 << 16))
        block = edge_grad[lo:hi].double()
        assert float((block.sum(2) - node_grad[parents[lo:hi]]).abs().max()) <= 1e-5
        assert float((block.sum(1) - node_grad[children[lo:hi]]).abs().max()) <= 1e-5
    del x, y, edge_grad
    torch.cuda.empty_cache()

    ws64 = reserve_for(plan, k, 1, torch.float64)
    with torch.no_grad():
        log_z64 = log_partition(log_nodes.double(), log_edges.double(), ws64)
    rel = ((log_z32 - log_z64).abs() / log_z64.abs().clamp(min=1)).max()
    print(f"K=64 FP32 versus FP64 log Z, 16 trees of 16,383 nodes: max rel {float(rel):.3g}")
    assert float(rel) <= 1e-5
