"""Randomised GPU-vs-oracle parity over state counts, topologies, forests and
batch shapes, for every kernel family:

* log Z and marginals (the adjoint): FP32 form within the stated
  tolerances (|dlogZ| <= 1e-5 * max(1, |logZ|), |dmarginal| <= 1e-5),
  plain FP64 form within 1e-11;
* MAP states and posterior samples: bit-exact against the oracle, except
  where the oracle reports a near-tie / a variate within 1e-12 of a CDF
  boundary (counted and required to be rare, never silently accepted);
* device-resident entry points and the log-potential forward/backward pair.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
import parallel_tree_hmm_amd as thmm
from conftest import caterpillar, forest, path, random_potentials, random_recursive, root_index_of_nodes, star, yule

pytestmark = pytest.mark.gpu

LOGZ_TOL = {"f32": 1e-5, "f64": 1e-11}
MARG_TOL = {"f32": 1e-5, "f64": 1e-11}


def topology(name, n, rng):
    if name == "recursive":
        return random_recursive(n, rng)
    if name == "path":
        return path(n)
    if name == "star":
        return star(n)
    if name == "caterpillar":
        return caterpillar(n)
    if name == "forest":
        parts = [random_recursive(max(1, n // 4), rng), path(max(1, n // 4)), star(max(1, n // 4)), caterpillar(max(1, n - 3 * (n // 4)))]
        return forest(parts)
    raise KeyError(name)


def as_precision(x, precision):
    return np.asarray(x, np.float32).astype(np.float64) if precision == "f32" else np.asarray(x, np.float64)


CASES = [
    # (K, topology, nodes, batch)
    (1, "path", 300, 3),
    (2, "recursive", 200, 5),
    (3, "star", 100, 4),
    (4, "caterpillar", 301, 33),
    (4, "forest", 400, 2),
    (5, "recursive", 150, 7),
    (8, "path", 1000, 4),
    (8, "forest", 300, 3),
    (9, "recursive", 120, 3),
    (16, "caterpillar", 121, 2),
    (16, "forest", 200, 1),
    (17, "recursive", 60, 2),
    (32, "path", 70, 2),
    (33, "recursive", 40, 1),
    (64, "caterpillar", 41, 1),
    (64, "recursive", 40, 2),
]


@pytest.mark.parametrize("precision", ["f32", "f64"])
@pytest.mark.parametrize("k,topo,n,batch", CASES)
def test_log_partition_and_marginals(require_gpu, precision, k, topo, n, batch):
    rng = np.random.default_rng(k * 1000 + n)
    plan = thmm.Plan(topology(topo, n, rng))
    nodes, edges = random_potentials(rng, batch, plan.num_nodes, plan.num_edges, k, zero_fraction=0.1, sigma=1.0)
    nodes, edges = as_precision(nodes, precision), as_precision(edges, precision)
    ws = thmm.Workspace()
    ws.reserve_marginals(plan, k, batch, precision=precision)
    out = thmm.posterior_marginals_prepared(thmm.BatchedModelView(plan, k, batch, nodes, edges), ws)
    roots = plan.roots.size
    log_z = ws.log_partitions_f64()[: batch * roots].reshape(batch, roots)
    gnodes = out.nodes.reshape(batch, plan.num_nodes, k)
    gedges = out.edges.reshape(batch, plan.num_edges, k, k)
    olz, onm, oem = oracle.marginals_batched(plan, k, nodes, edges, threads=4)
    assert np.all(np.abs(log_z - olz) <= LOGZ_TOL[precision] * np.maximum(1.0, np.abs(olz)))
    assert np.max(np.abs(gnodes - onm)) <= MARG_TOL[precision]
    if plan.num_edges:
        assert np.max(np.abs(gedges - oem)) <= MARG_TOL[precision]
    # log partition alone agrees with the marginal pass
    lp = thmm.log_partition_function_prepared(thmm.BatchedModelView(plan, k, batch, nodes, edges), ws)
    assert np.allclose(ws.log_partitions_f64()[: batch * roots].reshape(batch, roots), log_z, rtol=0, atol=0)
    assert lp.values.shape == (batch * roots,)


# MAP and sampling: every case above plus the K >= 32 team and group-per-item
# kernels on raked subtrees, multi-child absorptions, forests and
# batch-varying items.
DECODE_CASES = CASES + [
    (32, "forest", 200, 3),
    (32, "recursive", 150, 3),
    (32, "star", 120, 4),
    (48, "recursive", 60, 2),
    (48, "forest", 80, 2),
    (64, "forest", 80, 2),
    (64, "star", 60, 3),
    (64, "path", 50, 2),
]


@pytest.mark.parametrize("precision", ["f32", "f64"])
@pytest.mark.parametrize("k,topo,n,batch", DECODE_CASES)
def test_maximum_matches_oracle(require_gpu, precision, k, topo, n, batch):
    rng = np.random.default_rng(7 * k + n)
    plan = thmm.Plan(topology(topo, n, rng))
    nodes, edges = random_potentials(rng, batch, plan.num_nodes, plan.num_edges, k, zero_fraction=0.1, sigma=2.0)
    nodes, edges = as_precision(nodes, precision), as_precision(edges, precision)
    ws = thmm.Workspace()
    ws.reserve_maximum(plan, k, batch, precision=precision)
    out = thmm.maximum_a_posteriori_prepared(thmm.BatchedModelView(plan, k, batch, nodes, edges), ws)
    states = out.states.reshape(batch, -1)
    weights = out.log_weights.reshape(batch, -1)
    for b in range(batch):
        lw, expected = oracle.maximum(plan, k, nodes[b], edges)
        assert np.all(np.abs(weights[b] - lw) <= (1e-6 if precision == "f32" else 1e-12) * np.maximum(1, np.abs(lw)))
        assert np.array_equal(states[b], expected), f"MAP states differ for item {b}"


@pytest.mark.parametrize("precision", ["f32", "f64"])
@pytest.mark.parametrize("k,topo,n,batch", DECODE_CASES)
def test_samples_match_oracle(require_gpu, precision, k, topo, n, batch):
    rng = np.random.default_rng(11 * k + n)
    plan = thmm.Plan(topology(topo, n, rng))
    nodes, edges = random_potentials(rng, batch, plan.num_nodes, plan.num_edges, k, zero_fraction=0.1, sigma=2.0)
    nodes, edges = as_precision(nodes, precision), as_precision(edges, precision)
    u = as_precision(rng.random((batch, plan.num_nodes)), precision)
    ws = thmm.Workspace()
    ws.reserve_sampling(plan, k, batch, precision=precision)
    out = thmm.posterior_sample_prepared(thmm.BatchedModelView(plan, k, batch, nodes, edges), u, ws)
    states = out.states.reshape(batch, -1)
    near_boundary = 0
    for b in range(batch):
        expected, dist = oracle.sample(plan, k, nodes[b], edges, u[b])
        if np.min(dist) < 1e-12:
            near_boundary += 1  # reported separately, not compared
            continue
        assert np.array_equal(states[b], expected), f"sample differs for item {b}"
    assert near_boundary <= max(1, batch // 10)


@pytest.mark.parametrize("precision", ["f32", "f64"])
@pytest.mark.parametrize("topo,n", [("star", 30000), ("caterpillar", 20001), ("forest", 24000)])
def test_dynamic_range_of_large_trees(require_gpu, precision, topo, n):
    """Products of tens of thousands of messages span far more than FP32's
    linear range (a 30,000-leaf star's root factor spans > e^200); the log
    vectors with FP64 scales keep log Z and marginals within tolerance."""
    rng = np.random.default_rng(n)
    k = 16
    plan = thmm.Plan(topology(topo, n, rng))
    nodes, edges = random_potentials(rng, 1, plan.num_nodes, plan.num_edges, k)
    nodes, edges = as_precision(nodes, precision), as_precision(edges, precision)
    ws = thmm.Workspace()
    ws.reserve_marginals(plan, k, 1, precision=precision)
    out = thmm.posterior_marginals_prepared(thmm.BatchedModelView(plan, k, 1, nodes, edges), ws)
    log_z = ws.log_partitions_f64()[: plan.roots.size]
    olz, onm, oem = oracle.marginals(plan, k, nodes[0], edges)
    assert np.all(np.isfinite(out.nodes)) and np.all(np.isfinite(out.edges))
    assert np.all(np.abs(log_z - olz) <= LOGZ_TOL[precision] * np.maximum(1.0, np.abs(olz)))
    assert np.max(np.abs(out.nodes.reshape(onm.shape) - onm)) <= MARG_TOL[precision]
    assert np.max(np.abs(out.edges.reshape(oem.shape) - oem)) <= MARG_TOL[precision]


def test_device_entry_points_match_host(require_gpu):
    import torch

    rng = np.random.default_rng(5)
    plan = thmm.Plan(forest([random_recursive(300, rng), path(200), star(100)]))
    k, batch = 8, 3
    nodes, edges = random_potentials(rng, batch, plan.num_nodes, plan.num_edges, k)
    nodes, edges = nodes.astype(np.float32), edges.astype(np.float32)
    ws = thmm.Workspace()
    ws.reserve_marginals(plan, k, batch)
    host = thmm.posterior_marginals_prepared(thmm.BatchedModelView(plan, k, batch, nodes, edges), ws)
    host_logz = ws.log_partitions_f64()[: batch * 3].copy()
    host_nodes, host_edges = host.nodes.copy(), host.edges.copy()

    dws = thmm.Workspace()
    dws.reserve_marginals(plan, k, batch, device_resident=True)
    d_nodes = torch.from_numpy(nodes).cuda()
    d_edges = torch.from_numpy(edges).cuda()
    d_logz = torch.empty(batch * 3, dtype=torch.float64, device="cuda")
    d_nm = torch.empty(batch * plan.num_nodes * k, dtype=torch.float32, device="cuda")
    d_em = torch.empty(batch * plan.num_edges * k * k, dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    dws.marginals_device(
        d_nodes.data_ptr(), d_edges.data_ptr(), batch, d_logz.data_ptr(), d_nm.data_ptr(), d_em.data_ptr(), stream
    )
    torch.cuda.synchronize()
    assert np.array_equal(d_logz.cpu().numpy(), host_logz)
    assert np.array_equal(d_nm.cpu().numpy(), host_nodes)
    assert np.array_equal(d_em.cpu().numpy(), host_edges)
    d_lz2 = torch.empty_like(d_logz)
    dws.log_partition_device(d_nodes.data_ptr(), d_edges.data_ptr(), batch, d_lz2.data_ptr(), stream)
    torch.cuda.synchronize()
    assert np.array_equal(d_lz2.cpu().numpy(), host_logz)


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_log_potential_forward_backward(require_gpu, precision):
    import torch

    rng = np.random.default_rng(9)
    plan = thmm.Plan(yule(40, rng))
    k, batch = 6, 3
    dtype = torch.float32 if precision == "f32" else torch.float64
    log_nodes = torch.tensor(rng.normal(size=(batch, plan.num_nodes, k)), dtype=dtype)
    log_edges = torch.tensor(rng.normal(size=(plan.num_edges, k, k)), dtype=dtype)
    ws = thmm.Workspace()
    ws.reserve_marginals(plan, k, batch, precision=precision)
    d_ln, d_le = log_nodes.cuda(), log_edges.cuda()
    d_logz = torch.empty(batch, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    ws.logz_forward_device(d_ln.data_ptr(), d_le.data_ptr(), batch, d_logz.data_ptr(), stream)
    grad = torch.tensor([1.0, -0.5, 2.0], dtype=torch.float64, device="cuda")
    g_nodes = torch.empty_like(d_ln)
    g_edges = torch.empty_like(d_le)
    ws.logz_backward_device(grad.data_ptr(), g_nodes.data_ptr(), g_edges.data_ptr(), True, stream)
    torch.cuda.synchronize()
    nodes = np.exp(log_nodes.double().numpy())
    edges = np.exp(log_edges.double().numpy())
    olz, onm, oem = oracle.marginals_batched(plan, k, nodes, edges)
    tol = 1e-5 if precision == "f32" else 1e-11
    assert np.all(np.abs(d_logz.cpu().numpy() - olz[:, 0]) <= tol * np.maximum(1, np.abs(olz[:, 0])))
    g = grad.cpu().numpy()
    assert np.max(np.abs(g_nodes.cpu().double().numpy() - g[:, None, None] * onm)) <= tol * 2
    assert np.max(np.abs(g_edges.cpu().double().numpy() - np.einsum("b,bekl->ekl", g, oem))) <= tol * 4


@pytest.mark.parametrize("k", [16, 32, 64])
def test_gradient_seeds_of_either_sign(require_gpu, k):
    """The adjoint pair with per-(item, tree) seeds of both signs: the MFMA
    compression adjoints normalise their ratios by magnitude, so a negative
    seed scales every node and edge gradient of its tree, compressions
    included (ungrouped compressions: compress_depth 1)."""
    import torch

    rng = np.random.default_rng(k)
    plan = thmm.Plan(forest([caterpillar(121), path(90), random_recursive(60, rng)]))
    batch = 2
    log_nodes = torch.tensor(rng.normal(size=(batch, plan.num_nodes, k)), dtype=torch.float32, device="cuda")
    log_edges = torch.tensor(rng.normal(size=(plan.num_edges, k, k)), dtype=torch.float32, device="cuda")
    seeds = np.array([[-1.5, 0.7, -0.2], [2.0, -0.6, 1.1]])
    grad = torch.tensor(seeds.ravel(), dtype=torch.float64, device="cuda")
    ws = thmm.Workspace()
    ws.reserve_marginals(plan, k, batch, precision="f32", device_resident=True)
    if k == 16:
        ws.set_tuning(compress_depth=1)
    lz = torch.empty(batch * 3, dtype=torch.float64, device="cuda")
    ws.logz_forward_device(log_nodes.data_ptr(), log_edges.data_ptr(), batch, lz.data_ptr(), 0)
    gn = torch.empty_like(log_nodes)
    ge = torch.empty((batch, plan.num_edges, k, k), device="cuda")
    ws.logz_backward_device(grad.data_ptr(), gn.data_ptr(), ge.data_ptr(), False, 0)
    torch.cuda.synchronize()
    olz, onm, oem = oracle.marginals_batched(plan, k, np.exp(log_nodes.double().cpu().numpy()),
                                             np.exp(log_edges.double().cpu().numpy()), threads=4)
    per_node = seeds[:, root_index_of_nodes(plan)]
    assert np.all(np.abs(lz.cpu().numpy().reshape(batch, 3) - olz) <= 1e-5 * np.maximum(1, np.abs(olz)))
    assert np.max(np.abs(gn.cpu().double().numpy() - per_node[:, :, None] * onm)) <= 4e-5
    per_edge = per_node[:, plan.edge_children]
    assert np.max(np.abs(ge.cpu().double().numpy() - per_edge[:, :, None, None] * oem)) <= 4e-5


def test_unaligned_device_buffers_fall_back(require_gpu):
    """Edge potentials and outputs that are only 4-byte aligned take the
    generic kernels (the coalesced ones use 16-byte accesses) and agree with
    the aligned call."""
    import torch

    rng = np.random.default_rng(17)
    plan = thmm.Plan(forest([random_recursive(300, rng), star(200), caterpillar(201)]))
    k = 16
    nodes, edges = random_potentials(rng, 1, plan.num_nodes, plan.num_edges, k)
    ws = thmm.Workspace()
    ws.reserve_marginals(plan, k, 1, precision="f32", device_resident=True)
    d_nodes = torch.tensor(nodes, dtype=torch.float32, device="cuda")
    storage = torch.zeros(plan.num_edges * k * k + 1, dtype=torch.float32, device="cuda")
    shifted = storage[1:]  # 4 bytes past a 256-byte boundary
    shifted.copy_(torch.tensor(edges.ravel(), dtype=torch.float32))
    aligned = shifted.clone()
    results = []
    for e in (aligned, shifted):
        assert (e.data_ptr() % 16 == 0) == (e is aligned)
        lz = torch.empty(plan.roots.size, dtype=torch.float64, device="cuda")
        nm = torch.empty(plan.num_nodes * k + 1, dtype=torch.float32, device="cuda")[1:]
        em = torch.empty(plan.num_edges * k * k + 1, dtype=torch.float32, device="cuda")[1:]
        ws.marginals_device(d_nodes.data_ptr(), e.data_ptr(), 1, lz.data_ptr(), nm.data_ptr(), em.data_ptr(), 0)
        torch.cuda.synchronize()
        results.append((lz.cpu().numpy(), nm.cpu().numpy(), em.cpu().numpy()))
    (lz0, n0, e0), (lz1, n1, e1) = results
    assert np.max(np.abs(lz0 - lz1) / np.maximum(1, np.abs(lz0))) <= 1e-6
    assert np.max(np.abs(n0 - n1)) <= 1e-6 and np.max(np.abs(e0 - e1)) <= 1e-6


@pytest.mark.parametrize("k,precision", [(8, "f32"), (16, "f32"), (32, "f32"), (64, "f32"), (16, "f64")])
def test_log_edges_read_in_place(require_gpu, k, precision):
    """The adjoint pair over log potentials with the rake kernels reading the
    log edges in place (tuning rake_log_edges = 1; at K = 8 and 16 and FP64
    only the edges compressions read are staged, at K = 32 and 64 FP32 the
    compressions read theirs in place too) gives the same bits as staging
    every edge first (rake_log_edges = 0), also through graph replay; with
    unaligned gradient outputs the backward stages the edges itself, and
    with unaligned log edges the forward stages every edge. Dead
    rows (all -inf) sit in a rake edge and in both inputs of the first
    compression. Values against the FP64 oracle."""
    import torch

    rng = np.random.default_rng(k)
    plan = thmm.Plan(forest([random_recursive(150, rng), caterpillar(121), path(60), star(40)]))
    batch = 2
    dtype = torch.float32 if precision == "f32" else torch.float64
    log_nodes = torch.tensor(rng.normal(size=(batch, plan.num_nodes, k)), dtype=dtype, device="cuda")
    log_edges = torch.tensor(rng.normal(size=(plan.num_edges, k, k)), dtype=dtype, device="cuda")
    log_edges[0, 1, :] = -float("inf")  # a dead row
    first = plan.compressions[0]  # reads two edges, not composed paths
    log_edges[int(first["left_edge"]), 2, :] = -float("inf")
    log_edges[int(first["right_edge"]), 3, :] = -float("inf")
    grad = torch.tensor(rng.normal(size=batch * plan.roots.size), dtype=torch.float64, device="cuda")

    # The same log edges 4 or 8 bytes past 16-byte alignment: the forward
    # then stages every edge.
    edge_storage = torch.empty(log_edges.numel() + 1, dtype=dtype, device="cuda")
    shifted_edges = edge_storage[1:]
    shifted_edges.copy_(log_edges.reshape(-1))
    assert shifted_edges.data_ptr() % 16 != 0

    def run(in_place, graphs=False, shift=0, edges=log_edges):
        ws = thmm.Workspace()
        ws.reserve_marginals(plan, k, batch, precision=precision, device_resident=True, graphs=graphs)
        ws.set_tuning(rake_log_edges=in_place)
        outs = []
        for _ in range(2 if graphs else 1):
            lz = torch.empty(batch * plan.roots.size, dtype=torch.float64, device="cuda")
            gn = torch.empty(batch * plan.num_nodes * k + shift, dtype=dtype, device="cuda")[shift:]
            ge = torch.empty(batch * plan.num_edges * k * k + shift, dtype=dtype, device="cuda")[shift:]
            ws.logz_forward_device(log_nodes.data_ptr(), edges.data_ptr(), batch, lz.data_ptr(), 0)
            ws.logz_backward_device(grad.data_ptr(), gn.data_ptr(), ge.data_ptr(), False, 0)
            torch.cuda.synchronize()
            outs.append((lz.cpu(), gn.cpu(), ge.cpu()))
        return outs[-1]

    staged, direct, replayed = run(0), run(1), run(1, graphs=True)
    for a, b, c in zip(staged, direct, replayed):
        assert torch.equal(a, b) and torch.equal(a, c)
    # Other kernels (and, for unaligned log edges, the generic staging
    # kernel) round differently: log Z within a relative, gradients within
    # an absolute FP32 tolerance.
    close = 1e-6 if precision == "f32" else 1e-12
    for other in (run(1, shift=1), run(1, edges=shifted_edges)):
        assert float(((staged[0] - other[0]).abs() / staged[0].abs().clamp(min=1.0)).max()) <= close
        for a, b in zip(staged[1:], other[1:]):
            assert float((a.double() - b.double()).abs().max()) <= close

    nodes = np.exp(log_nodes.double().cpu().numpy())
    edges = np.exp(log_edges.double().cpu().numpy())
    olz, onm, oem = oracle.marginals_batched(plan, k, nodes, edges, threads=4)
    g = grad.cpu().numpy().reshape(batch, -1)
    per_node = g[:, root_index_of_nodes(plan)]
    tol = 1e-5 if precision == "f32" else 1e-11
    lz = direct[0].numpy().reshape(batch, -1)
    assert np.all(np.abs(lz - olz) <= tol * np.maximum(1.0, np.abs(olz)))
    gn = direct[1].double().numpy().reshape(batch, plan.num_nodes, k)
    assert np.max(np.abs(gn - per_node[:, :, None] * onm)) <= 4 * tol
    ge = direct[2].double().numpy().reshape(batch, plan.num_edges, k, k)
    assert np.max(np.abs(ge - per_node[:, plan.edge_children][:, :, None, None] * oem)) <= 4 * tol
