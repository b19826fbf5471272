"""Launches past the 32-bit work-item count of a dispatch.

A dispatch's grid size is a 32-bit count of work-items. Kernels over arrays
of more than 2^32 values step through them by the grid's size, and level
launches split their op range below 2^31 work-items. Before that, the
widening of the sampling edges and the max-product log transform touched
only the first E*K*K - 2^32 values past the limit: trees beyond it sampled
from unwritten memory (log Z = -inf, failure flag set) and MAP read
unwritten log factors.

Checked through properties that hold at any size (the oracle is too slow
here):

* MAP at K = 64 on 1.1 M nodes (E*K*K = 4.5e9): every tree's log weight
  equals the log-likelihood of the returned states, summed in FP64 on the
  device, and the states are in range;
* sampling on the same forest: log Z equals the FP32 log partition of the
  same inputs within 1e-6, and no conditional is reported empty;
* a 600,000-leaf star at batch 64 and K = 4 (one rake level of 38.4 M
  (op, item) pairs, split into two launches): log Z equals its closed form
  log sum_s root[s] prod_l sum_c E_l[s][c] leaf_l[c] (FP64, torch), node
  marginals sum to one and edge marginals' rows sum to the root marginal.
"""

from __future__ import annotations

import numpy as np
import pytest

import parallel_tree_hmm_amd as thmm
from parallel_tree_hmm_amd import _abi
from conftest import forest, random_recursive

pytestmark = pytest.mark.gpu


def device_reservation(plan, k, batch, flag):
    ws = thmm.Workspace()
    ws.reserve(plan, k, batch, precision="f32", flags=flag | _abi.THMM_RESERVE_DEVICE_RESIDENT)
    return ws


def test_map_and_sampling_past_four_billion_edge_values(require_gpu):
    import torch

    rng = np.random.default_rng(4)
    plan = thmm.Plan(forest([random_recursive(70_000, rng) for _ in range(16)]))
    k = 64
    n, e = plan.num_nodes, plan.num_edges
    assert e * k * k > 2**32
    print(f"{n} nodes, {e * k * k:.3e} edge values", flush=True)
    gen = torch.Generator(device="cuda").manual_seed(4)
    nodes = torch.rand((n, k), device="cuda", generator=gen) + 0.05
    edges = torch.rand((e, k, k), device="cuda", generator=gen) + 0.05
    roots = plan.roots.size
    parents = torch.tensor(plan.edge_parents.astype(np.int64), device="cuda")
    children = torch.tensor(plan.edge_children.astype(np.int64), device="cuda")
    root_of = torch.tensor(np.repeat(np.arange(16), 70_000), device="cuda")

    ws = device_reservation(plan, k, 1, _abi.THMM_RESERVE_MAXIMUM)
    weights = torch.empty(roots, dtype=torch.float64, device="cuda")
    states = torch.empty(n, dtype=torch.int32, device="cuda")
    ws.maximum_device(nodes.data_ptr(), edges.data_ptr(), 1, weights.data_ptr(), states.data_ptr(), 0)
    torch.cuda.synchronize()
    ws.close()
    del ws
    s = states.long()
    assert bool(((s >= 0) & (s < k)).all())
    # log-likelihood of the MAP states per tree, FP64
    node_terms = torch.log(nodes.double().gather(1, s[:, None])[:, 0])
    edge_terms = torch.log(edges[torch.arange(e, device="cuda"), s[parents], s[children]].double())
    ll = torch.zeros(roots, dtype=torch.float64, device="cuda")
    ll.index_add_(0, root_of, node_terms)
    ll.index_add_(0, root_of[children], edge_terms)
    gap = ((ll - weights).abs() / weights.abs().clamp(min=1.0)).max().item()
    print(f"MAP weight vs likelihood of its states: {gap:.2e}", flush=True)
    assert gap <= 1e-12

    ws = device_reservation(plan, k, 1, _abi.THMM_RESERVE_SAMPLING)
    uniforms = torch.rand(n, device="cuda", generator=gen)
    log_z = torch.empty(roots, dtype=torch.float64, device="cuda")
    failure = torch.zeros(1, dtype=torch.int32, device="cuda")
    ws.sample_device(nodes.data_ptr(), edges.data_ptr(), uniforms.data_ptr(), 1, log_z.data_ptr(),
                     states.data_ptr(), failure.data_ptr(), 0)
    torch.cuda.synchronize()
    ws.close()
    del ws
    assert int(failure.item()) == 0
    assert bool(torch.isfinite(log_z).all())
    ws = device_reservation(plan, k, 1, _abi.THMM_RESERVE_MARGINALS)
    reference = torch.empty(roots, dtype=torch.float64, device="cuda")
    ws.log_partition_device(nodes.data_ptr(), edges.data_ptr(), 1, reference.data_ptr(), 0)
    torch.cuda.synchronize()
    ws.close()
    rel = ((log_z - reference).abs() / reference.abs().clamp(min=1.0)).max().item()
    print(f"sampling log Z vs FP32 log partition: {rel:.2e}", flush=True)
    assert rel <= 1e-6


def test_level_split_into_launches_on_a_large_star(require_gpu):
    import torch

    leaves, k, batch = 600_000, 4, 64
    plan = thmm.Plan(np.concatenate([[-1], np.zeros(leaves, np.int64)]))
    assert leaves * batch * 64 >= 2**31
    gen = torch.Generator(device="cuda").manual_seed(8)
    nodes = torch.rand((batch, leaves + 1, k), device="cuda", generator=gen) + 0.1
    edges = torch.rand((leaves, k, k), device="cuda", generator=gen) + 0.1
    ws = thmm.Workspace()
    ws.reserve_marginals(plan, k, batch, precision="f32", device_resident=True)
    log_z = torch.empty(batch, dtype=torch.float64, device="cuda")
    node_out = torch.empty((batch, leaves + 1, k), device="cuda")
    edge_out = torch.empty((batch, leaves, k, k), device="cuda")
    ws.marginals_device(nodes.data_ptr(), edges.data_ptr(), batch, log_z.data_ptr(), node_out.data_ptr(),
                        edge_out.data_ptr(), 0)
    torch.cuda.synchronize()
    # log Z_b = log sum_s root[b, s] prod_l sum_c E[l, s, c] leaf[b, l, c]
    x = nodes.double()
    messages = torch.zeros((batch, k), dtype=torch.float64, device="cuda")
    for start in range(0, leaves, 100_000):
        stop = min(leaves, start + 100_000)
        m = (edges[start:stop].double()[None] * x[:, 1 + start:1 + stop, None, :]).sum(dim=3)  # [b, l, s]
        messages += torch.log(m).sum(dim=1)
    expected = torch.logsumexp(torch.log(x[:, 0]) + messages, dim=1)
    rel = ((log_z - expected).abs() / expected.abs().clamp(min=1.0)).max().item()
    print(f"star log Z vs closed form: {rel:.2e}", flush=True)
    assert rel <= 1e-6
    assert (node_out.double().sum(dim=2) - 1.0).abs().max().item() <= 1e-5
    rows = edge_out.double().sum(dim=3)  # [b, l, s]: rows sum to the root marginal
    assert (rows - node_out[:, :1, :].double()).abs().max().item() <= 1e-5
    ws.close()


def test_k32_team_kernels_and_their_large_launch_forms_agree_with_the_oracle(require_gpu):
    """At K = 32 the max-product rake and compression kernels and the FP64
    compression forward of sampling use 256 lanes per (op, item); a level of
    2^23 (op, item) pairs would pass 2^31 work-items, so those launches keep
    the group-per-item kernels. A broom (2^15 arms root -> a -> b -> leaf: one
    level of 2^15 rakes, one of 2^15 compressions, one of 2^15 rakes into the
    root) at batch 256 takes that route; the same items at batch 2 take the
    team kernels. States and draws of both are bit-exact against the oracle
    for two items, and the log weights and log partitions agree."""
    import oracle
    import torch

    arms, k, batch = 1 << 15, 32, 256
    parents = np.full(1 + 3 * arms, -1, np.int64)
    a = 1 + 3 * np.arange(arms)
    parents[a], parents[a + 1], parents[a + 2] = 0, a, a + 1
    plan = thmm.Plan(parents)
    assert arms * batch * k * k // 4 >= 2**31
    n, e = plan.num_nodes, plan.num_edges
    gen = torch.Generator(device="cuda").manual_seed(32)
    nodes = torch.exp(torch.randn((batch, n, k), device="cuda", generator=gen) * 2.0)
    edges = torch.exp(torch.randn((e, k, k), device="cuda", generator=gen))
    uniforms = torch.rand((batch, n), device="cuda", generator=gen)
    uniforms = (uniforms * 2.0**24).floor() * 2.0**-24  # exact in FP32, below 1
    host_edges = edges.double().cpu().numpy()
    items = (0, 1)
    host_nodes = [nodes[b].double().cpu().numpy() for b in items]
    host_u = [uniforms[b].double().cpu().numpy() for b in items]
    results = {}
    for size in (batch, 2):
        ws = device_reservation(plan, k, size, _abi.THMM_RESERVE_MAXIMUM)
        weights = torch.empty(size, dtype=torch.float64, device="cuda")
        states = torch.empty((size, n), dtype=torch.int32, device="cuda")
        ws.maximum_device(nodes.data_ptr(), edges.data_ptr(), size, weights.data_ptr(), states.data_ptr(), 0)
        torch.cuda.synchronize()
        ws.close()
        del ws
        results[("map", size)] = (weights[:2].cpu().numpy(), states[:2].cpu().numpy().astype(np.uint32))
        del weights, states
        torch.cuda.empty_cache()
        ws = device_reservation(plan, k, size, _abi.THMM_RESERVE_SAMPLING)
        log_z = torch.empty(size, dtype=torch.float64, device="cuda")
        states = torch.empty((size, n), dtype=torch.int32, device="cuda")
        failure = torch.ones(1, dtype=torch.int32, device="cuda")
        ws.sample_device(nodes.data_ptr(), edges.data_ptr(), uniforms.data_ptr(), size, log_z.data_ptr(),
                         states.data_ptr(), failure.data_ptr(), 0)
        torch.cuda.synchronize()
        ws.close()
        del ws
        assert int(failure.item()) == 0
        results[("sample", size)] = (log_z[:2].cpu().numpy(), states[:2].cpu().numpy().astype(np.uint32))
        del log_z, states
        torch.cuda.empty_cache()
        print(f"batch {size}: MAP and sampling done", flush=True)
    for i, b in enumerate(items):
        lw, expected = oracle.maximum(plan, k, host_nodes[i], host_edges)
        for size in (batch, 2):
            weights, states = results[("map", size)]
            assert np.array_equal(states[i], expected), f"MAP states of item {b} at batch {size}"
            assert abs(weights[i] - lw[0]) <= 1e-12 * abs(lw[0])
        drawn, dist = oracle.sample(plan, k, host_nodes[i], host_edges, host_u[i])
        near = int((dist < 1e-12).sum())
        print(f"item {b}: {near} variates within 1e-12 of a boundary", flush=True)
        if near:
            continue  # reported separately, not compared
        for size in (batch, 2):
            log_z, states = results[("sample", size)]
            assert np.array_equal(states[i], drawn), f"draws of item {b} at batch {size}"

    assert np.array_equal(results[("map", batch)][0], results[("map", 2)][0])
    lz_big, lz_small = results[("sample", batch)][0], results[("sample", 2)][0]
    assert np.max(np.abs(lz_big - lz_small) / np.abs(lz_small)) <= 1e-13
