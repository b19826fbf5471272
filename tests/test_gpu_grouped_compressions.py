"""Grouped compressions on the K = 16 FP32 path (tuning compress_depth 2 and
3: up to 4 or 8 slots composed in one launch, inner compression slots never
stored). Checked here:

* log Z and marginals agree with the FP64 oracle within the FP32 tolerances
  (|dlogZ| <= 1e-5 * max(1, |logZ|), |dmarginal| <= 1e-5) and with the
  ungrouped program (depth 1) within 1e-6, on deep compression trees
  (paths, caterpillars), shallow ones (random recursive trees) and forests,
  for one and several batch items;
* the log-potential adjoint pair agrees with the oracle at every depth,
  with gradient seeds of both signs, and a depth change between its forward
  and backward call is refused;
* the profiled grouped levels move exactly the K-vectors that a separate
  derivation of the groups from the plan arrays predicts (group sizes, their
  levels, and the reverse launches split into groups of 2, of 3 or 4, and
  of more slots).
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
import parallel_tree_hmm_amd as thmm
from conftest import caterpillar, forest, path, random_potentials, random_recursive, root_index_of_nodes, star

pytestmark = pytest.mark.gpu

K = 16


def plans(rng):
    return {
        "path": path(3000),
        "caterpillar": caterpillar(2001),
        "recursive": random_recursive(2000, rng),
        "forest": forest([caterpillar(1501), path(700), random_recursive(900, rng), star(300)]),
    }


def device_marginals(torch, ws, plan, nodes, edges, batch):
    d_nodes = torch.tensor(nodes, dtype=torch.float32, device="cuda")
    d_edges = torch.tensor(edges, dtype=torch.float32, device="cuda")
    lz = torch.empty(batch * plan.roots.size, dtype=torch.float64, device="cuda")
    nm = torch.empty((batch, plan.num_nodes, K), dtype=torch.float32, device="cuda")
    em = torch.empty((batch, plan.num_edges, K, K), dtype=torch.float32, device="cuda")
    ws.marginals_device(d_nodes.data_ptr(), d_edges.data_ptr(), batch, lz.data_ptr(), nm.data_ptr(), em.data_ptr(), 0)
    torch.cuda.synchronize()
    return lz.cpu().numpy().reshape(batch, -1), nm.cpu().numpy(), em.cpu().numpy()


@pytest.mark.parametrize("batch", [1, 3])
@pytest.mark.parametrize("name", ["path", "caterpillar", "recursive", "forest"])
def test_depths_match_oracle_and_each_other(require_gpu, name, batch):
    import torch

    rng = np.random.default_rng(sum(map(ord, name)) + batch)
    plan = thmm.Plan(plans(rng)[name])
    nodes, edges = random_potentials(rng, batch, plan.num_nodes, plan.num_edges, K, zero_fraction=0.1, sigma=1.0)
    nodes, edges = nodes.astype(np.float32), edges.astype(np.float32)
    olz, onm, oem = oracle.marginals_batched(plan, K, nodes.astype(np.float64), edges.astype(np.float64), threads=4)
    ws = thmm.Workspace()
    ws.reserve_marginals(plan, K, batch, precision="f32", device_resident=True)
    assert ws.tuning()["compress_depth"] == 2
    results = {}
    for depth in (1, 2, 3):
        ws.set_tuning(compress_depth=depth)
        lz, nm, em = results[depth] = device_marginals(torch, ws, plan, nodes, edges, batch)
        print(f"{name} batch {batch} depth {depth}: |dlogZ| {np.max(np.abs(lz - olz)):.2e} "
              f"|dnode| {np.max(np.abs(nm - onm)):.2e} |dedge| {np.max(np.abs(em - oem)):.2e}")
        assert np.all(np.abs(lz - olz) <= 1e-5 * np.maximum(1.0, np.abs(olz)))
        assert np.max(np.abs(nm - onm)) <= 1e-5
        assert np.max(np.abs(em - oem)) <= 1e-5
    for depth in (2, 3):
        lz, nm, em = results[depth]
        assert np.max(np.abs(lz - results[1][0]) / np.maximum(1.0, np.abs(results[1][0]))) <= 1e-6
        assert np.max(np.abs(nm - results[1][1])) <= 1e-6
        assert np.max(np.abs(em - results[1][2])) <= 1e-6


def test_adjoint_pair_at_every_depth(require_gpu):
    import torch

    rng = np.random.default_rng(3)
    plan = thmm.Plan(forest([caterpillar(801), path(600), random_recursive(500, rng)]))
    batch = 2
    log_nodes = torch.tensor(rng.normal(size=(batch, plan.num_nodes, K)), dtype=torch.float32, device="cuda")
    log_edges = torch.tensor(rng.normal(size=(plan.num_edges, K, K)), dtype=torch.float32, device="cuda")
    grad = torch.tensor(rng.normal(size=batch * plan.roots.size), dtype=torch.float64, device="cuda")
    nodes = np.exp(log_nodes.double().cpu().numpy())
    edges = np.exp(log_edges.double().cpu().numpy())
    olz, onm, oem = oracle.marginals_batched(plan, K, nodes, edges, threads=4)
    ws = thmm.Workspace()
    ws.reserve_marginals(plan, K, batch, precision="f32", device_resident=True)
    g = grad.cpu().numpy().reshape(batch, -1)
    # Each node's gradient is its marginal times its tree's seed (random
    # signs: negative seeds reach the compression adjoints); an edge takes
    # its child's tree.
    seeds_nodes = g[:, root_index_of_nodes(plan)]
    seeds_edges = seeds_nodes[:, plan.edge_children]
    lz = torch.empty(batch * plan.roots.size, dtype=torch.float64, device="cuda")
    gn = torch.empty((batch, plan.num_nodes, K), device="cuda")
    ge = torch.empty((batch, plan.num_edges, K, K), device="cuda")
    for depth in (1, 2, 3):
        ws.set_tuning(compress_depth=depth)
        ws.logz_forward_device(log_nodes.data_ptr(), log_edges.data_ptr(), batch, lz.data_ptr(), 0)
        ws.logz_backward_device(grad.data_ptr(), gn.data_ptr(), ge.data_ptr(), False, 0)
        torch.cuda.synchronize()
        got = lz.cpu().numpy().reshape(batch, -1)
        assert np.all(np.abs(got - olz) <= 1e-5 * np.maximum(1.0, np.abs(olz)))
        assert np.max(np.abs(gn.cpu().double().numpy() - seeds_nodes[:, :, None] * onm)) <= 2e-5
        assert np.max(np.abs(ge.cpu().double().numpy() - seeds_edges[:, :, None, None] * oem)) <= 4e-5
    # A new depth between the two calls of the pair drops the held forward.
    ws.logz_forward_device(log_nodes.data_ptr(), log_edges.data_ptr(), batch, lz.data_ptr(), 0)
    ws.set_tuning(compress_depth=2)
    with pytest.raises(thmm.LogicError):
        ws.logz_backward_device(grad.data_ptr(), gn.data_ptr(), ge.data_ptr(), False, 0)
    torch.cuda.synchronize()


def expected_group_work(plan, depth):
    """Per compression batch with groups: (forward, pair reverse, small
    reverse, large reverse) K-vectors per item, from the plan arrays alone: slot renaming,
    heights, then groups formed top-down (an input compression is taken in,
    left first, when its height is not a multiple of `depth` and the group
    stays within 2**depth slots; otherwise it starts its own group)."""
    E = plan.num_edges
    current = np.arange(E)
    comps = plan.compressions
    batches = [(int(b["offset"]), int(b["count"])) for b in plan.primitive_batches if int(b["primitive"]) == 3]
    ops = {}
    for off, cnt in batches:
        for c in range(off, off + cnt):
            ops[c] = (int(current[comps[c]["left_edge"]]), int(comps[c]["middle"]), int(current[comps[c]["right_edge"]]))
        for c in range(off, off + cnt):
            current[comps[c]["left_edge"]] = E + c
    height, starts = {}, {c: True for c in ops}
    for off, cnt in batches:
        for c in range(off, off + cnt):
            inputs = [s - E for s in (ops[c][0], ops[c][2]) if s >= E]
            height[c] = 1 + max([height[x] for x in inputs], default=0)
            for x in inputs:
                starts[x] = False
    limit = 2 ** depth
    sizes = {}
    for off, cnt in reversed(batches):
        for c in reversed(range(off, off + cnt)):
            if not starts[c]:
                continue
            slots = [ops[c][0], ops[c][2]]
            k = 0
            while k < len(slots):
                x = slots[k] - E
                if x >= 0 and not starts[x]:
                    if height[x] % depth != 0 and len(slots) < limit:
                        slots[k:k + 1] = [ops[x][0], ops[x][2]]
                        continue
                    starts[x] = True
                k += 1
            sizes[c] = len(slots)
    out = []
    for off, cnt in batches:
        n = np.array([sizes[c] for c in range(off, off + cnt) if starts[c]], np.int64)
        assert np.all((2 <= n) & (n <= limit))
        if n.size:
            forward = (n + 1) * K + n - 1
            reverse = (2 * n + 1) * K + 2 * (n - 1)
            pair, small = n == 2, (n == 3) | (n == 4)
            out.append((int(forward.sum()), int(reverse[pair].sum()), int(reverse[small].sum()),
                        int(reverse[n > 4].sum())))
    return out


@pytest.mark.parametrize("depth", [2, 3])
def test_profiled_group_levels_match_plan(require_gpu, depth):
    import torch

    rng = np.random.default_rng(depth)
    plan = thmm.Plan(forest([caterpillar(1001), path(500), random_recursive(700, rng)]))
    nodes, edges = random_potentials(rng, 1, plan.num_nodes, plan.num_edges, K)
    ws = thmm.Workspace()
    ws.reserve_marginals(plan, K, 1, precision="f32", device_resident=True)
    ws.set_tuning(compress_depth=depth)
    d_nodes = torch.tensor(nodes, dtype=torch.float32, device="cuda")
    d_edges = torch.tensor(edges, dtype=torch.float32, device="cuda")
    lz = torch.empty(plan.roots.size, dtype=torch.float64, device="cuda")
    nm = torch.empty((plan.num_nodes, K), device="cuda")
    em = torch.empty((plan.num_edges, K, K), device="cuda")
    timings = ws.profile_marginals_device(d_nodes.data_ptr(), d_edges.data_ptr(), 1, lz.data_ptr(), nm.data_ptr(),
                                          em.data_ptr(), 0)
    forward = [count for phase, primitive, count, batch, ms in timings if phase == 1 and primitive == 6]
    reverse = [(primitive, count) for phase, primitive, count, batch, ms in timings
               if phase == 3 and primitive in (6, 7, 8)]
    assert not any(primitive == 3 for phase, primitive, count, batch, ms in timings)
    expected = expected_group_work(plan, depth)
    assert forward == [f for f, pair, small, large in expected]
    # Reverse: batches in reverse order, each as its groups of more than 4
    # slots (primitive 6), then those of 3 or 4 (primitive 7), then those
    # of 2 (primitive 8).
    want = []
    for f, pair, small, large in reversed(expected):
        want += [(6, large)] if large else []
        want += [(7, small)] if small else []
        want += [(8, pair)] if pair else []
    assert reverse == want
