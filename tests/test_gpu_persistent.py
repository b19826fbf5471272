"""Persistent passes (tuning field `persistent`), switched off in this build
(every test here is skipped, see the marker):  on a batch-1 K = 16 FP32
forest each pass of the contraction runs as one launch whose blocks pull
tasks from a device queue, a task waiting only for the earlier levels of its
own group of trees. The results must equal the level-by-level launches bit
for bit:

* thmm_marginals_device (log Z, node and edge marginals) on an uneven forest
  (random recursive, caterpillar and star trees of 16 to 6,000 nodes, inputs
  from the portable SplitMix64 recipe), eagerly and replayed from a captured
  graph, and thmm_log_partition_device;
* the log-potential pair (forward, then the backward half twice);
* a forest whose trees do not occupy consecutive node ranges (the ops of a
  level are then not ordered by tree group, and every tree forms one group);
* a state count the persistent kernels do not cover (K = 8), which keeps the
  level launches.
thmm_synchronize reports a wait past the spin limit; none may occur.
"""

from __future__ import annotations

import numpy as np
import pytest

import bench
import parallel_tree_hmm_amd as thmm
from conftest import forest, random_recursive

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skip(reason="persistent passes are switched off: two MI355X runs of the 115k-node forest did not "
                            "complete (120 s, then 60 s with every dependency wait bounded at 2 s); cause not shown"),
]


def forest_inputs(torch, parents, k, seed):
    plan = thmm.Plan(parents)
    nodes = bench.fill_lognormal_device(torch, torch.empty((plan.num_nodes, k), device="cuda"), seed)
    edges = bench.fill_dirichlet_rows_device(torch, torch.empty((plan.num_edges, k, k), device="cuda"), seed + 1)
    return plan, nodes, edges


def marginals(torch, plan, k, nodes, edges, persistent, graphs=False, calls=1):
    ws = thmm.Workspace()
    ws.reserve_marginals(plan, k, 1, device_resident=True, graphs=graphs)
    ws.set_tuning(persistent=persistent)
    out = []
    for _ in range(calls):
        lz = torch.full((plan.roots.size,), float("nan"), dtype=torch.float64, device="cuda")
        nm = torch.full((plan.num_nodes, k), float("nan"), device="cuda")
        em = torch.full((plan.num_edges, k, k), float("nan"), device="cuda")
        ws.marginals_device(nodes.data_ptr(), edges.data_ptr(), 1, lz.data_ptr(), nm.data_ptr(), em.data_ptr(), 0)
        pz = torch.full_like(lz, float("nan"))
        ws.log_partition_device(nodes.data_ptr(), edges.data_ptr(), 1, pz.data_ptr(), 0)
        ws.synchronize()
        out.append((lz, nm, em, pz))
    ws.close()
    return out


def same(a, b):
    import torch

    return all(torch.equal(x, y) for x, y in zip(a, b))


def test_persistent_marginals_equal_level_launches(require_gpu):
    import torch

    parts = bench.forest_parts(4, trees=120, high=6000)
    plan, nodes, edges = forest_inputs(torch, bench.concatenate(parts)[0], 16, 41)
    print(f"{plan.roots.size} trees, {plan.num_nodes} nodes, {plan.num_compressions} compressions", flush=True)
    (levels,) = marginals(torch, plan, 16, nodes, edges, 0)
    assert bool(torch.isfinite(levels[0]).all())
    (persistent,) = marginals(torch, plan, 16, nodes, edges, 1)
    assert same(levels, persistent)
    for replay in marginals(torch, plan, 16, nodes, edges, 1, graphs=True, calls=3):
        assert same(levels, replay)


def test_persistent_log_potential_pair(require_gpu):
    import torch

    parts = bench.forest_parts(4, trees=60, high=4000)
    plan = thmm.Plan(bench.concatenate(parts)[0])
    k = 16
    log_nodes = bench.fill_lognormal_device(torch, torch.empty((plan.num_nodes, k), device="cuda"), 7, log=True)
    log_edges = bench.fill_lognormal_device(torch, torch.empty((plan.num_edges, k, k), device="cuda"), 8, log=True)
    results = []
    for persistent in (0, 1):
        ws = thmm.Workspace()
        ws.reserve_marginals(plan, k, 1, device_resident=True)
        ws.set_tuning(persistent=persistent)
        lz = torch.empty(plan.roots.size, dtype=torch.float64, device="cuda")
        ws.logz_forward_device(log_nodes.data_ptr(), log_edges.data_ptr(), 1, lz.data_ptr(), 0)
        outs = [lz]
        for _ in range(2):
            gn = torch.empty((plan.num_nodes, k), device="cuda")
            ge = torch.empty((plan.num_edges, k, k), device="cuda")
            ws.logz_backward_device(0, gn.data_ptr(), ge.data_ptr(), False, 0)
            outs += [gn, ge]
        ws.synchronize()
        ws.close()
        results.append(outs)
    assert same(results[0], results[1])


def test_persistent_with_interleaved_trees_and_uncovered_state_counts(require_gpu):
    import torch

    rng = np.random.default_rng(5)
    trees = [random_recursive(n, rng) for n in (300, 800, 150, 1200)]
    parents = forest(trees)
    # Renumber the nodes so that the trees interleave.
    perm = rng.permutation(parents.size)
    perm[np.flatnonzero(parents < 0)] = np.sort(perm[np.flatnonzero(parents < 0)])
    inverse = np.empty_like(perm)
    inverse[perm] = np.arange(perm.size)
    shuffled = np.full(parents.size, -1, np.int64)
    for v in range(parents.size):
        if parents[v] >= 0:
            shuffled[perm[v]] = perm[parents[v]]
    for k in (16, 8):
        plan, nodes, edges = forest_inputs(torch, shuffled, k, 90 + k)
        (levels,) = marginals(torch, plan, k, nodes, edges, 0)
        (persistent,) = marginals(torch, plan, k, nodes, edges, 1)
        assert same(levels, persistent), k
