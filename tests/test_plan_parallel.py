"""Forest planning on host threads (THMM_PLAN_THREADS): a forest whose trees
occupy consecutive node ranges is planned tree by tree and merged round by
round; the merged plan must equal the single contraction's plan array for
array (ops, ids, batches, round offsets), for forests of every topology the
workloads use, and forests in any other node order keep the single
contraction. Errors keep their messages."""

from __future__ import annotations

import os
import time

import numpy as np
import pytest

import parallel_tree_hmm_amd as thmm
from conftest import caterpillar, forest, path, random_recursive, star, yule

FIELDS = ("roots", "edge_parents", "edge_children", "rakes", "branch_combinations", "branch_absorptions",
          "compressions", "primitive_batches", "round_batch_offsets")


def planned(parents, threads):
    old = os.environ.get("THMM_PLAN_THREADS")
    os.environ["THMM_PLAN_THREADS"] = str(threads)
    try:
        start = time.perf_counter()
        plan = thmm.Plan(parents)
        return plan, time.perf_counter() - start
    finally:
        if old is None:
            del os.environ["THMM_PLAN_THREADS"]
        else:
            os.environ["THMM_PLAN_THREADS"] = old


def assert_same(a, b):
    assert (a.num_nodes, a.num_edges, a.num_branches, a.num_compressions, a.num_rounds) == (
        b.num_nodes, b.num_edges, b.num_branches, b.num_compressions, b.num_rounds)
    for name in FIELDS:
        assert np.array_equal(getattr(a, name), getattr(b, name)), name


def mixed_forest(rng, trees, low, high):
    parts = []
    for t in range(trees):
        n = int(np.exp(rng.uniform(np.log(low), np.log(high))))
        parts.append([random_recursive(n, rng), caterpillar(n), star(n), path(n), yule(max(2, n // 2), rng)][t % 5])
    return forest(parts)


@pytest.mark.parametrize("threads", [2, 3, 8])
def test_parallel_plan_equals_single_contraction(threads):
    rng = np.random.default_rng(threads)
    parents = mixed_forest(rng, 300, 2, 3000)
    assert parents.size >= 1 << 16
    single, t1 = planned(parents, 1)
    split, tn = planned(parents, threads)
    print(f"{parents.size} nodes, {single.roots.size} trees: 1 thread {t1:.3f} s, {threads} threads {tn:.3f} s")
    assert_same(single, split)


def test_single_node_trees_and_one_big_tree():
    rng = np.random.default_rng(5)
    parents = forest([np.array([-1])] * 50 + [random_recursive(70000, rng)] + [np.array([-1])] * 50)
    assert_same(planned(parents, 1)[0], planned(parents, 4)[0])


def test_interleaved_forest_keeps_single_contraction():
    # Two paths whose nodes alternate: not consecutive ranges.
    n = 40000
    parents = np.empty(2 * n, np.int64)
    parents[0], parents[1] = -1, -1
    parents[2::2] = np.arange(0, 2 * n - 2, 2)
    parents[3::2] = np.arange(1, 2 * n - 2, 2)
    assert_same(planned(parents, 1)[0], planned(parents, 4)[0])


def test_errors_keep_their_messages():
    bad = mixed_forest(np.random.default_rng(1), 100, 500, 2000)
    bad[12345] = bad.size + 7
    with pytest.raises(thmm.InvalidArgument, match="parent of node 12345 is out of range"):
        planned(bad, 4)
