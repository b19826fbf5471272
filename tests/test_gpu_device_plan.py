"""The device planner (thmm_plan_create_device) against the host planner
(thmm_plan_create): identical arrays (roots, edges, rakes, combinations,
absorptions, compressions, batches, round offsets) on every topology the
host planner's tests use, on shuffled node orders (children before their
parents), forests with single nodes and interleaved trees, parents given
on the device, the same errors, and the full 4,096-tree K = 16 forest
(33 M nodes) with both planners timed."""

from __future__ import annotations

import time

import numpy as np
import pytest

import bench
import parallel_tree_hmm_amd as thmm
from conftest import balanced_binary, caterpillar, forest, path, random_recursive, star, yule

pytestmark = pytest.mark.gpu

FIELDS = ("roots", "edge_parents", "edge_children", "rakes", "branch_combinations", "branch_absorptions",
          "compressions", "primitive_batches", "round_batch_offsets")


def assert_same_plan(host, device, label):
    for name in FIELDS:
        a, b = getattr(host, name), getattr(device, name)
        assert a.shape == b.shape and a.tobytes() == b.tobytes(), f"{label}: {name} differs"
    assert (host.num_nodes, host.num_edges, host.num_branches, host.num_rounds) == (
        device.num_nodes, device.num_edges, device.num_branches, device.num_rounds), label


def shuffled(parents, rng):
    """The same tree with its nodes renumbered at random."""
    n = parents.size
    perm = rng.permutation(n)  # new id of old node i
    out = np.empty(n, np.int64)
    out[perm] = np.where(parents < 0, -1, perm[np.maximum(parents, 0)])
    return out


def topologies(rng):
    yield "single", np.array([-1])
    yield "pair", np.array([-1, 0])
    yield "path", path(1000)
    yield "star", star(5000)
    yield "caterpillar", caterpillar(2001)
    yield "binary", balanced_binary(4095)
    yield "recursive", random_recursive(3000, rng)
    yield "yule", yule(700, rng)
    yield "forest", forest([random_recursive(300, rng), np.array([-1]), star(64), path(257), np.array([-1, 0]),
                            caterpillar(101), yule(50, rng)])
    yield "shuffled_recursive", shuffled(random_recursive(2500, rng), rng)
    yield "shuffled_path", shuffled(path(3000), rng)
    yield "shuffled_forest", shuffled(forest([path(400), star(300), random_recursive(500, rng)]), rng)
    yield "long_path", path(1 << 20)
    yield "wide_star", star(70000)


def test_device_plans_equal_host_plans(require_gpu):
    rng = np.random.default_rng(17)
    for name, parents in topologies(rng):
        host = thmm.Plan(parents)
        device = thmm.Plan.on_device(parents)
        assert_same_plan(host, device, name)
    for seed in range(20):  # random forests, including interleaved trees
        r = np.random.default_rng(seed)
        trees = [random_recursive(int(r.integers(1, 200)), r) for _ in range(int(r.integers(1, 12)))]
        parents = forest(trees)
        if seed % 2:
            parents = shuffled(parents, r)
        assert_same_plan(thmm.Plan(parents), thmm.Plan.on_device(parents), f"forest {seed}")


def test_parents_on_the_device(require_gpu):
    import torch

    rng = np.random.default_rng(3)
    parents = forest([random_recursive(5000, rng), caterpillar(999), star(300)])
    device = thmm.Plan.on_device(torch.from_numpy(parents).cuda())
    assert_same_plan(thmm.Plan(parents), device, "torch parents")
    with pytest.raises(thmm.InvalidArgument):
        thmm.Plan.on_device(torch.from_numpy(parents.astype(np.int32)).cuda())


@pytest.mark.parametrize("parents", [
    [0],                   # self parent
    [-1, 5],               # out of range
    [-1, 0, -2],           # negative, not -1
    [1, 0],                # no root
    [-1, 0, 3, 2],         # cycle beside a valid tree
    [-1, 0, 1, 4, 5, 3],   # cycle of three
])
def test_errors_match_the_host_planner(require_gpu, parents):
    with pytest.raises(thmm.InvalidArgument) as host:
        thmm.Plan(parents)
    with pytest.raises(thmm.InvalidArgument) as device:
        thmm.Plan.on_device(parents)
    assert str(host.value) == str(device.value)


def test_empty_plan_is_rejected(require_gpu):
    with pytest.raises(thmm.InvalidArgument, match="at least one node"):
        thmm.Plan.on_device([])


def test_full_forest(require_gpu):
    parts = bench.forest_parts(4)
    parents, _ = bench.concatenate(parts)
    start = time.perf_counter()
    host = thmm.Plan(parents)
    host_s = time.perf_counter() - start
    thmm.Plan.on_device(parents[:100_000])  # first use of the kernels
    start = time.perf_counter()
    device = thmm.Plan.on_device(parents)
    device_s = time.perf_counter() - start
    print(f"{parents.size} nodes, {host.num_rounds} rounds: host planner {host_s:.3f} s, device planner "
          f"{device_s:.3f} s (both including the arrays' copy into the plan object)", flush=True)
    assert_same_plan(host, device, "forest of 4096 trees")
