"""Budgeted path storage (thmm_set_path_budget): a forest whose composed
paths and path adjoints exceed the budget runs in groups of consecutive
trees that reuse one path store; the backward half of the log-potential pair
recomputes each group's paths from the stored node state and the log edges.

Every case compares the budgeted workspace with an unbudgeted one on the same
inputs, bit for bit: log Z, node and edge gradients of the pair (forward
once, backward twice), log Z with node and edge marginals of
thmm_marginals_device, thmm_log_partition_device, and a captured-graph
replay of the budgeted pair. Shapes: K = 64 FP32 random full binary (Yule)
trees with log potentials N(0, 1) (the dense-state workload's shape at a
smaller size; compressions read log edges in place), K = 16 FP32 uneven
forest (fused forward, grouped compressions, staged compression edges),
K = 8 FP32 paths and recursive trees, K = 4 FP64 (plan-level kernels). Also:
a budget below one tree, batch 2, and a MAP reservation are length errors.
"""

from __future__ import annotations

import numpy as np
import pytest

import bench
import parallel_tree_hmm_amd as thmm
from parallel_tree_hmm_amd import _abi
from conftest import forest, path, random_recursive

pytestmark = pytest.mark.gpu


def yule_forest(trees, leaves, seed):
    u = bench.uniforms_host(seed, 0, trees * (leaves - 1)).reshape(trees, leaves - 1)
    return forest([bench.yule_tree(leaves, u[t]) for t in range(trees)])


def run_all(torch, plan, k, precision, log_nodes, log_edges, budget, graphs=False):
    dtype = torch.float32 if precision == "f32" else torch.float64
    ws = thmm.Workspace()
    ws.set_path_budget(budget)
    ws.reserve_marginals(plan, k, 1, precision=precision, device_resident=True, graphs=graphs)
    groups, held = ws.path_storage()
    n, e, roots = plan.num_nodes, plan.num_edges, plan.roots.size
    out = {}
    stream = torch.cuda.current_stream().cuda_stream
    lz = torch.full((roots,), float("nan"), dtype=torch.float64, device="cuda")
    ws.logz_forward_device(log_nodes.data_ptr(), log_edges.data_ptr(), 1, lz.data_ptr(), stream)
    for rep in range(2):  # the backward half twice after one forward half
        gn = torch.full((n, k), float("nan"), dtype=dtype, device="cuda")
        ge = torch.full((e, k, k), float("nan"), dtype=dtype, device="cuda")
        seed = torch.linspace(-1.0, 2.0, roots, dtype=torch.float64, device="cuda")
        ws.logz_backward_device(seed.data_ptr(), gn.data_ptr(), ge.data_ptr(), False, stream)
        torch.cuda.synchronize()
        out[f"pair{rep}"] = (lz.clone(), gn, ge)
    nodes, edges = torch.exp(log_nodes), torch.exp(log_edges)
    mz = torch.empty_like(lz)
    mn = torch.empty((n, k), dtype=dtype, device="cuda")
    me = torch.empty((e, k, k), dtype=dtype, device="cuda")
    ws.marginals_device(nodes.data_ptr(), edges.data_ptr(), 1, mz.data_ptr(), mn.data_ptr(), me.data_ptr(), stream)
    pz = torch.empty_like(lz)
    ws.log_partition_device(nodes.data_ptr(), edges.data_ptr(), 1, pz.data_ptr(), stream)
    torch.cuda.synchronize()
    out["marginals"] = (mz, mn, me)
    out["partition"] = (pz,)
    ws.close()
    return out, groups, held


def assert_equal(a, b, label):
    for name in a:
        for x, y in zip(a[name], b[name]):
            assert torch_equal(x, y), f"{label}: {name} differs"


def torch_equal(x, y):
    import torch

    return bool(torch.equal(torch.nan_to_num(x, nan=7.0), torch.nan_to_num(y, nan=7.0))) and not bool(
        torch.isnan(x).any())


CASES = {
    "k64_yule": dict(k=64, precision="f32", trees=lambda: yule_forest(8, 256, 5)),
    "k16_forest": dict(k=16, precision="f32",
                       trees=lambda: bench.concatenate(bench.forest_parts(4, trees=24, high=6000))[0]),
    "k8_paths": dict(k=8, precision="f32",
                     trees=lambda: forest([path(900), random_recursive(700, np.random.default_rng(3)), path(1200),
                                           random_recursive(500, np.random.default_rng(4)), path(300)])),
    "k4_fp64": dict(k=4, precision="f64",
                    trees=lambda: forest([random_recursive(400, np.random.default_rng(t)) for t in range(6)])),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_budgeted_groups_equal_the_stored_form(require_gpu, case):
    import torch

    spec = CASES[case]
    k, precision = spec["k"], spec["precision"]
    plan = thmm.Plan(spec["trees"]())
    n, e = plan.num_nodes, plan.num_edges
    dtype = torch.float32 if precision == "f32" else torch.float64
    log_nodes = bench.fill_lognormal_device(torch, torch.empty((n, k), dtype=dtype, device="cuda"), 11, log=True)
    log_edges = bench.fill_lognormal_device(torch, torch.empty((e, k, k), dtype=dtype, device="cuda"), 12, log=True)
    full, g0, held_full = run_all(torch, plan, k, precision, log_nodes, log_edges, 0)
    assert g0 == 0
    s = 4 if precision == "f32" else 8
    slot = k * k * s * 2 + k * 8
    tree = np.searchsorted(plan.roots, plan.compressions["middle"], side="right") - 1
    largest = int(np.bincount(tree, minlength=plan.roots.size).max())
    budget = max(largest * slot, plan.num_compressions * slot // 3)
    grouped, groups, held = run_all(torch, plan, k, precision, log_nodes, log_edges, budget)
    print(f"{case}: {plan.roots.size} trees, {plan.num_compressions} compressions, {groups} groups, path storage "
          f"{held} bytes against {held_full} unbudgeted", flush=True)
    assert groups >= 3 and held <= budget and held < held_full
    assert_equal(full, grouped, case)
    graphed, _, _ = run_all(torch, plan, k, precision, log_nodes, log_edges, budget, graphs=True)
    assert_equal(full, graphed, case + " (graph replay)")


def test_budget_errors(require_gpu):
    rng = np.random.default_rng(9)
    plan = thmm.Plan(forest([path(400), random_recursive(300, rng), path(500)]))
    k = 8
    ws = thmm.Workspace()
    ws.set_path_budget(1000)  # below the largest tree's path storage
    with pytest.raises(thmm.LengthError, match="alone needs"):
        ws.reserve_marginals(plan, k, 1, device_resident=True)
    ws.set_path_budget(200_000)
    with pytest.raises(thmm.LengthError, match="batch-1 forest"):
        ws.reserve_marginals(plan, k, 2, device_resident=True)
    with pytest.raises(thmm.LengthError, match="batch-1 forest"):
        ws.reserve(plan, k, 1, flags=_abi.THMM_RESERVE_MAXIMUM | _abi.THMM_RESERVE_DEVICE_RESIDENT)
    with pytest.raises(thmm.LengthError, match="batch-1 forest"):
        ws.reserve_marginals(plan, k, 1)  # host-staged
    ws.set_path_budget(10**12)  # fits: not grouped
    ws.reserve_marginals(plan, k, 1, device_resident=True)
    assert ws.path_storage()[0] == 0
    ws.close()
