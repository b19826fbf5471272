"""Launch variants (thmm_set_tuning) and the autotuner: every variant gives
the same marginals (bit-exact where only the launch shape changes; within
1e-6 where the fused forward's chunking reorders FP64 sums or compression
grouping reorders FP32 products), invalid values are rejected, and the
autotuner leaves a valid setting and valid outputs."""

from __future__ import annotations

import numpy as np
import pytest

import parallel_tree_hmm_amd as thmm
from conftest import caterpillar, forest, random_recursive, star

pytestmark = pytest.mark.gpu


def setup(torch, k=16):
    rng = np.random.default_rng(21)
    plan = thmm.Plan(forest([star(3000), caterpillar(1001), random_recursive(2000, rng)]))
    nodes = torch.tensor(np.exp(rng.normal(size=(1, plan.num_nodes, k))), dtype=torch.float32, device="cuda")
    rows = rng.exponential(1.0, size=(plan.num_edges, k, k))
    edges = torch.tensor(rows / rows.sum(axis=2, keepdims=True), dtype=torch.float32, device="cuda")
    ws = thmm.Workspace()
    ws.reserve_marginals(plan, k, 1, precision="f32", device_resident=True)
    return plan, nodes, edges, ws


def run(torch, ws, plan, nodes, edges, k=16):
    lz = torch.empty(plan.roots.size, dtype=torch.float64, device="cuda")
    nm = torch.empty((plan.num_nodes, k), device="cuda")
    em = torch.empty((plan.num_edges, k, k), device="cuda")
    ws.marginals_device(nodes.data_ptr(), edges.data_ptr(), 1, lz.data_ptr(), nm.data_ptr(), em.data_ptr(), 0)
    torch.cuda.synchronize()
    return lz, nm, em


def test_variants_agree(require_gpu):
    import torch

    plan, nodes, edges, ws = setup(torch)
    assert ws.tuning() == {"compress_items_forward": 2, "compress_items_reverse": 2, "rake_block": 256,
                           "gather_chunk": 16, "compress_depth": 2, "rake_log_edges": 1, "persistent": 0,
                           "input_leaves": 1}
    base = run(torch, ws, plan, nodes, edges)
    # Launch shapes only: bit-exact. The binary compression kernels'
    # variants are seen with grouping off (depth 1).
    for depth in (2, 1):
        ws.set_tuning(compress_depth=depth)
        same = run(torch, ws, plan, nodes, edges)
        for fields in ({"compress_items_forward": 1}, {"compress_items_reverse": 1}, {"rake_block": 64},
                       {"rake_block": 128}):
            ws.set_tuning(**fields)
            for a, b in zip(same, run(torch, ws, plan, nodes, edges)):
                assert torch.equal(a, b), (depth, fields)
            ws.set_tuning(compress_items_forward=2, compress_items_reverse=2, rake_block=256)
    # Reordered FP64 sums (chunk size) or FP32 products (grouping depth).
    for fields in ({"gather_chunk": 1}, {"gather_chunk": 4}, {"gather_chunk": 32}, {"gather_chunk": 64},
                   {"compress_depth": 1}, {"compress_depth": 3}):
        ws.set_tuning(**{"gather_chunk": 16, "compress_depth": 2, **fields})
        lz, nm, em = run(torch, ws, plan, nodes, edges)
        assert float(((lz - base[0]).abs() / base[0].abs().clamp(min=1)).max()) <= 1e-6, fields
        assert float((nm - base[1]).abs().max()) <= 1e-6, fields
        assert float((em - base[2]).abs().max()) <= 1e-6, fields
    for bad in ({"compress_items_forward": 3}, {"rake_block": 96}, {"gather_chunk": 0}, {"gather_chunk": 65},
                {"compress_depth": 0}, {"compress_depth": 4}, {"rake_log_edges": 2}, {"persistent": 2},
                {"input_leaves": 2}):
        with pytest.raises(thmm.InvalidArgument):
            ws.set_tuning(**bad)
    # persistent passes are switched off in this build
    with pytest.raises(thmm.LogicError, match="switched off"):
        ws.set_tuning(persistent=1)


def test_autotuner_keeps_a_valid_setting(require_gpu):
    import torch

    plan, nodes, edges, ws = setup(torch)
    base = run(torch, ws, plan, nodes, edges)
    lz = torch.empty(plan.roots.size, dtype=torch.float64, device="cuda")
    nm = torch.empty((plan.num_nodes, 16), device="cuda")
    em = torch.empty((plan.num_edges, 16, 16), device="cuda")
    chosen = ws.autotune_marginals_device(nodes.data_ptr(), edges.data_ptr(), 1, lz.data_ptr(), nm.data_ptr(),
                                          em.data_ptr(), 0)
    torch.cuda.synchronize()
    print(f"autotuned: {chosen}")
    assert chosen["ms"] > 0
    assert {k: v for k, v in chosen.items() if k != "ms"} == ws.tuning()
    assert float((nm - base[1]).abs().max()) <= 1e-6 and float((em - base[2]).abs().max()) <= 1e-6
    again = run(torch, ws, plan, nodes, edges)
    assert float((again[1] - base[1]).abs().max()) <= 1e-6
