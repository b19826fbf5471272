"""Max-product compression pieces (MAP reservations at K = 8, 16, 32).

A piece evaluates a root compression together with the input compressions
absorbed into it, holding their paths in registers. Each compression keeps
the level-by-level arithmetic ((L + mid) + R, strict '>' over ascending
middle states), so with pieces on (compress_depth 2, the default) MAP
states, log weights and every choice the backtrack reads must equal the
plain level form (compress_depth 1) bit for bit, and the FP64 oracle
(src/inference.cc:598-639 compression, :651-667 backtrack).

Topologies with long unary chains (paths, caterpillars) give deep
compression trees, so pieces of three compressions, of two (one input
absorbed on either side) and single roots all occur; potentials drawn from a
few values make ties in the middle-state scans common.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
import parallel_tree_hmm_amd as thmm
from parallel_tree_hmm_amd import _abi
from conftest import balanced_binary, caterpillar, forest, path, random_potentials, random_recursive, star

pytestmark = pytest.mark.gpu


def pieces_plan(rng):
    return thmm.Plan(forest([path(257), caterpillar(301), random_recursive(400, rng), balanced_binary(127),
                             star(40), path(2), path(3), path(5)]))


def run_map(torch, ws, plan, k, batch, d_nodes, d_edges):
    weights = torch.empty(batch * plan.roots.size, dtype=torch.float64, device="cuda")
    states = torch.empty((batch, plan.num_nodes), dtype=torch.int32, device="cuda")
    ws.maximum_device(d_nodes.data_ptr(), d_edges.data_ptr(), batch, weights.data_ptr(), states.data_ptr(), 0)
    torch.cuda.synchronize()
    return states.cpu().numpy().view(np.uint32).copy(), weights.cpu().numpy().copy()


@pytest.mark.parametrize("precision", ["f32", "f64"])
@pytest.mark.parametrize("k", [8, 16, 32])
@pytest.mark.parametrize("ties", [False, True])
def test_pieces_equal_plain_levels_and_oracle(require_gpu, precision, k, ties):
    import torch

    rng = np.random.default_rng(100 * k + 10 * ties + (precision == "f64"))
    plan = pieces_plan(rng)
    batch = 3
    if ties:
        # Few distinct values: equal candidates in most scans.
        choices = np.array([0.25, 0.5, 1.0, 2.0])
        nodes = rng.choice(choices, size=(batch, plan.num_nodes, k))
        edges = rng.choice(choices, size=(plan.num_edges, k, k))
    else:
        nodes, edges = random_potentials(rng, batch, plan.num_nodes, plan.num_edges, k, zero_fraction=0.1,
                                         sigma=2.0)
    dtype = np.float32 if precision == "f32" else np.float64
    nodes, edges = nodes.astype(dtype), edges.astype(dtype)
    tdtype = torch.float32 if precision == "f32" else torch.float64
    d_nodes = torch.tensor(nodes, dtype=tdtype, device="cuda")
    d_edges = torch.tensor(edges, dtype=tdtype, device="cuda")
    ws = thmm.Workspace()
    ws.reserve(plan, k, batch, precision=precision,
               flags=_abi.THMM_RESERVE_MAXIMUM | _abi.THMM_RESERVE_DEVICE_RESIDENT)
    assert ws.tuning()["compress_depth"] == 2
    with_pieces = run_map(torch, ws, plan, k, batch, d_nodes, d_edges)
    ws.set_tuning(compress_depth=1)
    plain = run_map(torch, ws, plan, k, batch, d_nodes, d_edges)
    assert np.array_equal(with_pieces[0], plain[0])
    assert np.array_equal(with_pieces[1], plain[1])
    for b in range(batch):
        lw, expected = oracle.maximum(plan, k, nodes[b].astype(np.float64), edges.astype(np.float64))
        assert np.array_equal(with_pieces[0][b], expected)
        got = with_pieces[1].reshape(batch, -1)[b]
        assert np.all(np.abs(got - lw) <= 1e-12 * np.maximum(1, np.abs(lw)))


@pytest.mark.parametrize("k", [8, 16])
def test_pieces_host_staged_call(require_gpu, k):
    """The host-staged MAP call runs the same pieces: states equal the
    oracle's and the device call's."""
    rng = np.random.default_rng(7 + k)
    plan = pieces_plan(rng)
    batch = 2
    nodes, edges = random_potentials(rng, batch, plan.num_nodes, plan.num_edges, k, zero_fraction=0.1, sigma=2.0)
    nodes, edges = nodes.astype(np.float32), edges.astype(np.float32)
    ws = thmm.Workspace()
    ws.reserve_maximum(plan, k, batch)
    out = thmm.maximum_a_posteriori_prepared(thmm.BatchedModelView(plan, k, batch, nodes, edges), ws)
    for b in range(batch):
        _, expected = oracle.maximum(plan, k, nodes[b].astype(np.float64), edges.astype(np.float64))
        assert np.array_equal(np.asarray(out.states).reshape(batch, -1)[b], expected)
