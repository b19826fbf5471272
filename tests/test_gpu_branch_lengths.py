"""The branch-length gradient pair (thmm_categorical_logz_forward /
thmm_branch_length_backward) against the oracle's chain over its edge
marginals (the chain rule itself is pinned by finite differences in
test_branch_length_chain.py), at a small size and at the full categorical
workload: one balanced binary tree with 4,096 leaves (8,191 nodes), 1,024
sites, K = 4, 5 % missing leaves, FP32 with FP64 log scales. Edge factors
formed on the device from branch lengths (thmm_categorical_set_branch_lengths)
give what the same factors staged from the host give.

Tolerances: FP64 workspaces 1e-9 relative; FP32 workspaces 1e-4 relative to
the largest gradient entry (each entry sums FP32 edge marginals of every
site, weighted by (Q P)/P, which reaches 1/t on off-diagonal entries)."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
import parallel_tree_hmm_amd as thmm
from conftest import dense_categorical_nodes, phylo_model, stationary_of, transition_matrices
from parallel_tree_hmm_amd.api import BatchedCategoricalModelView, CategoricalInputUpdate

pytestmark = pytest.mark.gpu

DTYPES = {"f32": np.float32, "f64": np.float64}
GRAD_TOL = {"f32": 1e-4, "f64": 1e-9}


def oracle_chain(plan, q, p, nodes, seeds):
    """sum_b seeds[b] sum_{p,c} M_be[p][c] (Q P_e)[p][c] / P_e[p][c], FP64,
    over the sites with a nonzero seed."""
    ratio = np.matmul(q, p) / p
    grad = np.zeros(p.shape[0])
    log_z = {}
    for b in np.flatnonzero(seeds):
        lz, _, em = oracle.marginals(plan, q.shape[0], nodes[b], p)
        log_z[int(b)] = lz[0]
        grad += seeds[b] * np.einsum("epc,epc->e", em, ratio)
    return grad, log_z


def reserve(plan, leaves, sites, precision):
    ws = thmm.Workspace()
    ws.reserve_categorical_marginals(plan, 4, sites, 5, leaves, precision=precision)
    return ws


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_branch_length_gradient_matches_oracle_chain(require_gpu, precision):
    rng = np.random.default_rng(41)
    parents, leaves, obs, root, emission, q, t = phylo_model(rng, 5, 12)
    plan = thmm.Plan(parents)
    dtype = DTYPES[precision]
    p = transition_matrices(q, t).astype(dtype)
    ws = reserve(plan, leaves, 12, precision)
    model = BatchedCategoricalModelView(plan, 4, 12, 5, leaves, obs, root, emission, p)
    log_z = ws.categorical_logz_forward(model)
    seeds = rng.normal(size=12)
    grad = ws.branch_length_backward(q, seeds)

    # The oracle sees exactly the staged values (rounded to the workspace type).
    nodes = dense_categorical_nodes(parents.size, leaves, obs, root.astype(dtype).astype(np.float64),
                                    emission.astype(dtype).astype(np.float64))
    expected, oracle_log_z = oracle_chain(plan, q, p.astype(np.float64), nodes, seeds)
    for b, lz in oracle_log_z.items():
        assert abs(log_z[b] - lz) <= 1e-5 * max(1.0, abs(lz))
    scale = max(1.0, float(np.max(np.abs(expected))))
    assert np.max(np.abs(grad - expected)) <= GRAD_TOL[precision] * scale

    # A second backward reuses the same forward state; ones = no seeds.
    ones = ws.branch_length_backward(q, None)
    expected_ones, _ = oracle_chain(plan, q, p.astype(np.float64), nodes, np.ones(12))
    assert np.max(np.abs(ones - expected_ones)) <= GRAD_TOL[precision] * max(1.0, np.max(np.abs(expected_ones)))


def test_backward_needs_its_forward(require_gpu):
    rng = np.random.default_rng(42)
    parents, leaves, obs, root, emission, q, t = phylo_model(rng, 3, 2)
    plan = thmm.Plan(parents)
    ws = reserve(plan, leaves, 2, "f64")
    model = BatchedCategoricalModelView(plan, 4, 2, 5, leaves, obs, root, emission, transition_matrices(q, t))
    with pytest.raises(thmm.LogicError):
        ws.branch_length_backward(q)
    ws.categorical_logz_forward(model)
    # Any other call overwrites the contraction state.
    thmm.log_partition_function_prepared(model, ws)
    with pytest.raises(thmm.LogicError):
        ws.branch_length_backward(q)
    with pytest.raises(thmm.InvalidArgument):
        ws.categorical_logz_forward(model)
        ws.branch_length_backward(np.eye(3))


def test_autograd_harness_and_stale_state(require_gpu):
    import torch

    from parallel_tree_hmm_amd.autograd import branch_length_log_partition

    rng = np.random.default_rng(43)
    parents, leaves, obs, root, emission, q, t = phylo_model(rng, 4, 6)
    plan = thmm.Plan(parents)
    ws = reserve(plan, leaves, 6, "f64")
    model = BatchedCategoricalModelView(plan, 4, 6, 5, leaves, obs, root, emission, transition_matrices(q, t))
    tt = torch.tensor(t, dtype=torch.float64, requires_grad=True)
    weights = torch.tensor(rng.normal(size=6), dtype=torch.float64)
    (branch_length_log_partition(tt, q, model, ws) * weights).sum().backward()
    nodes = dense_categorical_nodes(parents.size, leaves, obs, root, emission)
    expected, _ = oracle_chain(plan, q, transition_matrices(q, t), nodes, weights.numpy())
    assert np.max(np.abs(tt.grad.numpy() - expected)) <= 1e-9 * max(1.0, np.max(np.abs(expected)))

    # kFactors reuse of the staged observations gives the same values.
    first = branch_length_log_partition(tt, q, model, ws, CategoricalInputUpdate.kFactors)
    second = branch_length_log_partition(tt.detach() * 1.01, q, model, ws, CategoricalInputUpdate.kFactors)
    with pytest.raises(thmm.LogicError):
        first.sum().backward()
    fresh = thmm.Workspace()
    fresh.reserve_categorical_marginals(plan, 4, 6, 5, leaves, precision="f64")
    again = branch_length_log_partition(tt.detach() * 1.01, q, model, fresh)
    assert torch.equal(second, again)


def test_branch_length_gradient_full_workload(require_gpu):
    """4,096 leaves, 1,024 sites, FP32: log Z against the FP64 form and the
    oracle; the gradient seeded on 16 sites against the oracle's chain over
    those sites; the all-sites gradient against the FP64 form."""
    rng = np.random.default_rng(0x54484D4D00000003)
    sites = 1024
    parents, leaves, obs, root, emission, q, t = phylo_model(rng, 12, sites)
    plan = thmm.Plan(parents)
    assert plan.num_nodes == 8191 and leaves.size == 4096
    p64 = transition_matrices(q, t)
    p32 = p64.astype(np.float32)
    ws32 = reserve(plan, leaves, sites, "f32")
    ws64 = reserve(plan, leaves, sites, "f64")
    model32 = BatchedCategoricalModelView(plan, 4, sites, 5, leaves, obs, root, emission, p32)
    # FP64 form on the same (float-rounded) factors.
    model64 = BatchedCategoricalModelView(plan, 4, sites, 5, leaves, obs, root.astype(np.float32),
                                          emission, p32.astype(np.float64))
    lz32 = ws32.categorical_logz_forward(model32)
    lz64 = ws64.categorical_logz_forward(model64)
    assert np.max(np.abs(lz32 - lz64) / np.maximum(1.0, np.abs(lz64))) <= 1e-5

    picked = rng.choice(sites, size=16, replace=False)
    seeds = np.zeros(sites)
    seeds[picked] = rng.normal(size=16)
    grad32 = ws32.branch_length_backward(q, seeds)
    nodes = dense_categorical_nodes(parents.size, leaves, obs[picked], root.astype(np.float32).astype(np.float64),
                                    emission)
    expected = np.zeros(plan.num_edges)
    ratio = np.matmul(q, p32.astype(np.float64)) / p32.astype(np.float64)
    for i, b in enumerate(picked):
        lz, _, em = oracle.marginals(plan, 4, nodes[i], p32.astype(np.float64))
        assert abs(lz32[b] - lz[0]) <= 1e-5 * max(1.0, abs(lz[0]))
        expected += seeds[b] * np.einsum("epc,epc->e", em, ratio)
    assert np.max(np.abs(grad32 - expected)) <= 1e-4 * max(1.0, np.max(np.abs(expected)))

    all32 = ws32.branch_length_backward(q, None)
    all64 = ws64.branch_length_backward(q, None)
    assert np.max(np.abs(all32 - all64)) <= 1e-4 * max(1.0, np.max(np.abs(all64)))


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_device_branch_lengths_match_staged_factors(require_gpu, precision):
    rng = np.random.default_rng(44)
    sites = 32
    parents, leaves, obs, root, emission, q, t = phylo_model(rng, 6, sites)
    pi = stationary_of(q)
    plan = thmm.Plan(parents)
    dtype = DTYPES[precision]
    seeds = rng.normal(size=sites)

    def staged(lengths):
        ws = reserve(plan, leaves, sites, precision)
        factors = thmm.expm_reversible(q, pi, lengths).astype(dtype)
        model = BatchedCategoricalModelView(plan, 4, sites, 5, leaves, obs, root, emission, factors)
        return ws.categorical_logz_forward(model), ws.branch_length_backward(q, seeds), model

    ws = reserve(plan, leaves, sites, precision)
    with pytest.raises(thmm.LogicError):  # nothing resident yet
        ws.categorical_set_branch_lengths(q, pi, t)
    first_log_z, first_grad, model = staged(t)
    assert np.array_equal(ws.categorical_logz_forward(model), first_log_z)

    longer = t * 1.3 + 0.01
    ws.categorical_set_branch_lengths(q, pi, longer)
    with pytest.raises(thmm.LogicError):  # the held forward pass was dropped
        ws.branch_length_backward(q, seeds)
    log_z = ws.categorical_logz_forward(model, CategoricalInputUpdate.kNone)
    grad = ws.branch_length_backward(q, seeds)
    want_log_z, want_grad, _ = staged(longer)
    tol = 1e-12 if precision == "f64" else 1e-6
    assert np.max(np.abs(log_z - want_log_z) / np.maximum(1.0, np.abs(want_log_z))) <= tol
    assert np.max(np.abs(grad - want_grad)) <= 10 * tol * max(1.0, np.max(np.abs(want_grad)))
    assert not np.allclose(log_z, first_log_z)

    # kFactors restages the host edges (the first lengths' factors).
    again = ws.categorical_logz_forward(model, CategoricalInputUpdate.kFactors)
    assert np.array_equal(again, first_log_z)
    with pytest.raises(thmm.InvalidArgument):
        ws.categorical_set_branch_lengths(q, pi, longer[:-1])
    with pytest.raises(thmm.InvalidArgument):
        ws.categorical_set_branch_lengths(q, pi, -longer)
