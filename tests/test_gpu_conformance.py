"""Accelerator conformance on the GPU: the scenarios of
tests/accelerator_test.h (joaospinto/parallel_tree_hmm), with the FP64 oracle
in place of the reference CPU path. Each scenario runs for the FP32 form
(the reference's tolerances) and the plain FP64 form (1e-11)."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
import parallel_tree_hmm_amd as thmm
from conftest import path, star

pytestmark = pytest.mark.gpu

PRECISIONS = ["f32", "f64"]


def tol(precision, f32):
    return f32 if precision == "f32" else 1e-11


def rel_ok(actual, expected, tolerance):
    actual, expected = np.asarray(actual, float), np.asarray(expected, float)
    scale = np.maximum(1.0, np.maximum(np.abs(actual), np.abs(expected)))
    return np.all(np.abs(actual - expected) <= tolerance * scale)


def f32(x):
    return np.asarray(x, np.float32).astype(np.float64)


def cast(x, precision):
    """Inputs as the workspace Scalar would hold them (the oracle then sees
    exactly the same values)."""
    return f32(x) if precision == "f32" else np.asarray(x, np.float64)


# ---------------------------------------------------------------------------
# TestAccelerator (tests/accelerator_test.h:14-274)
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("precision", PRECISIONS)
def test_accelerator_partition(require_gpu, precision):
    plan = thmm.Plan([-1, 0, 0, 1, 1, 2, 2])
    k, batch = 4, 257
    b, n, s = np.meshgrid(np.arange(batch), np.arange(plan.num_nodes), np.arange(k), indexing="ij")
    nodes = cast(0.2 + 0.01 * (1 + b + 2 * n + s), precision)
    p, c = np.meshgrid(np.arange(k), np.arange(k), indexing="ij")
    edges = cast(np.broadcast_to(np.where(p == c, 0.82, 0.06), (plan.num_edges, k, k)), precision)
    ws = thmm.Workspace()
    ws.reserve(plan, k, batch, precision=precision)
    result = thmm.partition_function_prepared(thmm.BatchedModelView(plan, k, batch, nodes, edges), ws)
    assert result.values.shape == (batch,)
    expected = np.array([oracle.partition(plan, k, nodes[i], edges)[0] for i in range(batch)])
    assert rel_ok(result.values, expected, tol(precision, 2e-5))
    assert result.timings.kernel_ms >= 0.0

    # zero-copy staged inputs (:69-87)
    staged = ws.inputs(batch)
    staged.node_potentials[...] = nodes
    staged.edge_potentials[...] = edges
    again = thmm.partition_function_prepared(staged, ws)
    assert rel_ok(again.values, expected, tol(precision, 2e-5))

    # capacity-backed tail batch (:89-114)
    tail = ws.inputs(3)
    tail.node_potentials[...] = nodes[:3]
    tail.edge_potentials[...] = edges
    tail_result = thmm.log_partition_function_prepared(tail, ws)
    assert tail_result.values.shape == (3,)
    expected_log = np.array([oracle.log_partition(plan, k, nodes[i], edges)[0] for i in range(batch)])
    assert rel_ok(tail_result.values, expected_log[:3], tol(precision, 5e-5))

    log_result = thmm.log_partition_function_prepared(thmm.BatchedModelView(plan, k, batch, nodes, edges), ws)
    assert rel_ok(log_result.values, expected_log, tol(precision, 5e-5))
    ws.close()


@pytest.mark.parametrize("precision", PRECISIONS)
def test_accelerator_chain_and_generic_states(require_gpu, precision):
    ws = thmm.Workspace()
    # unary chain with per-edge factors (:135-176)
    plan = thmm.Plan(path(7))
    k, batch = 4, 5
    nodes = cast((0.15 + 0.007 * (np.arange(batch * 7 * k) % 17)).reshape(batch, 7, k), precision)
    e, p, c = np.meshgrid(np.arange(6), np.arange(k), np.arange(k), indexing="ij")
    edges = cast(0.02 + 0.01 * (1 + e + 2 * p + 3 * c), precision)
    ws.reserve(plan, k, batch, precision=precision)
    out = thmm.log_partition_function_prepared(thmm.BatchedModelView(plan, k, batch, nodes, edges), ws)
    expected = [oracle.log_partition(plan, k, nodes[i], edges)[0] for i in range(batch)]
    assert rel_ok(out.values, expected, tol(precision, 5e-5))

    # 4096-node K=1 chain must not underflow (:178-193)
    long_plan = thmm.Plan(path(4096))
    ws.reserve(long_plan, 1, 1, precision=precision)
    out = thmm.log_partition_function_prepared(
        thmm.BatchedModelView(long_plan, 1, 1, np.full((1, 4096, 1), 0.5), np.ones((4095, 1, 1))), ws
    )
    assert abs(out.values[0] - 4096 * np.log(0.5)) <= (2e-3 if precision == "f32" else 1e-9)

    # generic K=9 with path composition (:195-230)
    gplan = thmm.Plan(path(5))
    k, batch = 9, 2
    nodes = cast((0.1 + 0.003 * (np.arange(batch * 5 * k) % 23)).reshape(batch, 5, k), precision)
    edges = cast((0.02 + 0.001 * (np.arange(4 * k * k) % 31)).reshape(4, k, k), precision)
    ws.reserve(gplan, k, batch, precision=precision)
    out = thmm.log_partition_function_prepared(thmm.BatchedModelView(gplan, k, batch, nodes, edges), ws)
    expected = [oracle.log_partition(gplan, k, nodes[i], edges)[0] for i in range(batch)]
    assert rel_ok(out.values, expected, tol(precision, 1e-4))

    # wide star, several sibling-reduction stages (:232-273)
    wplan = thmm.Plan(star(18))
    assert len(wplan.rounds()[0]["branch_reduction_stages"]) >= 4
    k, batch = 3, 3
    nodes = cast((0.14 + 0.006 * (np.arange(batch * 18 * k) % 19)).reshape(batch, 18, k), precision)
    edges = cast((0.03 + 0.004 * (np.arange(17 * k * k) % 23)).reshape(17, k, k), precision)
    ws.reserve(wplan, k, batch, precision=precision)
    out = thmm.log_partition_function_prepared(thmm.BatchedModelView(wplan, k, batch, nodes, edges), ws)
    expected = [oracle.log_partition(wplan, k, nodes[i], edges)[0] for i in range(batch)]
    assert rel_ok(out.values, expected, tol(precision, 1e-4))
    ws.close()


# ---------------------------------------------------------------------------
# TestCategoricalAccelerator (:276-518)
# ---------------------------------------------------------------------------

CAT_PARENTS = [-1, 0, 1, 1, 0, 4, 5, 5]
CAT_OBS_NODES = [2, 3, 6, 7]
CAT_ROOT = [0.1, 0.2, 0.3, 0.4]
CAT_EMISSIONS = [
    [1.0, 1.0, 1.0, 1.0],
    [1.0, 0.0, 0.0, 0.0],
    [0.0, 1.0, 0.0, 0.0],
    [0.0, 0.0, 1.0, 0.0],
    [0.0, 0.0, 0.0, 1.0],
    [1.0, 0.0, 1.0, 0.0],
]


def categorical_case(precision):
    plan = thmm.Plan(CAT_PARENTS)
    k, batch, cats = 4, 7, 6
    obs = np.array([[(b + 2 * o) % cats for o in range(4)] for b in range(batch)], np.uint8)
    p, c = np.meshgrid(np.arange(k), np.arange(k), indexing="ij")
    edges = cast(np.broadcast_to(np.where(p == c, 0.82, 0.06), (plan.num_edges, k, k)), precision)
    root = cast(CAT_ROOT, precision)
    em = cast(CAT_EMISSIONS, precision)
    dtype = np.float32 if precision == "f32" else np.float64
    dense = np.ones((batch, plan.num_nodes, k), dtype)
    dense[:, 0, :] = root.astype(dtype)
    for b in range(batch):
        for o, node in enumerate(CAT_OBS_NODES):
            dense[b, node] = dense[b, node] * em[obs[b, o]].astype(dtype)
    return plan, k, batch, cats, obs, root, em, edges, dense.astype(np.float64)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_categorical_accelerator(require_gpu, precision):
    plan, k, batch, cats, obs, root, em, edges, dense = categorical_case(precision)
    ws = thmm.Workspace()
    ws.reserve_categorical_maximum(plan, k, batch, cats, CAT_OBS_NODES, precision=precision)
    staged = ws.categorical_inputs(batch)
    staged.observations[...] = obs
    staged.root_potential[...] = root
    staged.emission_potentials[...] = em
    staged.edge_potentials[...] = edges
    log_values = thmm.log_partition_function_prepared(staged, ws).values.copy()
    maximum = thmm.maximum_a_posteriori_prepared(staged, ws)
    max_states = maximum.states.reshape(batch, plan.num_nodes).copy()
    max_weights = maximum.log_weights.copy()

    dws = thmm.Workspace()
    dws.reserve(plan, k, batch, precision=precision)
    dense_values = thmm.log_partition_function_prepared(thmm.BatchedModelView(plan, k, batch, dense, edges), dws).values
    for b in range(batch):
        expected = oracle.log_partition(plan, k, dense[b], edges)[0]
        assert rel_ok(log_values[b], expected, tol(precision, 5e-5))
        assert dense_values[b] == log_values[b]  # categorical == dense bitwise (:377-380)
        lw, states = oracle.maximum(plan, k, dense[b], edges)
        assert abs(max_weights[b] - lw[0]) <= (2e-5 if precision == "f32" else 1e-11)
        weight = oracle.assignment_weight(plan, k, dense[b], edges, max_states[b])
        assert abs(np.log(weight) - lw[0]) <= 2e-5  # the assignment attains the optimum
        assert np.array_equal(max_states[b], states)

    tail = ws.categorical_inputs(3)
    assert tail.observations.shape == (3, 4)

    # samples equal the oracle exactly (:422-465)
    ws.reserve_categorical_sampling(plan, k, batch, cats, CAT_OBS_NODES, precision=precision)
    staged = ws.categorical_inputs(batch)
    staged.observations[...] = obs
    staged.root_potential[...] = root
    staged.emission_potentials[...] = em
    staged.edge_potentials[...] = edges
    u = ws.uniforms(batch)
    u[...] = cast((0.13 + 0.12 * (np.arange(u.size) % 7)).reshape(u.shape), precision)
    variates = np.asarray(u, np.float64).copy()
    samples = thmm.posterior_sample_prepared(staged, u, ws).states.reshape(batch, plan.num_nodes)
    for b in range(batch):
        expected, dist = oracle.sample(plan, k, dense[b], edges, variates[b])
        assert np.min(dist) > 1e-12
        assert np.array_equal(samples[b], expected)

    # marginals (:467-517)
    ws.reserve_categorical_marginals(plan, k, batch, cats, CAT_OBS_NODES, precision=precision)
    staged = ws.categorical_inputs(batch)
    staged.observations[...] = obs
    staged.root_potential[...] = root
    staged.emission_potentials[...] = em
    staged.edge_potentials[...] = edges
    m = thmm.posterior_marginals_prepared(staged, ws)
    assert m.log_partitions.shape == (batch,)
    nodes_out = m.nodes.reshape(batch, plan.num_nodes, k)
    edges_out = m.edges.reshape(batch, plan.num_edges, k, k)
    for b in range(batch):
        lz, nm, emg = oracle.marginals(plan, k, dense[b], edges)
        assert abs(m.log_partitions[b] - lz[0]) <= (1e-4 if precision == "f32" else 1e-11) * max(1, abs(lz[0]))
        assert np.max(np.abs(nodes_out[b] - nm)) <= tol(precision, 8e-5)
        assert np.max(np.abs(edges_out[b] - emg)) <= tol(precision, 8e-5)


# ---------------------------------------------------------------------------
# TestCategoricalResidentInputs (:520-650)
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("precision", PRECISIONS)
def test_categorical_resident_inputs(require_gpu, precision):
    plan = thmm.Plan([-1, 0, 0, 1, 1, 2, 2])
    k, batch, cats = 4, 5, 4
    obs_nodes = [3, 4, 5, 6]
    ws = thmm.Workspace()
    ws.reserve_categorical(plan, k, batch, cats, obs_nodes, precision=precision)
    obs = (np.arange(batch * 4) % cats).astype(np.uint8).reshape(batch, 4)
    root = 0.1 * (np.arange(k) + 1)
    em = np.eye(cats, k)
    e, p, c = np.meshgrid(np.arange(plan.num_edges), np.arange(k), np.arange(k), indexing="ij")
    edges = 0.02 + 0.003 * (1 + e + 2 * p + 3 * c)
    model = thmm.BatchedCategoricalModelView(plan, k, batch, cats, np.array(obs_nodes), obs, root, em, edges)
    U = thmm.CategoricalInputUpdate
    for update in (U.kFactors, U.kNone):
        with pytest.raises(thmm.LogicError):
            thmm.log_partition_function_prepared(model, ws, update)
    baseline = thmm.log_partition_function_prepared(model, ws, U.kAll).values.copy()
    obs[0, 0] = (obs[0, 0] + 1) % cats
    reused = thmm.log_partition_function_prepared(model, ws, U.kFactors).values.copy()
    assert np.array_equal(reused, baseline)
    updated = thmm.log_partition_function_prepared(model, ws, U.kAll).values.copy()
    assert not np.array_equal(updated, baseline)
    edges *= 0.9
    factors_reused = thmm.log_partition_function_prepared(model, ws, U.kNone).values.copy()
    assert np.array_equal(factors_reused, updated)
    factors_updated = thmm.log_partition_function_prepared(model, ws, U.kFactors).values.copy()
    assert not np.array_equal(factors_updated, updated)
    tail = thmm.BatchedCategoricalModelView(plan, k, 3, cats, np.array(obs_nodes), obs[:3], root, em, edges)
    with pytest.raises(thmm.LogicError):
        thmm.log_partition_function_prepared(tail, ws, U.kFactors)
    ws.reserve_categorical(plan, k, batch, cats, obs_nodes, precision=precision)
    with pytest.raises(thmm.LogicError):
        thmm.log_partition_function_prepared(model, ws, U.kNone)


# ---------------------------------------------------------------------------
# TestMaximumAccelerator (:652-712), TestSamplingAccelerator (:714-790)
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("precision", PRECISIONS)
def test_maximum_accelerator(require_gpu, precision):
    plan = thmm.Plan([-1, 0, 1, 1, 0, 4, 5, 5])
    k, batch = 4, 5
    b, n, s = np.meshgrid(np.arange(batch), np.arange(plan.num_nodes), np.arange(k), indexing="ij")
    nodes = cast(0.11 + 0.013 * (1 + 3 * b + 5 * n + s), precision)
    e, p, c = np.meshgrid(np.arange(plan.num_edges), np.arange(k), np.arange(k), indexing="ij")
    edges = cast(0.03 + 0.007 * (1 + e + 2 * p + 3 * c), precision)
    ws = thmm.Workspace()
    ws.reserve_maximum(plan, k, batch, precision=precision)
    staged = ws.inputs(batch)
    staged.node_potentials[...] = nodes
    staged.edge_potentials[...] = edges
    out = thmm.maximum_a_posteriori_prepared(staged, ws)
    assert out.log_weights.shape == (batch,) and out.states.shape == (batch * plan.num_nodes,)
    states = out.states.reshape(batch, -1)
    for i in range(batch):
        lw, expected = oracle.maximum(plan, k, nodes[i], edges)
        assert abs(out.log_weights[i] - lw[0]) <= (2e-5 if precision == "f32" else 1e-12)
        assert np.array_equal(states[i], expected)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_sampling_accelerator(require_gpu, precision):
    plan = thmm.Plan([-1, 0, 1, 1, 0, 4, 5, 5])
    k, batch = 4, 5
    b, n, s = np.meshgrid(np.arange(batch), np.arange(plan.num_nodes), np.arange(k), indexing="ij")
    nodes = cast(0.2 + 0.01 * (1 + 2 * b + 3 * n + 5 * s), precision)
    e, p, c = np.meshgrid(np.arange(plan.num_edges), np.arange(k), np.arange(k), indexing="ij")
    edges = cast(0.1 + 0.015 * (1 + e + 2 * p + 4 * c), precision)
    variates = cast((0.07 + 0.11 * (np.arange(batch * plan.num_nodes) % 8)).reshape(batch, -1), precision)
    ws = thmm.Workspace()
    ws.reserve_sampling(plan, k, batch, precision=precision)
    staged = ws.inputs(batch)
    staged.node_potentials[...] = nodes
    staged.edge_potentials[...] = edges
    u = ws.uniforms(batch)
    u[...] = variates
    out = thmm.posterior_sample_prepared(staged, u, ws)
    states = out.states.reshape(batch, -1)
    for i in range(batch):
        expected, dist = oracle.sample(plan, k, nodes[i], edges, variates[i])
        assert np.min(dist) > 1e-12
        assert np.array_equal(states[i], expected)
    u[0, 0] = 1.0
    with pytest.raises(thmm.InvalidArgument):
        thmm.posterior_sample_prepared(staged, u, ws)


# ---------------------------------------------------------------------------
# TestMarginalAccelerator (:792-965)
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("precision", PRECISIONS)
def test_marginal_accelerator(require_gpu, precision):
    plan = thmm.Plan([-1, 0, 1, 1, 0, 4, 5, 5])
    k, batch = 4, 5
    b, n, s = np.meshgrid(np.arange(batch), np.arange(plan.num_nodes), np.arange(k), indexing="ij")
    nodes = cast(np.where((b + 2 * n + 3 * s) % 13 == 0, 0.0, 0.17 + 0.011 * (1 + b + 2 * n + 5 * s)), precision)
    e, p, c = np.meshgrid(np.arange(plan.num_edges), np.arange(k), np.arange(k), indexing="ij")
    edges = cast(np.where((e + 3 * p + 5 * c) % 17 == 0, 0.0, 0.09 + 0.008 * (1 + e + p + 2 * c)), precision)
    ws = thmm.Workspace()
    ws.reserve_marginals(plan, k, batch, precision=precision)
    staged = ws.inputs(batch)
    staged.node_potentials[...] = nodes
    staged.edge_potentials[...] = edges
    out = thmm.posterior_marginals_prepared(staged, ws)
    nm_all = out.nodes.reshape(batch, plan.num_nodes, k)
    em_all = out.edges.reshape(batch, plan.num_edges, k, k)
    for i in range(batch):
        lz, nm, em = oracle.marginals(plan, k, nodes[i], edges)
        assert abs(out.log_partitions[i] - lz[0]) <= (1e-4 if precision == "f32" else 1e-11) * max(1, abs(lz[0]))
        assert np.max(np.abs(nm_all[i] - nm)) <= tol(precision, 8e-5)
        assert np.max(np.abs(em_all[i] - em)) <= tol(precision, 8e-5)
        assert np.all(np.abs(nm_all[i].sum(1) - 1.0) <= tol(precision, 1e-4))

    # root-only plan (:883-908)
    rplan = thmm.Plan([-1])
    root_nodes = cast(np.array([0.2, 0.3, 0.5, 0.1, 0.0, 0.9, 0.4, 0.4, 0.2]).reshape(3, 1, 3), precision)
    ws.reserve_marginals(rplan, 3, 3, precision=precision)
    out = thmm.posterior_marginals_prepared(thmm.BatchedModelView(rplan, 3, 3, root_nodes, np.zeros((0, 3, 3))), ws)
    for i in range(3):
        _, nm, _ = oracle.marginals(rplan, 3, root_nodes[i], np.zeros((0, 3, 3)))
        assert np.max(np.abs(out.nodes.reshape(3, 3)[i] - nm[0])) <= tol(precision, 5e-6)

    # wide-star reverse (:910-965)
    wplan = thmm.Plan(star(18))
    k, batch = 3, 2
    idx = np.arange(batch * 18 * k)
    wnodes = cast(np.where(idx % 29 == 0, 0.0, 0.16 + 0.005 * (idx % 17)).reshape(batch, 18, k), precision)
    wedges = cast((0.04 + 0.003 * (np.arange(17 * k * k) % 31)).reshape(17, k, k), precision)
    ws.reserve_marginals(wplan, k, batch, precision=precision)
    out = thmm.posterior_marginals_prepared(thmm.BatchedModelView(wplan, k, batch, wnodes, wedges), ws)
    for i in range(batch):
        _, nm, em = oracle.marginals(wplan, k, wnodes[i], wedges)
        assert np.max(np.abs(out.nodes.reshape(batch, 18, k)[i] - nm)) <= tol(precision, 8e-5)
        assert np.max(np.abs(out.edges.reshape(batch, 17, k, k)[i] - em)) <= tol(precision, 8e-5)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_error_paths(require_gpu, precision):
    plan = thmm.Plan([-1, 0, 0])
    other = thmm.Plan([-1, 0, 0])
    ws = thmm.Workspace()
    ws.reserve(plan, 2, 4, precision=precision)
    nodes = np.ones((2, 3, 2))
    edges = np.ones((2, 2, 2))
    with pytest.raises(thmm.InvalidArgument):  # wrong plan object
        thmm.log_partition_function_prepared(thmm.BatchedModelView(other, 2, 2, nodes, edges), ws)
    with pytest.raises(thmm.InvalidArgument):  # batch above capacity
        thmm.log_partition_function_prepared(thmm.BatchedModelView(plan, 2, 5, np.ones((5, 3, 2)), edges), ws)
    with pytest.raises(thmm.InvalidArgument):  # MAP needs ReserveMaximum
        thmm.maximum_a_posteriori_prepared(thmm.BatchedModelView(plan, 2, 2, nodes, edges), ws)
    with pytest.raises(thmm.InvalidArgument):  # marginals need ReserveMarginals
        thmm.posterior_marginals_prepared(thmm.BatchedModelView(plan, 2, 2, nodes, edges), ws)
    with pytest.raises(thmm.LogicError):
        ws.uniforms(2)
    ws.reserve_marginals(plan, 2, 4, precision=precision)
    with pytest.raises(thmm.DomainError):  # nonpositive partition function
        thmm.posterior_marginals_prepared(thmm.BatchedModelView(plan, 2, 2, np.zeros((2, 3, 2)), edges), ws)
    ws.reserve_categorical(plan, 2, 2, 2, [1])
    with pytest.raises(thmm.LogicError):  # dense Inputs on a categorical workspace
        ws.inputs(2)
