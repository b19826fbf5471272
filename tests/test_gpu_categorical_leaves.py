"""Categorical leaf fusion (thmm_tuning.input_leaves) and the
site-summed log likelihood (thmm_categorical_log_likelihood).

With leaf fusion on (the default) the rake kernels of a categorical
partition, marginal or branch-length call form the leaf vectors of the leaf
rakes from the observations, emissions and root potential, and the node
initialisation skips those leaves (the reference writes every node's factor
first: src/gpu_backend_impl.inc:2917-2941, :326-382). Every case compares
the fused form with the unfused one (input_leaves = 0) on the same
workspace and inputs, bit for bit:

* log partitions (Scalar and FP64), node and edge marginals;
* the branch-length pair: per-item log Z and d/dt_e with seeds;
* the dense form on the materialised node factors (the unfused dense
  initialisation) gives the same log partitions bit for bit.

The site sum equals the FP64 sum, in the documented fixed order, of the
per-item log partitions of the same inputs, and its backward with no seed equals the
pair's backward with no seed after thmm_categorical_logz_forward.

Shapes: K in {2, 3, 4, 5, 7} (the generic rake kernels) and K = 8, 16 (no
fusion: the team kernels; the flag changes nothing), FP32 and FP64, batch
37 and 1 (at K = 4 and batch 1 the dense initialisation takes FP32 logs, and
so does the categorical one), a
forest of a random recursive tree, a star, a path, a two-node tree and a
single node, observations on leaves, internal nodes and a root, emission
rows with zeros, and an out-of-range category (zero factor). The full
categorical workload (8,191-node balanced binary tree, 1,024 sites, K = 4,
FP32, device expm) is compared the same way and both forms are timed.
"""

from __future__ import annotations

import time

import numpy as np
import pytest

import bench
import parallel_tree_hmm_amd as thmm
from conftest import forest, path, random_recursive, reversible_rate_matrix, star, transition_matrices
from parallel_tree_hmm_amd.api import BatchedCategoricalModelView, CategoricalInputUpdate

pytestmark = pytest.mark.gpu

DTYPES = {"f32": np.float32, "f64": np.float64}


def small_case(k, seed, batch):
    rng = np.random.default_rng(seed)
    parents = forest([random_recursive(60, rng), star(9), path(7), np.array([-1, 0]), np.array([-1])])
    n = parents.size
    children = np.bincount(parents[parents >= 0], minlength=n)
    leaves = np.flatnonzero(children == 0)
    internal = np.flatnonzero(children > 0)
    roots = np.flatnonzero(parents < 0)
    # most leaves, a few internal nodes and one root observed
    observed = np.union1d(np.union1d(leaves[rng.random(leaves.size) < 0.85], internal[:: 7]), roots[:1])
    categories = 6
    obs = rng.integers(0, categories, size=(batch, observed.size)).astype(np.uint8)
    root = rng.dirichlet(np.ones(k))
    emission = rng.random((categories, k)) + 0.05
    emission[0] = 0.0
    emission[0, 0] = 1.0  # a row with zeros (a state ruled out)
    emission[categories - 1] = 1.0  # missing data
    q = reversible_rate_matrix(rng, k)
    t = rng.exponential(0.1, size=n - roots.size)
    return dict(parents=parents, observed=observed.astype(np.uint32), obs=obs, categories=categories, root=root,
                emission=emission, q=q, t=t, leaves=leaves)


def model_of(plan, k, c, dtype, batch, obs=None, factors=None):
    if factors is None:
        factors = transition_matrices(c["q"], c["t"])
    return BatchedCategoricalModelView(plan, k, batch, c["categories"], c["observed"], c["obs"] if obs is None else obs,
                                       c["root"].astype(dtype), c["emission"].astype(dtype), factors.astype(dtype))


def run_all(ws, model, q, seeds):
    out = {}
    view = thmm.log_partition_function_prepared(model, ws)
    out["log_partition"] = view.values.copy()
    out["log_partition_f64"] = ws.log_partitions_f64()[: model.batch * model.plan.roots.size].copy()
    m = thmm.posterior_marginals_prepared(model, ws)
    out["m_log_z"] = m.log_partitions.copy()
    out["m_nodes"] = m.nodes.copy()
    out["m_edges"] = m.edges.copy()
    out["pair_log_z"] = ws.categorical_logz_forward(model)
    out["pair_grad"] = ws.branch_length_backward(q, seeds)
    ws.categorical_logz_forward(model)
    out["pair_grad_ones"] = ws.branch_length_backward(q, None)
    out["sum"] = ws.categorical_log_likelihood(model)
    out["sum_grad"] = ws.branch_length_backward(q, None)
    return out


def batch_sum(values):
    """The device's fixed summation order of thmm_categorical_log_likelihood
    over one root's per-item values (Python floats: IEEE FP64 adds)."""
    part = [0.0] * 256
    for t in range(256):
        for b in range(t, len(values), 256):
            part[t] += float(values[b])
    s = 128
    while s > 0:
        for t in range(s):
            part[t] += part[t + s]
        s //= 2
    return part[0]


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and a.tobytes() == b.tobytes()


@pytest.mark.parametrize("batch", [37, 1])
@pytest.mark.parametrize("precision", ["f32", "f64"])
@pytest.mark.parametrize("k", [2, 3, 4, 5, 7, 8, 16])
def test_fused_leaves_equal_the_unfused_form(require_gpu, k, precision, batch):
    c = small_case(k, 100 + k, batch)
    plan = thmm.Plan(c["parents"])
    dtype = DTYPES[precision]
    ws = thmm.Workspace()
    ws.reserve_categorical_marginals(plan, k, batch, c["categories"], c["observed"], precision=precision)
    model = model_of(plan, k, c, dtype, batch)
    seeds = np.random.default_rng(k).normal(size=batch * plan.roots.size)
    assert ws.tuning()["input_leaves"] == 1
    fused = run_all(ws, model, c["q"], seeds)
    ws.set_tuning(input_leaves=0)
    plain = run_all(ws, model, c["q"], seeds)
    for name in fused:
        assert same(fused[name], plain[name]), f"K={k} {precision}: {name} differs"
    # The site sum is the fixed-order FP64 sum of the per-item log
    # partitions; its unseeded gradient is the pair's.
    lz = fused["pair_log_z"].reshape(batch, plan.roots.size)
    for r in range(plan.roots.size):
        assert fused["sum"][r] == batch_sum(lz[:, r]), r
    assert same(fused["sum_grad"], fused["pair_grad_ones"])
    # The dense form on the materialised factors (unfused dense
    # initialisation) gives the same log partitions.
    factors = transition_matrices(c["q"], c["t"]).astype(dtype)
    dense = np.ones((batch, plan.num_nodes, k), dtype)
    dense[:, plan.roots, :] = c["root"].astype(dtype)
    em = c["emission"].astype(dtype)
    dense[:, c["observed"], :] = dense[:, c["observed"], :] * em[c["obs"]]
    dws = thmm.Workspace()
    dws.reserve(plan, k, batch, precision=precision)
    dense_values = thmm.log_partition_function_prepared(thmm.BatchedModelView(plan, k, batch, dense, factors), dws)
    assert same(dense_values.values, fused["log_partition"])
    dws.close()
    ws.close()


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_zero_factor_leaves(require_gpu, precision):
    """An out-of-range category gives a zero leaf factor: log Z = -inf for
    that item in both forms, and the other items are unchanged."""
    k, batch = 4, 6
    c = small_case(k, 7, batch)
    plan = thmm.Plan(c["parents"])
    dtype = DTYPES[precision]
    obs = c["obs"].copy()
    j = int(np.flatnonzero(np.isin(c["observed"], c["leaves"]))[0])  # an observed leaf of the first tree
    assert c["observed"][j] < 60
    obs[2, j] = 200  # >= categories
    ws = thmm.Workspace()
    ws.reserve_categorical(plan, k, batch, c["categories"], c["observed"], precision=precision)
    model = model_of(plan, k, c, dtype, batch, obs=obs)
    fused = thmm.log_partition_function_prepared(model, ws).values.copy()
    ws.set_tuning(input_leaves=0)
    plain = thmm.log_partition_function_prepared(model, ws).values.copy()
    assert same(fused, plain)
    per_root = fused.reshape(batch, plan.roots.size)
    assert np.isneginf(per_root[2]).sum() == 1 and np.all(np.isfinite(np.delete(per_root, 2, axis=0)))
    ws.close()


def test_log_likelihood_errors(require_gpu):
    k, batch = 4, 3
    c = small_case(k, 9, batch)
    plan = thmm.Plan(c["parents"])
    model = model_of(plan, k, c, np.float32, batch)
    ws = thmm.Workspace()
    ws.reserve_categorical(plan, k, batch, c["categories"], c["observed"])  # no marginal storage
    with pytest.raises(thmm.LogicError):
        ws.categorical_log_likelihood(model)
    ws.reserve_categorical_marginals(plan, k, batch, c["categories"], c["observed"])
    obs = c["obs"].copy()
    obs[1, 0] = 200
    with pytest.raises(thmm.DomainError, match="nonpositive partition function"):
        ws.categorical_log_likelihood(model_of(plan, k, c, np.float32, batch, obs=obs))
    with pytest.raises(thmm.LogicError):
        ws.branch_length_backward(c["q"], None)  # the failed forward holds no state
    ws.categorical_log_likelihood(model)
    ws.set_tuning(input_leaves=0)  # drops the held forward pass
    with pytest.raises(thmm.LogicError):
        ws.branch_length_backward(c["q"], None)
    ws.close()


def test_full_categorical_workload_fused_equals_unfused(require_gpu):
    sites, k, categories = 1024, 4, 5
    parents, leaves, obs, root, emission, q, pi, t = bench.phylo_inputs(sites)
    plan = thmm.Plan(parents)
    factors = thmm.expm_reversible(q, pi, t).astype(np.float32)
    model = BatchedCategoricalModelView(plan, k, sites, categories, leaves, obs, root.astype(np.float32),
                                        emission.astype(np.float32), factors)
    ws = thmm.Workspace()
    ws.reserve_categorical_marginals(plan, k, sites, categories, leaves, precision="f32")
    ws.categorical_logz_forward(model)  # stages observations and factors
    ws.categorical_set_branch_lengths(q, pi, t)
    seeds = np.linspace(0.5, 1.5, sites)

    def step(summed):
        if summed:
            s = ws.categorical_log_likelihood(model, CategoricalInputUpdate.kNone)
        else:
            s = ws.categorical_logz_forward(model, CategoricalInputUpdate.kNone)
        return s, ws.branch_length_backward(q, None if summed else seeds)

    results, times = {}, {}
    for leaves_flag in (1, 0):
        ws.set_tuning(input_leaves=leaves_flag)
        for summed in (False, True):
            out = step(summed)
            step(summed)
            start = time.perf_counter()
            for _ in range(10):
                step(summed)
            times[(leaves_flag, summed)] = (time.perf_counter() - start) / 10 * 1e3
            results[(leaves_flag, summed)] = out
    print("ms per forward + backward step (fused / unfused leaves): per-site log Z "
          f"{times[(1, False)]:.3f} / {times[(0, False)]:.3f}; site-summed "
          f"{times[(1, True)]:.3f} / {times[(0, True)]:.3f}", flush=True)
    for summed in (False, True):
        for a, b in zip(results[(1, summed)], results[(0, summed)]):
            assert same(a, b), summed
    lz = results[(1, False)][0]
    assert results[(1, True)][0][0] == batch_sum(lz)
    assert np.all(np.isfinite(results[(1, True)][1]))
    ws.close()
