"""Pins the FP64 oracle (CPU) against what fixes it independently:

* exhaustive enumeration (Z, node/edge marginals, MAP weight) on random tiny
  trees and forests (method of tests/inference_test.cc:55-144);
* the literal fixtures of tests/inference_test.cc:159-193 with expected
  values from tests/golden/make_golden.py (numpy enumeration);
* closed forms of tests/inference_test.cc:286-341;
* marginals == d log Z / d log psi by central finite differences;
* exact sample states against an exhaustive conditional sampler that
  follows the plan's expansion order, and the 30,000-draw frequency test of
  tests/inference_test.cc:249-274 with its LCG;
* tie-breaking (tests/inference_test.cc:210-217) and invalid variates
  (:276-284).

Each check is chosen so that a plausible slip (a dropped factor, a
transposed edge matrix, a wrong association, an off-by-one in the inverse
CDF) fails it.
"""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

import oracle
import parallel_tree_hmm_amd as thmm
from conftest import forest, path, random_recursive, random_potentials, star

GOLDEN = json.loads((Path(__file__).parent / "golden" / "inference_fixtures.json").read_text())


def near(a, b, tol):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.all(np.abs(a - b) <= tol * np.maximum(1.0, np.maximum(np.abs(a), np.abs(b))))


@pytest.mark.parametrize("case", sorted(GOLDEN))
def test_literal_fixtures(case):
    c = GOLDEN[case]
    plan = thmm.Plan(c["parents"])
    k = c["states"]
    exp = c["expected"]
    assert near(oracle.partition(plan, k, c["nodes"], c["edges"])[0], exp["partition"], 1e-11)
    assert near(oracle.log_partition(plan, k, c["nodes"], c["edges"])[0], exp["log_partition"], 1e-11)
    log_z, nm, em = oracle.marginals(plan, k, c["nodes"], c["edges"])
    assert near(log_z[0], exp["log_partition"], 1e-11)
    assert near(nm.ravel(), exp["node_marginals"], 1e-11)
    assert near(em.ravel(), exp["edge_marginals"], 1e-11)
    lw, states = oracle.maximum(plan, k, c["nodes"], c["edges"])
    assert near(np.exp(lw[0]), exp["max_weight"], 1e-11)
    assert near(oracle.assignment_weight(plan, k, c["nodes"], c["edges"], states), exp["max_weight"], 1e-11)
    if case == "ties":
        assert np.all(states == 0)  # tests/inference_test.cc:210-217
    else:
        assert list(states) == exp["max_states"]


@pytest.mark.parametrize("seed", range(40))
def test_brute_force_random(seed):
    rng = np.random.default_rng(seed)
    k = int(rng.integers(1, 4))
    n = int(rng.integers(1, 8 if k <= 2 else 6))
    if seed % 4 == 3 and n >= 3:
        parents = forest([random_recursive(n // 2, rng), random_recursive(n - n // 2, rng)])
    else:
        parents = random_recursive(n, rng)
    plan = thmm.Plan(parents)
    nodes, edges = random_potentials(rng, 1, n, n - plan.roots.size, k, zero_fraction=0.15 if seed % 2 else 0.0)
    nodes, edges = nodes[0], edges
    z, bnm, bem, top, _ = oracle.brute_force(plan, k, nodes, edges)
    if z <= 0.0:
        with pytest.raises(oracle.OracleDomainError):
            oracle.marginals(plan, k, nodes, edges)
        return
    zr = oracle.partition(plan, k, nodes, edges)
    assert near(np.prod(zr), z, 1e-11)
    assert near(np.sum(oracle.log_partition(plan, k, nodes, edges)), np.log(z), 1e-11)
    log_z, nm, em = oracle.marginals(plan, k, nodes, edges)
    assert near(np.sum(log_z), np.log(z), 1e-11)
    assert near(nm, bnm, 1e-11)
    assert near(em, bem, 1e-11)
    lw, states = oracle.maximum(plan, k, nodes, edges)
    assert near(np.exp(np.sum(lw)), top, 1e-11)
    assert near(oracle.assignment_weight(plan, k, nodes, edges, states), top, 1e-11)


@pytest.mark.parametrize("seed", range(12))
def test_marginals_are_log_partition_gradients(seed):
    rng = np.random.default_rng(100 + seed)
    n, k = int(rng.integers(2, 30)), int(rng.integers(2, 5))
    plan = thmm.Plan(random_recursive(n, rng) if seed % 3 else path(n))
    nodes, edges = random_potentials(rng, 1, n, n - 1, k)
    nodes = nodes[0]
    _, nm, em = oracle.marginals(plan, k, nodes, edges)
    h = 1e-6
    for _ in range(6):
        i, s = int(rng.integers(n)), int(rng.integers(k))
        up, dn = nodes.copy(), nodes.copy()
        up[i, s] *= np.exp(h)
        dn[i, s] *= np.exp(-h)
        fd = (oracle.log_partition(plan, k, up, edges)[0] - oracle.log_partition(plan, k, dn, edges)[0]) / (2 * h)
        assert abs(fd - nm[i, s]) < 1e-7
        e, a, c = int(rng.integers(n - 1)), int(rng.integers(k)), int(rng.integers(k))
        up, dn = edges.copy(), edges.copy()
        up[e, a, c] *= np.exp(h)
        dn[e, a, c] *= np.exp(-h)
        fd = (oracle.log_partition(plan, k, nodes, up)[0] - oracle.log_partition(plan, k, nodes, dn)[0]) / (2 * h)
        assert abs(fd - em[e, a, c]) < 1e-7


def test_long_chain_closed_forms():
    # tests/inference_test.cc:286-308: K=1, psi = 0.5, edges 1
    n = 4096
    plan = thmm.Plan(path(n))
    assert near(oracle.log_partition(plan, 1, np.full(n, 0.5), np.ones(n - 1))[0], n * np.log(0.5), 1e-10)
    log_z, nm, em = oracle.marginals(plan, 1, np.full(n, 0.5), np.ones(n - 1))
    assert near(log_z[0], n * np.log(0.5), 1e-10)
    assert np.allclose(nm, 1.0, atol=1e-11) and np.allclose(em, 1.0, atol=1e-11)
    lw, _ = oracle.maximum(plan, 1, np.full(n, 0.5), np.ones(n - 1))
    assert near(lw[0], n * np.log(0.5), 1e-10)
    # :310-341: K=2 stochastic chain, uniform node factors 0.5
    edges = np.tile(np.array([[0.9, 0.1], [0.2, 0.8]]), (n - 1, 1, 1))
    log_z, nm, em = oracle.marginals(plan, 2, np.full((n, 2), 0.5), edges)
    assert near(log_z[0], (n - 1) * np.log(0.5), 1e-10)
    assert np.allclose(nm.sum(1), 1.0, atol=1e-10)
    assert np.allclose(em.sum((1, 2)), 1.0, atol=1e-10)
    parents, children = plan.edge_parents, plan.edge_children
    assert np.allclose(em.sum(2), nm[parents], atol=1e-10)
    assert np.allclose(em.sum(1), nm[children], atol=1e-10)


@pytest.mark.parametrize("seed", range(25))
def test_sample_matches_exhaustive_conditional_sampler(seed):
    rng = np.random.default_rng(200 + seed)
    k = int(rng.integers(2, 4))
    n = int(rng.integers(2, 8 if k == 2 else 6))
    parents = path(n) if seed % 5 == 0 else (star(n) if seed % 5 == 1 else random_recursive(n, rng))
    plan = thmm.Plan(parents)
    nodes, edges = random_potentials(rng, 1, n, n - 1, k, zero_fraction=0.2 if seed % 2 else 0.0)
    nodes = nodes[0]
    if oracle.brute_force(plan, k, nodes, edges)[0] <= 0:
        pytest.skip("zero partition")
    for _ in range(8):
        u = rng.random(n)
        states, dist = oracle.sample(plan, k, nodes, edges, u)
        expected = oracle.brute_conditional_sample(plan, k, nodes, edges, u)
        if np.min(dist) < 1e-12:
            continue  # a variate on a CDF boundary is reported, not compared
        assert np.array_equal(states, expected)
        assert oracle.assignment_weight(plan, k, nodes, edges, states) > 0.0


def test_sample_frequencies():
    # tests/inference_test.cc:249-274: LCG variates, 30,000 draws, 0.012.
    c = GOLDEN["literal"]
    plan = thmm.Plan(c["parents"])
    k = 2
    exp = c["expected"]
    state = 0xD1B54A32D192ED03
    counts_n = np.zeros((5, 2))
    counts_e = np.zeros((4, 2, 2))
    draws = 30000
    u = np.empty(5)
    for _ in range(draws):
        for i in range(5):
            state = (state * 6364136223846793005 + 1442695040888963407) & 0xFFFFFFFFFFFFFFFF
            u[i] = (state >> 11) * 2.0**-53
        s, _ = oracle.sample(plan, k, c["nodes"], c["edges"], u)
        counts_n[np.arange(5), s] += 1
        counts_e[np.arange(4), s[plan.edge_parents], s[plan.edge_children]] += 1
    assert near(counts_n.ravel() / draws, exp["node_marginals"], 0.012)
    assert near(counts_e.ravel() / draws, exp["edge_marginals"], 0.012)


def test_invalid_uniform_rejected():
    c = GOLDEN["literal"]
    plan = thmm.Plan(c["parents"])
    u = np.full(5, 0.5)
    u[0] = 1.0
    with pytest.raises(oracle.OracleError):
        oracle.sample(plan, 2, c["nodes"], c["edges"], u)


def test_planted_slips_are_caught():
    """Mutations of the oracle's inputs standing in for typical slips: a
    transposed edge matrix and a dropped node factor both move log Z away
    from enumeration, so the brute-force check above is sensitive to them."""
    c = GOLDEN["literal"]
    plan = thmm.Plan(c["parents"])
    edges = np.asarray(c["edges"]).reshape(4, 2, 2)
    nodes = np.asarray(c["nodes"]).reshape(5, 2)
    exact = c["expected"]["log_partition"]
    transposed = oracle.log_partition(plan, 2, nodes, edges.transpose(0, 2, 1))[0]
    dropped = nodes.copy()
    dropped[3] = 1.0
    assert abs(transposed - exact) > 1e-3
    assert abs(oracle.log_partition(plan, 2, dropped, edges)[0] - exact) > 1e-3


def test_batched_helper_matches_single(rng):
    plan = thmm.Plan(random_recursive(40, rng))
    nodes, edges = random_potentials(rng, 6, 40, 39, 3)
    log_z, nm, em = oracle.marginals_batched(plan, 3, nodes, edges, threads=3)
    for b in range(6):
        lz, n1, e1 = oracle.marginals(plan, 3, nodes[b], edges)
        assert np.array_equal(log_z[b], lz) and np.array_equal(nm[b], n1) and np.array_equal(em[b], e1)
