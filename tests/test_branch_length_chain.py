"""CPU pin of the branch-length chain rule used by the branch-length
gradient: with edge factors P_e = expm(Q t_e),
    d log Z / d t_e = sum_{p,c} M_e[p][c] (Q P_e)[p][c] / P_e[p][c],
M_e the oracle's edge marginals, checked against central finite differences
of the oracle's log Z in t (the FP64 oracle restates src/inference.cc)."""

from __future__ import annotations

import numpy as np

import oracle
import parallel_tree_hmm_amd as thmm
from conftest import dense_categorical_nodes, phylo_model, transition_matrices


def chain(q, p, edge_marginals):
    return np.einsum("epc,epc->e", edge_marginals, np.matmul(q, p) / p)


def test_chain_rule_matches_finite_differences():
    rng = np.random.default_rng(31)
    parents, leaves, obs, root, emission, q, t = phylo_model(rng, 3, 3)
    plan = thmm.Plan(parents)
    n = parents.size
    nodes = dense_categorical_nodes(n, leaves, obs, root, emission)
    p = transition_matrices(q, t)
    weights = np.array([0.5, -1.25, 2.0])
    grad = np.zeros(n - 1)
    for b in range(3):
        _, _, em = oracle.marginals(plan, 4, nodes[b], p)
        grad += weights[b] * chain(q, p, em)

    def objective(tt):
        pp = transition_matrices(q, tt)
        return sum(weights[b] * oracle.log_partition(plan, 4, nodes[b], pp)[0] for b in range(3))

    h = 1e-6
    for e in range(n - 1):
        up, dn = t.copy(), t.copy()
        up[e] += h
        dn[e] -= h
        fd = (objective(up) - objective(dn)) / (2 * h)
        assert abs(fd - grad[e]) <= 1e-6 * max(1.0, abs(fd)), (e, fd, grad[e])


def test_chain_rule_detects_a_transposed_rate_matrix():
    """A transposed Q (a plausible slip) changes the chain for a
    non-symmetric reversible Q, so the check above is sensitive to it."""
    rng = np.random.default_rng(32)
    parents, leaves, obs, root, emission, q, t = phylo_model(rng, 2, 1)
    plan = thmm.Plan(parents)
    nodes = dense_categorical_nodes(parents.size, leaves, obs, root, emission)
    p = transition_matrices(q, t)
    _, _, em = oracle.marginals(plan, 4, nodes[0], p)
    assert np.max(np.abs(chain(q, p, em) - chain(q.T, p, em))) > 1e-3
