"""Shared fixtures: topology generators and the `gpu` marker."""

from __future__ import annotations

import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


# ---------------------------------------------------------------------------
# Topologies (parent arrays, -1 = root)
# ---------------------------------------------------------------------------


def random_recursive(n: int, rng: np.random.Generator) -> np.ndarray:
    """parent(i) ~ U{0..i-1}."""
    p = np.empty(n, np.int64)
    p[0] = -1
    if n > 1:
        p[1:] = (rng.random(n - 1) * np.arange(1, n)).astype(np.int64)
    return p


def path(n: int) -> np.ndarray:
    return np.arange(-1, n - 1, dtype=np.int64)


def star(n: int) -> np.ndarray:
    p = np.zeros(n, np.int64)
    p[0] = -1
    return p


def caterpillar(n: int) -> np.ndarray:
    """Spine with one leaf per spine node (node 2i spine, 2i+1 its leaf)."""
    i = np.arange(n, dtype=np.int64)
    p = np.where(i % 2 == 0, i - 2, i - 1)  # spine link / leaf link
    p[0] = -1
    return p


def balanced_binary(n: int) -> np.ndarray:
    p = (np.arange(n, dtype=np.int64) - 1) // 2
    p[0] = -1
    return p


def yule(leaves: int, rng: np.random.Generator) -> np.ndarray:
    """Random full binary tree: split a uniform random leaf until `leaves`."""
    parents = [-1]
    current = [0]
    while len(current) < leaves:
        i = int(rng.integers(len(current)))
        leaf = current[i]
        a, b = len(parents), len(parents) + 1
        parents += [leaf, leaf]
        current[i] = a
        current.append(b)
    return np.asarray(parents, np.int64)


def forest(trees: list[np.ndarray]) -> np.ndarray:
    """Concatenates parent arrays into one forest."""
    out, offset = [], 0
    for t in trees:
        shifted = np.where(t < 0, -1, t + offset)
        out.append(shifted)
        offset += t.size
    return np.concatenate(out)


def root_index_of_nodes(plan) -> np.ndarray:
    """For every node, the position of its tree's root in plan.roots (the
    order of per-root log partitions and gradient seeds)."""
    parents = np.full(plan.num_nodes, -1, np.int64)
    parents[plan.edge_children] = plan.edge_parents
    position = {int(r): i for i, r in enumerate(plan.roots)}
    out = np.empty(plan.num_nodes, np.int64)
    for v in range(plan.num_nodes):
        u = v
        while parents[u] >= 0:
            u = parents[u]
        out[v] = position[u]
    return out


def random_potentials(rng, batch, n_nodes, n_edges, k, zero_fraction=0.0, sigma=1.0):
    nodes = np.exp(rng.normal(0.0, sigma, size=(batch, n_nodes, k)))
    if zero_fraction:
        nodes[rng.random(nodes.shape) < zero_fraction] = 0.0
        # keep at least one supported state per node so that Z > 0
        dead = np.all(nodes == 0.0, axis=2)
        nodes[dead, 0] = 1.0
    rows = rng.exponential(1.0, size=(n_edges, k, k))
    edges = rows / rows.sum(axis=2, keepdims=True)
    return nodes, edges


@pytest.fixture
def rng():
    return np.random.default_rng(0x54484D4D)


def gpu_available() -> bool:
    try:
        import parallel_tree_hmm_amd as thmm

        return thmm.available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def require_gpu():
    """GPU tests fail loudly when the library or the device is missing."""
    import parallel_tree_hmm_amd as thmm

    assert thmm.available(), "no HIP device visible: the gpu tests need an MI355X"
    return thmm


TOPOLOGIES = {
    "recursive": lambda n, rng: random_recursive(n, rng),
    "path": lambda n, rng: path(n),
    "star": lambda n, rng: star(n),
    "caterpillar": lambda n, rng: caterpillar(n),
    "balanced": lambda n, rng: balanced_binary(n),
}

os.environ.setdefault("PYTHONHASHSEED", "0")


# ---------------------------------------------------------------------------
# Categorical phylogenetic-style model (config 3 shape)
# ---------------------------------------------------------------------------


def reversible_rate_matrix(rng, k: int) -> np.ndarray:
    """Q = S Pi with S symmetric Exp(1) exchangeabilities and pi ~ Dirichlet(1),
    rows summing to 0, scaled to unit mean rate (-sum_i pi_i Q_ii = 1)."""
    s = rng.exponential(1.0, size=(k, k))
    s = (s + s.T) / 2.0
    np.fill_diagonal(s, 0.0)
    pi = rng.dirichlet(np.ones(k))
    q = s * pi[None, :]
    np.fill_diagonal(q, -q.sum(axis=1))
    return q / -np.sum(pi * np.diag(q))


def stationary_of(q: np.ndarray) -> np.ndarray:
    """pi with pi Q = 0 (the left null vector of the rate matrix), normalised."""
    w, v = np.linalg.eig(q.T)
    pi = np.real(v[:, np.argmin(np.abs(w))])
    return pi / pi.sum()


def transition_matrices(q: np.ndarray, t: np.ndarray) -> np.ndarray:
    """P_e = expm(Q t_e) in FP64 ([edge, parent, child])."""
    from scipy.linalg import expm

    return np.stack([expm(q * te) for te in t])


def phylo_model(rng, depth: int, sites: int, k: int = 4, missing: float = 0.05):
    """Balanced binary tree with 2**depth leaves (parent(i) = (i-1)//2), the
    leaves observed per site with values in {0..k-1} and a fraction coded k
    (missing: an all-ones emission row); root ~ Dirichlet(1); emission
    [I_k; 1^T]; a reversible rate matrix; branch lengths t_e ~ Exp(mean 0.1),
    t_e the length of the edge into node e + 1."""
    n = 2 ** (depth + 1) - 1
    parents = balanced_binary(n)
    leaves = np.arange(n // 2, n, dtype=np.uint32)
    obs = rng.integers(0, k, size=(sites, leaves.size)).astype(np.uint8)
    obs[rng.random(obs.shape) < missing] = k
    root = rng.dirichlet(np.ones(k))
    emission = np.vstack([np.eye(k), np.ones((1, k))])
    q = reversible_rate_matrix(rng, k)
    t = rng.exponential(0.1, size=n - 1)
    return parents, leaves, obs, root, emission, q, t


def dense_categorical_nodes(n, leaves, obs, root, emission) -> np.ndarray:
    """The dense node factors of a categorical model: (root at node 0, else 1)
    times the emission row of the observed category at observed nodes."""
    sites, k = obs.shape[0], root.size
    nodes = np.ones((sites, n, k))
    nodes[:, 0, :] = root
    nodes[:, leaves, :] *= emission[obs]
    return nodes
