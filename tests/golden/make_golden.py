"""Writes tests/golden/inference_fixtures.json.

Inputs are the literal arrays of joaospinto/parallel_tree_hmm
tests/inference_test.cc:159-168 (tree {-1,0,0,1,3}, K=2) and its zero-factor
variant (:187-193), plus the all-ones tie case (:210-217). The reference
stores no expected values (it enumerates at test time), so the expected
outputs here come from an exhaustive enumeration written independently in
numpy below, not from the oracle.

Edge order: the reference's edges follow btrc's plan order, which is not
available; this fixture uses edge e = the edge into node e + 1 (ascending
non-root nodes), the canonical order of this project's planner.
"""

from __future__ import annotations

import itertools
import json
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent

PARENTS = [-1, 0, 0, 1, 3]
NODES = [0.55, 0.45, 0.9, 0.2, 0.4, 0.8, 0.7, 0.3, 0.1, 0.95]
EDGES = [0.8, 0.2, 0.3, 0.7, 0.9, 0.1, 0.25, 0.75, 0.6, 0.4, 0.15, 0.85, 0.7, 0.3, 0.2, 0.8]
ZERO_NODES = [0.55, 0.45, 1.0, 0.0, 0.0, 1.0, 0.7, 0.3, 0.0, 1.0]
ZERO_EDGES = [0.8, 0.2, 0.3, 0.7, 1.0, 0.0, 0.25, 0.75, 0.6, 0.4, 0.0, 1.0, 0.7, 0.3, 0.2, 0.8]


def enumerate_model(parents, k, nodes, edges):
    n = len(parents)
    nodes = np.asarray(nodes, float).reshape(n, k)
    children = [c for c in range(n) if parents[c] >= 0]
    edges = np.asarray(edges, float).reshape(len(children), k, k)
    z = 0.0
    node_mass = np.zeros((n, k))
    edge_mass = np.zeros_like(edges)
    best_w, best = -1.0, None
    # node 0 varies fastest, as tests/inference_test.cc:67-73 decodes codes
    for code in range(k**n):
        x = [(code // k**i) % k for i in range(n)]
        w = 1.0
        for i in range(n):
            w *= nodes[i, x[i]]
        for e, c in enumerate(children):
            w *= edges[e, x[parents[c]], x[c]]
        z += w
        for i in range(n):
            node_mass[i, x[i]] += w
        for e, c in enumerate(children):
            edge_mass[e, x[parents[c]], x[c]] += w
        if w > best_w:
            best_w, best = w, x
    return {
        "partition": z,
        "log_partition": float(np.log(z)),
        "node_marginals": (node_mass / z).ravel().tolist(),
        "edge_marginals": (edge_mass / z).ravel().tolist(),
        "max_weight": best_w,
        "max_states": best,
    }


def main():
    cases = {}
    for name, nodes, edges in (
        ("literal", NODES, EDGES),
        ("zero", ZERO_NODES, ZERO_EDGES),
        ("ties", [1.0] * 10, [1.0] * 16),
    ):
        cases[name] = {
            "parents": PARENTS,
            "states": 2,
            "nodes": nodes,
            "edges": edges,
            "expected": enumerate_model(PARENTS, 2, nodes, edges),
        }
    (HERE / "inference_fixtures.json").write_text(json.dumps(cases, indent=1) + "\n")


if __name__ == "__main__":
    main()
