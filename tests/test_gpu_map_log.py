"""The MAP kernels' log of FP32 potentials equals LogOfFloat evaluated on the
host (tests/test_map_log.py checks that function against the C library log
and logl), bit for bit:

* single-node trees: the MAP weight of a tree is the largest log of its
  node's K potentials (the node log transform, then the finish);
* two-node trees: max over p of log x0[p] + max over c of (log e[p][c] +
  log x1[c]) (the rake kernel's logs of edge factors taken on load, after
  its linear-domain screen, and the node logs), at K = 8, 16 and 32.
Compressions log their edge operands through the same function; their MAP
results are compared with the oracle in test_gpu_parity.py and
test_gpu_map_pieces.py.

Potentials span the float range (exp of N(0, 20), values next to 1 and to
powers of two, subnormals) so that every table interval and both reduction
branches are met.
"""

from __future__ import annotations

import numpy as np
import pytest

import parallel_tree_hmm_amd as thmm
from parallel_tree_hmm_amd import _abi
from test_map_log import built, host_log  # noqa: F401  (fixture)

pytestmark = pytest.mark.gpu


def wide_floats(rng, shape):
    x = np.exp(rng.normal(0.0, 20.0, size=shape)).astype(np.float32)
    flat = x.reshape(-1)
    n = flat.size
    pick = rng.random(n)
    near_one = np.float32(1.0) + rng.integers(-300, 300, size=n).astype(np.float32) * np.float32(2.0 ** -23)
    flat[pick < 0.15] = near_one[pick < 0.15]
    pow2 = np.ldexp(np.float32(1.0), rng.integers(-120, 120, size=n)).astype(np.float32)
    below = np.nextafter(pow2, np.float32(0))
    flat[(pick >= 0.15) & (pick < 0.25)] = below[(pick >= 0.15) & (pick < 0.25)]
    flat[(pick >= 0.25) & (pick < 0.3)] = np.float32(3e-42)  # subnormal
    flat[~np.isfinite(flat)] = np.float32(1.0)
    flat[flat == 0] = np.float32(1e-40)
    return x


def run_map(torch, plan, k, nodes, edges):
    ws = thmm.Workspace()
    ws.reserve(plan, k, 1, precision="f32", flags=_abi.THMM_RESERVE_MAXIMUM | _abi.THMM_RESERVE_DEVICE_RESIDENT)
    d_nodes = torch.tensor(nodes, device="cuda")
    d_edges = torch.tensor(edges, device="cuda") if edges.size else torch.zeros(1, device="cuda")
    weights = torch.empty(plan.roots.size, dtype=torch.float64, device="cuda")
    states = torch.empty(plan.num_nodes, dtype=torch.int32, device="cuda")
    ws.maximum_device(d_nodes.data_ptr(), d_edges.data_ptr(), 1, weights.data_ptr(), states.data_ptr(), 0)
    torch.cuda.synchronize()
    return weights.cpu().numpy()


@pytest.mark.parametrize("k", [8, 16, 32])
def test_single_node_weights_are_host_logs(require_gpu, built, k):  # noqa: F811
    import torch

    _, lib = built
    rng = np.random.default_rng(k)
    n = 20_000
    plan = thmm.Plan(np.full(n, -1, np.int64))
    nodes = wide_floats(rng, (1, n, k))
    weights = run_map(torch, plan, k, nodes, np.zeros((0, k, k), np.float32))
    expected = host_log(lib, nodes[0]).max(axis=1)
    assert np.array_equal(weights, expected)


@pytest.mark.parametrize("k", [8, 16, 32])
def test_edge_logs_are_host_logs(require_gpu, built, k):  # noqa: F811
    import torch

    _, lib = built
    rng = np.random.default_rng(100 + k)
    pairs = 4000
    parents = np.empty(2 * pairs, np.int64)
    parents[0::2] = -1
    parents[1::2] = np.arange(0, 2 * pairs, 2)
    plan = thmm.Plan(parents)
    nodes = wide_floats(rng, (1, 2 * pairs, k))
    edges = wide_floats(rng, (plan.num_edges, k, k))
    weights = run_map(torch, plan, k, nodes, edges)
    ln = host_log(lib, nodes[0])
    le = host_log(lib, edges)
    edge_of_child = {int(c): e for e, c in enumerate(plan.edge_children)}
    expected = []
    for r in plan.roots:
        r = int(r)
        # Rake: branch[p] = max_c (log e[p][c] + log x1[c]); absorb, finish.
        branch = (le[edge_of_child[r + 1]] + ln[r + 1][None, :]).max(axis=1)
        expected.append((ln[r] + branch).max())
    assert np.array_equal(weights, np.array(expected))
