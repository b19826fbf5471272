"""IEEE binary16 edge outputs (THMM_RESERVE_HALF_EDGE_MARGINALS) of the
device-resident entry points. The kernels compute exactly what they compute
for wide outputs and round each edge value once at the store, so checked here:

* log Z and node marginals are bitwise those of the same reservation
  without the flag;
* every edge value is the wide value rounded to binary16:
  |half - wide| <= 2^-11 |wide| + 2^-25 (half an ulp, normal or subnormal),
  and within that plus the FP32 tolerance of the FP64 oracle;
* on every reverse kernel that stores edge adjoints: the K = 16 MFMA
  compressions (compress_depth 1) and grouped chains (depths 2 and 3), the
  K = 32 / 64 wide MFMA compressions, the team rake kernels for FP32 and
  FP64, the plan-level kernels (K = 5, and outputs off 16-byte alignment,
  which also take the chain kernels' per-value stores);
* the log-potential gradient pair, graph replays, and the refusal of
  outputs that are not 2-byte aligned.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
import parallel_tree_hmm_amd as thmm
from conftest import caterpillar, forest, path, random_potentials, random_recursive, star

pytestmark = pytest.mark.gpu

HALF_REL = 2.0**-11
HALF_ABS = 2.0**-25


def half_close(half, wide):
    """Max of |half - wide| / (2^-11 |wide| + 2^-25): at most 1 when each
    value is the wide value rounded to nearest binary16."""
    h = half.astype(np.float64)
    w = wide.astype(np.float64)
    return float(np.max(np.abs(h - w) / (HALF_REL * np.abs(w) + HALF_ABS)))


def make_plan(rng):
    return thmm.Plan(forest([caterpillar(601), path(400), random_recursive(500, rng), star(100)]))


def run_marginals(torch, ws, plan, k, batch, d_nodes, d_edges, edge_dtype, edge_offset=0):
    lz = torch.empty(batch * plan.roots.size, dtype=torch.float64, device="cuda")
    nm = torch.empty((batch, plan.num_nodes, k), dtype=d_nodes.dtype, device="cuda")
    count = batch * plan.num_edges * k * k
    store = torch.empty(count + edge_offset, dtype=edge_dtype, device="cuda")
    ws.marginals_device(d_nodes.data_ptr(), d_edges.data_ptr(), batch, lz.data_ptr(), nm.data_ptr(),
                        store[edge_offset:].data_ptr(), 0)
    torch.cuda.synchronize()
    em = store[edge_offset:].reshape(batch, plan.num_edges, k, k)
    return lz.cpu().numpy(), nm.cpu().numpy(), em.cpu().numpy()


CASES = [
    # (K, precision, compress_depth, edge offset in values)
    (16, "f32", 1, 0),
    (16, "f32", 2, 0),
    (16, "f32", 3, 0),
    (16, "f32", 2, 2),
    (8, "f32", 2, 0),
    (32, "f32", 2, 0),
    (64, "f32", 2, 0),
    (8, "f64", 2, 0),
    (16, "f64", 2, 0),
    (5, "f32", 2, 0),
]


@pytest.mark.parametrize("k,precision,depth,offset", CASES)
def test_half_edge_marginals_round_the_wide_values(require_gpu, k, precision, depth, offset):
    import torch

    rng = np.random.default_rng(100 * k + depth + offset)
    plan = make_plan(rng)
    batch = 2
    nodes, edges = random_potentials(rng, batch, plan.num_nodes, plan.num_edges, k, zero_fraction=0.1, sigma=1.0)
    dtype = torch.float32 if precision == "f32" else torch.float64
    d_nodes = torch.tensor(nodes, dtype=dtype, device="cuda")
    d_edges = torch.tensor(edges, dtype=dtype, device="cuda")
    wide_ws, half_ws = thmm.Workspace(), thmm.Workspace()
    wide_ws.reserve_marginals(plan, k, batch, precision=precision, device_resident=True)
    half_ws.reserve_marginals(plan, k, batch, precision=precision, device_resident=True, half_edge_marginals=True)
    wide_ws.set_tuning(compress_depth=depth)
    half_ws.set_tuning(compress_depth=depth)
    # Misaligned outputs take other kernels: the wide run gets the same offset.
    wlz, wnm, wem = run_marginals(torch, wide_ws, plan, k, batch, d_nodes, d_edges, dtype, offset)
    hlz, hnm, hem = run_marginals(torch, half_ws, plan, k, batch, d_nodes, d_edges, torch.float16, offset)
    ratio = half_close(hem, wem)
    print(f"K {k} {precision} depth {depth} offset {offset}: half/wide {ratio:.3f} of half an ulp")
    assert np.array_equal(hlz, wlz)
    assert np.array_equal(hnm, wnm)
    assert ratio <= 1.0
    if precision == "f32":
        _, _, oem = oracle.marginals_batched(plan, k, nodes.astype(np.float32).astype(np.float64),
                                             edges.astype(np.float32).astype(np.float64), threads=4)
        err = np.abs(hem.astype(np.float64) - oem)
        assert np.all(err <= HALF_REL * np.abs(oem) + HALF_ABS + 1e-5)
    wide_ws.close()
    half_ws.close()


def test_half_edge_gradients_of_the_log_potential_pair(require_gpu):
    import torch

    rng = np.random.default_rng(7)
    plan = make_plan(rng)
    batch, k = 2, 16
    log_nodes = torch.tensor(rng.normal(size=(batch, plan.num_nodes, k)), dtype=torch.float32, device="cuda")
    log_edges = torch.tensor(rng.normal(size=(plan.num_edges, k, k)), dtype=torch.float32, device="cuda")
    grad = torch.tensor(rng.normal(size=batch * plan.roots.size), dtype=torch.float64, device="cuda")
    outputs = {}
    for half in (False, True):
        ws = thmm.Workspace()
        ws.reserve_marginals(plan, k, batch, precision="f32", device_resident=True, half_edge_marginals=half)
        lz = torch.empty(batch * plan.roots.size, dtype=torch.float64, device="cuda")
        gn = torch.empty((batch, plan.num_nodes, k), device="cuda")
        ge = torch.empty((batch, plan.num_edges, k, k), dtype=torch.float16 if half else torch.float32, device="cuda")
        ws.logz_forward_device(log_nodes.data_ptr(), log_edges.data_ptr(), batch, lz.data_ptr(), 0)
        ws.logz_backward_device(grad.data_ptr(), gn.data_ptr(), ge.data_ptr(), False, 0)
        torch.cuda.synchronize()
        outputs[half] = (lz.cpu().numpy(), gn.cpu().numpy(), ge.cpu().numpy())
        if half:
            with pytest.raises(thmm.LogicError):  # batch-reduced edge gradients need workspace storage
                ws.logz_backward_device(grad.data_ptr(), gn.data_ptr(), ge.data_ptr(), True, 0)
        ws.close()
    assert np.array_equal(outputs[True][0], outputs[False][0])
    assert np.array_equal(outputs[True][1], outputs[False][1])
    assert half_close(outputs[True][2], outputs[False][2]) <= 1.0


def test_half_edge_marginals_replay_in_graphs(require_gpu):
    import torch

    rng = np.random.default_rng(11)
    plan = make_plan(rng)
    batch, k = 3, 16
    nodes, edges = random_potentials(rng, batch, plan.num_nodes, plan.num_edges, k)
    d_nodes = torch.tensor(nodes, dtype=torch.float32, device="cuda")
    d_edges = torch.tensor(edges, dtype=torch.float32, device="cuda")
    eager, graphs = thmm.Workspace(), thmm.Workspace()
    eager.reserve_marginals(plan, k, batch, precision="f32", device_resident=True, half_edge_marginals=True)
    graphs.reserve_marginals(plan, k, batch, precision="f32", device_resident=True, graphs=True,
                             half_edge_marginals=True)
    want = run_marginals(torch, eager, plan, k, batch, d_nodes, d_edges, torch.float16)
    lz = torch.empty(batch * plan.roots.size, dtype=torch.float64, device="cuda")
    nm = torch.empty((batch, plan.num_nodes, k), device="cuda")
    em = torch.empty((batch, plan.num_edges, k, k), dtype=torch.float16, device="cuda")
    for replay in range(3):
        em.fill_(float("nan"))
        graphs.marginals_device(d_nodes.data_ptr(), d_edges.data_ptr(), batch, lz.data_ptr(), nm.data_ptr(),
                                em.data_ptr(), 0)
        torch.cuda.synchronize()
        assert np.array_equal(lz.cpu().numpy(), want[0]), replay
        assert np.array_equal(nm.cpu().numpy(), want[1]), replay
        assert np.array_equal(em.cpu().numpy(), want[2]), replay
    eager.close()
    graphs.close()


def test_half_edge_output_alignment_is_checked(require_gpu):
    import torch

    plan = thmm.Plan(path(50))
    k = 4
    ws = thmm.Workspace()
    ws.reserve_marginals(plan, k, 1, precision="f32", device_resident=True, half_edge_marginals=True)
    nodes = torch.ones((1, plan.num_nodes, k), device="cuda")
    edges = torch.ones((plan.num_edges, k, k), device="cuda")
    lz = torch.empty(1, dtype=torch.float64, device="cuda")
    nm = torch.empty((1, plan.num_nodes, k), device="cuda")
    store = torch.empty(plan.num_edges * k * k + 2, dtype=torch.float16, device="cuda")
    with pytest.raises(thmm.InvalidArgument):
        ws.marginals_device(nodes.data_ptr(), edges.data_ptr(), 1, lz.data_ptr(), nm.data_ptr(),
                            store.data_ptr() + 1, 0)
    torch.cuda.synchronize()
    ws.close()
