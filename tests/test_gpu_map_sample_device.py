"""Device-resident MAP and posterior sampling (thmm_maximum_device,
thmm_sample_device): the kernels of the host-staged calls on caller device
buffers. Checked here, for FP32 and FP64 workspaces:

* states are those of the host-staged call bit for bit, and those of the
  FP64 oracle (MAP exactly; samples away from CDF boundaries, as in the host
  tests); MAP log weights and log Z are the host call's values (the host
  call rounds them once to the workspace precision);
* a tree whose potentials leave no positive weight gives -inf and, for
  sampling, the failure flag, where the host call raises; the other trees'
  results are unaffected;
* variates outside [0, 1), which the host call rejects, still give states
  in range;
* at K = 8 and 16, the edge logs (MAP) and FP64 widening (sampling) taken
  on load equal, bit for bit, the whole-array passes that edges off
  16-byte alignment take, also where the MAP rake kernel's linear-domain
  screen meets ties, zeros and extreme magnitudes;
* a caller's graph capture of either call replays to the eager results
  (sampling clears its failure flag with a kernel, not a memset node).
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
import parallel_tree_hmm_amd as thmm
from parallel_tree_hmm_amd import _abi
from conftest import caterpillar, forest, path, random_potentials, random_recursive, star

pytestmark = pytest.mark.gpu


def make_plan(rng):
    return thmm.Plan(forest([random_recursive(300, rng), caterpillar(201), path(150), star(80)]))


def device_inputs(torch, nodes, edges, precision):
    dtype = torch.float32 if precision == "f32" else torch.float64
    return torch.tensor(nodes, dtype=dtype, device="cuda"), torch.tensor(edges, dtype=dtype, device="cuda")


def as_precision(x, precision):
    return np.asarray(x, np.float32) if precision == "f32" else np.asarray(x, np.float64)


def reserve_device(plan, k, batch, precision, flag):
    ws = thmm.Workspace()
    ws.reserve(plan, k, batch, precision=precision, flags=flag | _abi.THMM_RESERVE_DEVICE_RESIDENT)
    return ws


@pytest.mark.parametrize("precision", ["f32", "f64"])
@pytest.mark.parametrize("k", [4, 16])
def test_maximum_device_matches_host_and_oracle(require_gpu, precision, k):
    import torch

    rng = np.random.default_rng(5 * k + (precision == "f64"))
    plan = make_plan(rng)
    batch = 3
    nodes, edges = random_potentials(rng, batch, plan.num_nodes, plan.num_edges, k, zero_fraction=0.1, sigma=2.0)
    nodes, edges = as_precision(nodes, precision), as_precision(edges, precision)
    host_ws = thmm.Workspace()
    host_ws.reserve_maximum(plan, k, batch, precision=precision)
    host = thmm.maximum_a_posteriori_prepared(thmm.BatchedModelView(plan, k, batch, nodes, edges), host_ws)
    ws = reserve_device(plan, k, batch, precision, _abi.THMM_RESERVE_MAXIMUM)
    d_nodes, d_edges = device_inputs(torch, nodes, edges, precision)
    weights = torch.empty(batch * plan.roots.size, dtype=torch.float64, device="cuda")
    states = torch.empty((batch, plan.num_nodes), dtype=torch.int32, device="cuda")
    ws.maximum_device(d_nodes.data_ptr(), d_edges.data_ptr(), batch, weights.data_ptr(), states.data_ptr(), 0)
    torch.cuda.synchronize()
    got_states = states.cpu().numpy().view(np.uint32)
    got_weights = weights.cpu().numpy()
    assert np.array_equal(got_states.reshape(-1), np.asarray(host.states).reshape(-1))
    assert np.array_equal(got_weights.astype(nodes.dtype), np.asarray(host.log_weights).reshape(-1))
    for b in range(batch):
        lw, expected = oracle.maximum(plan, k, nodes[b].astype(np.float64), edges.astype(np.float64))
        assert np.array_equal(got_states[b], expected)
        assert np.all(np.abs(got_weights.reshape(batch, -1)[b] - lw) <= 1e-12 * np.maximum(1, np.abs(lw)))
    host_ws.close()
    ws.close()


@pytest.mark.parametrize("precision", ["f32", "f64"])
@pytest.mark.parametrize("k", [4, 16])
def test_sample_device_matches_host_and_oracle(require_gpu, precision, k):
    import torch

    rng = np.random.default_rng(9 * k + (precision == "f64"))
    plan = make_plan(rng)
    batch = 3
    nodes, edges = random_potentials(rng, batch, plan.num_nodes, plan.num_edges, k, zero_fraction=0.1, sigma=2.0)
    nodes, edges = as_precision(nodes, precision), as_precision(edges, precision)
    u = as_precision(rng.random((batch, plan.num_nodes)), precision)
    host_ws = thmm.Workspace()
    host_ws.reserve_sampling(plan, k, batch, precision=precision)
    host = thmm.posterior_sample_prepared(thmm.BatchedModelView(plan, k, batch, nodes, edges), u, host_ws)
    host_states = np.asarray(host.states).reshape(batch, -1).copy()
    host_log_z = host_ws.log_partitions_f64()[: batch * plan.roots.size].copy()
    ws = reserve_device(plan, k, batch, precision, _abi.THMM_RESERVE_SAMPLING)
    d_nodes, d_edges = device_inputs(torch, nodes, edges, precision)
    d_u = torch.tensor(u, device="cuda")
    log_z = torch.empty(batch * plan.roots.size, dtype=torch.float64, device="cuda")
    states = torch.empty((batch, plan.num_nodes), dtype=torch.int32, device="cuda")
    failure = torch.full((1,), 7, dtype=torch.int32, device="cuda")
    ws.sample_device(d_nodes.data_ptr(), d_edges.data_ptr(), d_u.data_ptr(), batch, log_z.data_ptr(),
                     states.data_ptr(), failure.data_ptr(), 0)
    torch.cuda.synchronize()
    got = states.cpu().numpy().view(np.uint32)
    assert int(failure.item()) == 0
    assert np.array_equal(got, host_states)
    assert np.array_equal(log_z.cpu().numpy(), host_log_z)
    for b in range(batch):
        expected, dist = oracle.sample(plan, k, nodes[b].astype(np.float64), edges.astype(np.float64),
                                       u[b].astype(np.float64))
        if np.min(dist) >= 1e-12:
            assert np.array_equal(got[b], expected)
    host_ws.close()
    ws.close()


def test_device_calls_report_unsupported_trees_in_their_outputs(require_gpu):
    import torch

    rng = np.random.default_rng(2)
    plan = make_plan(rng)
    k, batch = 4, 2
    nodes, edges = random_potentials(rng, batch, plan.num_nodes, plan.num_edges, k)
    nodes, edges = nodes.astype(np.float32), edges.astype(np.float32)
    # Item 1: the first tree's root has no supported state.
    root = int(plan.roots[0])
    nodes[1, root, :] = 0.0
    d_nodes, d_edges = device_inputs(torch, nodes, edges, "f32")
    roots = plan.roots.size

    map_ws = reserve_device(plan, k, batch, "f32", _abi.THMM_RESERVE_MAXIMUM)
    weights = torch.empty(batch * roots, dtype=torch.float64, device="cuda")
    states = torch.empty((batch, plan.num_nodes), dtype=torch.int32, device="cuda")
    map_ws.maximum_device(d_nodes.data_ptr(), d_edges.data_ptr(), batch, weights.data_ptr(), states.data_ptr(), 0)
    torch.cuda.synchronize()
    w = weights.cpu().numpy().reshape(batch, roots)
    assert np.isneginf(w[1, 0]) and np.all(np.isfinite(w[0])) and np.all(np.isfinite(w[1, 1:]))
    assert np.all(states.cpu().numpy().view(np.uint32) < k)

    sample_ws = reserve_device(plan, k, batch, "f32", _abi.THMM_RESERVE_SAMPLING)
    # Variates outside [0, 1) are taken as given: states stay in range.
    u = rng.random((batch, plan.num_nodes)).astype(np.float32)
    u[0, ::7] = 1.5
    u[0, 3::7] = -0.5
    d_u = torch.tensor(u, device="cuda")
    log_z = torch.empty(batch * roots, dtype=torch.float64, device="cuda")
    failure = torch.zeros(1, dtype=torch.int32, device="cuda")
    sample_ws.sample_device(d_nodes.data_ptr(), d_edges.data_ptr(), d_u.data_ptr(), batch, log_z.data_ptr(),
                            states.data_ptr(), failure.data_ptr(), 0)
    torch.cuda.synchronize()
    lz = log_z.cpu().numpy().reshape(batch, roots)
    assert np.isneginf(lz[1, 0]) and np.all(np.isfinite(lz[0]))
    assert int(failure.item()) == 1
    assert np.all(states.cpu().numpy().view(np.uint32) < k)

    # The host-staged calls raise for the same inputs.
    host_ws = thmm.Workspace()
    host_ws.reserve_sampling(plan, k, batch, precision="f32")
    with pytest.raises(thmm.DomainError):
        thmm.posterior_sample_prepared(thmm.BatchedModelView(plan, k, batch, nodes, edges),
                                       np.clip(u, 0.0, 0.999), host_ws)
    for ws in (map_ws, sample_ws, host_ws):
        ws.close()


def test_device_calls_need_their_reservations(require_gpu):
    import torch

    plan = thmm.Plan(path(20))
    ws = thmm.Workspace()
    ws.reserve_marginals(plan, 4, 1, precision="f32", device_resident=True)
    x = torch.ones((1, plan.num_nodes, 4), device="cuda")
    e = torch.ones((plan.num_edges, 4, 4), device="cuda")
    out = torch.empty(plan.num_nodes, dtype=torch.int32, device="cuda")
    lz = torch.empty(1, dtype=torch.float64, device="cuda")
    with pytest.raises(thmm.InvalidArgument):
        ws.maximum_device(x.data_ptr(), e.data_ptr(), 1, lz.data_ptr(), out.data_ptr(), 0)
    with pytest.raises(thmm.InvalidArgument):
        ws.sample_device(x.data_ptr(), e.data_ptr(), x.data_ptr(), 1, lz.data_ptr(), out.data_ptr(), out.data_ptr(), 0)
    torch.cuda.synchronize()
    ws.close()


@pytest.mark.parametrize("k", [8, 16])
def test_on_load_edge_transforms_equal_the_whole_array_passes(require_gpu, k):
    """At K = 8 and 16 the MAP level kernels log FP32 edge values as they
    load them, and the sampling kernels widen them to FP64; edges off 16-byte
    alignment take the log / widening pass into FP64 copies instead. Both
    give the same bits."""
    import torch

    rng = np.random.default_rng(40 + k)
    plan = make_plan(rng)
    batch = 2
    nodes, edges = random_potentials(rng, batch, plan.num_nodes, plan.num_edges, k, zero_fraction=0.1, sigma=2.0)
    d_nodes = torch.tensor(nodes, dtype=torch.float32, device="cuda")
    store = torch.empty(edges.size + 1, dtype=torch.float32, device="cuda")
    store[1:] = torch.tensor(edges.reshape(-1), dtype=torch.float32, device="cuda")
    aligned = torch.tensor(edges, dtype=torch.float32, device="cuda")
    u = torch.tensor(rng.random((batch, plan.num_nodes)), dtype=torch.float32, device="cuda")
    results = {}
    for name, ptr in (("on load", aligned.data_ptr()), ("pass", store[1:].data_ptr())):
        out = []
        ws = reserve_device(plan, k, batch, "f32", _abi.THMM_RESERVE_MAXIMUM)
        weights = torch.empty(batch * plan.roots.size, dtype=torch.float64, device="cuda")
        states = torch.empty((batch, plan.num_nodes), dtype=torch.int32, device="cuda")
        ws.maximum_device(d_nodes.data_ptr(), ptr, batch, weights.data_ptr(), states.data_ptr(), 0)
        torch.cuda.synchronize()
        out += [weights.cpu().numpy(), states.cpu().numpy()]
        ws.close()
        ws = reserve_device(plan, k, batch, "f32", _abi.THMM_RESERVE_SAMPLING)
        failure = torch.zeros(1, dtype=torch.int32, device="cuda")
        ws.sample_device(d_nodes.data_ptr(), ptr, u.data_ptr(), batch, weights.data_ptr(), states.data_ptr(),
                         failure.data_ptr(), 0)
        torch.cuda.synchronize()
        out += [weights.cpu().numpy(), states.cpu().numpy(), int(failure.item())]
        ws.close()
        results[name] = out
    for a, b in zip(results["on load"], results["pass"]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("k,precision", [(8, "f32"), (16, "f32"), (16, "f64"), (32, "f32"), (32, "f64")])
def test_screened_rake_rows_match_the_logged_rows(require_gpu, k, precision):
    """The MAP rake kernel screens raw edge rows in the linear domain and
    logs only the columns that can win. On inputs built to stress the screen
    its results equal, bit for bit, those of the unscreened kernel (edges off
    16-byte alignment are logged whole first), and its states the oracle's.
    The inputs hold exact ties (duplicated states and columns), zeros,
    subnormal and near-maximal FP32 factors, leaf vectors spanning thousands
    of nats (a star's centre raked after absorbing extreme leaves), and, for
    FP64, factors above 2^128 (rows logged whole)."""
    import torch

    rng = np.random.default_rng(60 + k + (precision == "f64"))
    plan = thmm.Plan(forest([random_recursive(300, rng), caterpillar(201), star(120), path(40)]))
    batch = 2
    nodes, edges = random_potentials(rng, batch, plan.num_nodes, plan.num_edges, k, zero_fraction=0.1, sigma=2.0)
    n_tie = rng.random(plan.num_nodes) < 0.3
    nodes[:, n_tie, 1] = nodes[:, n_tie, 0]
    wide = rng.random(plan.num_nodes) < 0.2
    nodes[:, wide, :] *= np.exp(rng.choice([-80.0, 80.0], size=(batch, int(wide.sum()), k)))
    e_tie = rng.random(plan.num_edges) < 0.3
    edges[e_tie, :, 1] = edges[e_tie, :, 0]
    picks = rng.random(edges.shape)
    edges[picks < 0.02] = 1e-40 if precision == "f32" else 1e-300
    edges[(picks >= 0.02) & (picks < 0.04)] = 3e38 if precision == "f32" else 1e300
    if precision == "f64":
        edges[(picks >= 0.04) & (picks < 0.05)] = 1e45
    nodes, edges = as_precision(nodes, precision), as_precision(edges, precision)
    assert np.all(np.isfinite(nodes)) and np.all(np.isfinite(edges))
    dtype = torch.float32 if precision == "f32" else torch.float64
    d_nodes = torch.tensor(nodes, dtype=dtype, device="cuda")
    aligned = torch.tensor(edges, dtype=dtype, device="cuda")
    store = torch.empty(edges.size + 1, dtype=dtype, device="cuda")
    store[1:] = aligned.reshape(-1)
    results = {}
    for name, ptr in (("screened", aligned.data_ptr()), ("logged", store[1:].data_ptr())):
        ws = reserve_device(plan, k, batch, precision, _abi.THMM_RESERVE_MAXIMUM)
        weights = torch.empty(batch * plan.roots.size, dtype=torch.float64, device="cuda")
        states = torch.empty((batch, plan.num_nodes), dtype=torch.int32, device="cuda")
        ws.maximum_device(d_nodes.data_ptr(), ptr, batch, weights.data_ptr(), states.data_ptr(), 0)
        torch.cuda.synchronize()
        results[name] = (weights.cpu().numpy(), states.cpu().numpy().view(np.uint32))
        ws.close()
    assert np.array_equal(results["screened"][0], results["logged"][0])
    assert np.array_equal(results["screened"][1], results["logged"][1])
    got_weights = results["screened"][0].reshape(batch, -1)
    for b in range(batch):
        lw, expected = oracle.maximum(plan, k, nodes[b].astype(np.float64), edges.astype(np.float64))
        assert np.array_equal(results["screened"][1][b], expected)
        finite = np.isfinite(lw)
        assert np.array_equal(np.isfinite(got_weights[b]), finite)
        assert np.all(np.abs(got_weights[b][finite] - lw[finite]) <= 1e-12 * np.maximum(1, np.abs(lw[finite])))


@pytest.mark.parametrize("kind", ["map", "sample"])
def test_device_calls_replay_in_a_caller_graph(require_gpu, kind):
    """A caller may capture thmm_maximum_device or thmm_sample_device on its
    stream into a graph of its own: both enqueue kernels only (sampling
    clears its failure flag with a kernel, not a memset node). Every replay,
    after the outputs are overwritten eagerly, gives the eager call's
    results (and a cleared flag)."""
    import torch

    rng = np.random.default_rng(71)
    plan = make_plan(rng)
    k, batch = 16, 2
    nodes, edges = random_potentials(rng, batch, plan.num_nodes, plan.num_edges, k, zero_fraction=0.1, sigma=2.0)
    d_nodes, d_edges = device_inputs(torch, nodes, edges, "f32")
    u = torch.tensor(rng.random((batch, plan.num_nodes)), dtype=torch.float32, device="cuda")
    flag = _abi.THMM_RESERVE_MAXIMUM if kind == "map" else _abi.THMM_RESERVE_SAMPLING
    ws = reserve_device(plan, k, batch, "f32", flag)
    log_z = torch.empty(batch * plan.roots.size, dtype=torch.float64, device="cuda")
    states = torch.empty((batch, plan.num_nodes), dtype=torch.int32, device="cuda")
    failure = torch.ones(1, dtype=torch.int32, device="cuda") if kind == "sample" else torch.zeros(1, dtype=torch.int32,
                                                                                                   device="cuda")
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())

    def call():
        if kind == "map":
            ws.maximum_device(d_nodes.data_ptr(), d_edges.data_ptr(), batch, log_z.data_ptr(), states.data_ptr(),
                              stream.cuda_stream)
        else:
            ws.sample_device(d_nodes.data_ptr(), d_edges.data_ptr(), u.data_ptr(), batch, log_z.data_ptr(),
                             states.data_ptr(), failure.data_ptr(), stream.cuda_stream)

    with torch.cuda.stream(stream):
        call()
    torch.cuda.synchronize()
    want = (log_z.cpu().numpy().copy(), states.cpu().numpy().copy())
    assert int(failure.item()) == 0
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=stream):
        call()
    torch.cuda.synchronize()
    for replay in range(3):
        failure.fill_(1 if kind == "sample" else 0)
        states.fill_(-1)
        log_z.fill_(float("nan"))
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert int(failure.item()) == 0, replay
        assert np.array_equal(log_z.cpu().numpy(), want[0]), replay
        assert np.array_equal(states.cpu().numpy(), want[1]), replay
    del graph
    ws.close()


@pytest.mark.parametrize("k", [8, 16, 32])
def test_map_rake_keeps_a_nan_first_column(require_gpu, k):
    """The reference's argmax (src/inference.cc:113-126) starts at column 0's
    candidate: when that candidate is NaN (here log 0 + inf: a +inf node
    potential in state 0 meeting a zero edge column), no later candidate
    compares greater and every row of the rake keeps choice 0. The team
    rake kernels must do the same rather than pick the best finite column."""
    import torch

    plan = thmm.Plan(np.array([-1, 0, 0], np.int64))
    nodes = torch.ones((1, 3, k), dtype=torch.float32, device="cuda")
    nodes[0, 1, 0] = float("inf")
    nodes[0, 1, 5] = 50.0
    edges = torch.full((2, k, k), 0.5, dtype=torch.float32, device="cuda")
    edges[:, :, 5] = 2.0  # column 5 dominates every finite row
    edge1 = int(np.flatnonzero(plan.edge_children == 1)[0])
    edges[edge1, :, 0] = 0.0
    ws = thmm.Workspace()
    ws.reserve(plan, k, 1, precision="f32", flags=_abi.THMM_RESERVE_MAXIMUM | _abi.THMM_RESERVE_DEVICE_RESIDENT)
    weights = torch.empty(1, dtype=torch.float64, device="cuda")
    states = torch.full((1, 3), 99, dtype=torch.int32, device="cuda")
    ws.maximum_device(nodes.data_ptr(), edges.data_ptr(), 1, weights.data_ptr(), states.data_ptr(), 0)
    torch.cuda.synchronize()
    ws.close()
    got = states.cpu().numpy()[0]
    assert got[1] == 0, f"leaf with a NaN first candidate took state {got[1]}"
    assert got[2] == 5
    assert np.isnan(weights.cpu().numpy()[0])
