"""The uneven-forest workload at full size: 4,096 trees with node counts
log-uniform in [16, 65,536] (about 33 M nodes), topologies cycling random
recursive / caterpillar / star, K = 16, FP32, inputs from the portable
SplitMix64 / Box-Muller recipe (node potentials exp(N(0,1)), edge rows
Dirichlet(1), one variate per node), contracted as one forest.

Per call (device-resident entry points):

* marginals: log Z, node and edge marginals against the FP64 oracle on whole
  trees (the largest tree of each topology and ten small ones) within 1e-5;
  over the whole forest, every node marginal sums to one and every edge
  marginal's rows and columns sum to the parent and child marginals; a
  THMM_RESERVE_GRAPH workspace replays bit-equal to the eager call, with
  eager work between the replays;
* MAP: states bit-exact against the oracle on the same trees, and over the
  whole forest every tree's MAP log weight equals the log likelihood of the
  returned states;
* sampling: draws bit-exact against the oracle on the same trees, variates
  within 1e-12 of a CDF boundary counted separately; states in range and no
  failure over the whole forest.

The oracle runs over the plan of each tree alone; the forest's plan holds
the same per-tree structure, so MAP tie resolution and draws are comparable
op for op (tests/accelerator_test.h:652-790 checks MAP and samples exactly).
"""

from __future__ import annotations

import numpy as np
import pytest

import bench
import oracle
import parallel_tree_hmm_amd as thmm

pytestmark = pytest.mark.gpu

K = 16
TOL = 1e-5


@pytest.fixture(scope="module")
def forest(require_gpu):
    import torch

    bench.check_device_stream(torch)
    parts = bench.forest_parts(4)
    parents, offsets = bench.concatenate(parts)
    plan = thmm.Plan(parents)
    n, e = plan.num_nodes, plan.num_edges
    print(f"forest: {len(parts)} trees, {n} nodes, {plan.num_compressions} compressions", flush=True)
    nodes = bench.fill_lognormal_device(torch, torch.empty((n, K), device="cuda"), bench.config_seed(4, bench.STREAM_NODES))
    edges = bench.fill_dirichlet_rows_device(torch, torch.empty((e, K, K), device="cuda"),
                                             bench.config_seed(4, bench.STREAM_EDGES))
    variates = bench.variates_device(torch, bench.config_seed(4, bench.STREAM_VARIATES), n)
    torch.cuda.synchronize()
    sizes = np.array([p.size for p in parts])
    kinds = np.arange(len(parts)) % 3
    picks = [int(np.flatnonzero(kinds == kind)[np.argmax(sizes[kinds == kind])]) for kind in range(3)]
    small = [int(t) for t in np.flatnonzero(sizes < 2000)[:10]]
    tree_of_node = torch.from_numpy(np.repeat(np.arange(len(parts)), sizes)).cuda()
    yield dict(torch=torch, parts=parts, offsets=offsets, plan=plan, nodes=nodes, edges=edges, variates=variates,
               picks=picks + small, tree_of_node=tree_of_node)
    del nodes, edges, variates, tree_of_node
    torch.cuda.empty_cache()


def host_tree(f, t, *arrays):
    lo, hi, elo, ehi = bench.tree_slices(f["offsets"], t)
    out = [f["nodes"][lo:hi].double().cpu().numpy(), f["edges"][elo:ehi].double().cpu().numpy()]
    return out + [a(lo, hi, elo, ehi) for a in arrays]


def test_forest_marginals_and_graph_replay(forest):
    torch, plan = forest["torch"], forest["plan"]
    n, e = plan.num_nodes, plan.num_edges
    ws = thmm.Workspace()
    ws.reserve_marginals(plan, K, 1, precision="f32", device_resident=True)
    log_z = torch.empty(plan.roots.size, dtype=torch.float64, device="cuda")
    node_out = torch.empty((n, K), device="cuda")
    edge_out = torch.empty((e, K, K), device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    ws.marginals_device(forest["nodes"].data_ptr(), forest["edges"].data_ptr(), 1, log_z.data_ptr(),
                        node_out.data_ptr(), edge_out.data_ptr(), stream)
    torch.cuda.synchronize()
    lz = log_z.cpu().numpy()
    worst = {"log_z": 0.0, "nodes": 0.0, "edges": 0.0}
    for t in forest["picks"]:
        nn, ee, gn, ge = host_tree(forest, t, lambda lo, hi, elo, ehi: node_out[lo:hi].double().cpu().numpy(),
                                   lambda lo, hi, elo, ehi: edge_out[elo:ehi].double().cpu().numpy())
        olz, onm, oem = oracle.marginals(thmm.Plan(forest["parts"][t]), K, nn, ee)
        worst["log_z"] = max(worst["log_z"], abs(lz[t] - olz[0]) / max(1.0, abs(olz[0])))
        worst["nodes"] = max(worst["nodes"], float(np.max(np.abs(gn - onm))))
        if ee.size:
            worst["edges"] = max(worst["edges"], float(np.max(np.abs(ge - oem))))
    print(f"oracle parity on {len(forest['picks'])} trees: {worst}", flush=True)
    assert worst["log_z"] <= TOL and worst["nodes"] <= TOL and worst["edges"] <= TOL
    # Whole-forest invariants.
    assert np.all(np.isfinite(lz))
    assert float((node_out.double().sum(1) - 1.0).abs().max()) <= TOL
    eparents = torch.from_numpy(plan.edge_parents.astype(np.int64)).cuda()
    echildren = torch.from_numpy(plan.edge_children.astype(np.int64)).cuda()
    rows = cols = 0.0
    step = 1 << 22
    for i in range(0, e, step):
        block = edge_out[i:i + step].double()
        rows = max(rows, float((block.sum(2) - node_out[eparents[i:i + step]].double()).abs().max()))
        cols = max(cols, float((block.sum(1) - node_out[echildren[i:i + step]].double()).abs().max()))
        del block
    print(f"edge row / column sums against the endpoint marginals: {rows:.2e} {cols:.2e}", flush=True)
    assert rows <= TOL and cols <= TOL
    # hipGraph replay of the same call, bit-equal to the eager outputs.
    wg = thmm.Workspace()
    wg.reserve_marginals(plan, K, 1, precision="f32", device_resident=True, graphs=True)
    glz = torch.full_like(log_z, float("nan"))
    gnode = torch.full_like(node_out, float("nan"))
    gedge = torch.empty_like(edge_out)
    for replay in range(2):
        gedge.fill_(float("nan"))
        wg.marginals_device(forest["nodes"].data_ptr(), forest["edges"].data_ptr(), 1, glz.data_ptr(),
                            gnode.data_ptr(), gedge.data_ptr(), stream)
        torch.cuda.synchronize()
        assert torch.equal(glz, log_z) and torch.equal(gnode, node_out) and torch.equal(gedge, edge_out), replay
        # eager work of the program between the replays
        gnode.zero_()
        torch.cuda.synchronize()
    wg.close()
    ws.close()


def test_forest_map(forest):
    torch, plan = forest["torch"], forest["plan"]
    n = plan.num_nodes
    ws = thmm.Workspace()
    ws.reserve(plan, K, 1, precision="f32",
               flags=thmm._abi.THMM_RESERVE_MAXIMUM | thmm._abi.THMM_RESERVE_DEVICE_RESIDENT)
    log_w = torch.empty(plan.roots.size, dtype=torch.float64, device="cuda")
    states = torch.empty(n, dtype=torch.int32, device="cuda")
    ws.maximum_device(forest["nodes"].data_ptr(), forest["edges"].data_ptr(), 1, log_w.data_ptr(), states.data_ptr(),
                      torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    ws.close()
    lw = log_w.cpu().numpy()
    assert bool(((states >= 0) & (states < K)).all())
    for t in forest["picks"]:
        nn, ee, gs = host_tree(forest, t, lambda lo, hi, elo, ehi: states[lo:hi].cpu().numpy().astype(np.uint32))
        olw, ostates = oracle.maximum(thmm.Plan(forest["parts"][t]), K, nn, ee)
        assert abs(lw[t] - olw[0]) <= 1e-12 * max(1.0, abs(olw[0])), t
        assert np.array_equal(gs, ostates), f"MAP states of tree {t} ({nn.shape[0]} nodes) differ"
    # Whole forest: the MAP log weight of every tree is the log likelihood of
    # its states (FP64, summed per tree).
    s = states.long()
    node_terms = torch.log(forest["nodes"].double().gather(1, s[:, None])[:, 0])
    ep = torch.from_numpy(plan.edge_parents.astype(np.int64)).cuda()
    ec = torch.from_numpy(plan.edge_children.astype(np.int64)).cuda()
    flat = forest["edges"].view(-1, K * K)
    edge_terms = torch.log(flat.gather(1, (s[ep] * K + s[ec])[:, None])[:, 0].double())
    tree = forest["tree_of_node"]
    total = torch.zeros(plan.roots.size, dtype=torch.float64, device="cuda")
    total.index_add_(0, tree, node_terms)
    total.index_add_(0, tree[ec], edge_terms)
    gap = float(((total - log_w).abs() / log_w.abs().clamp(min=1.0)).max())
    print(f"MAP weight vs likelihood of the states, whole forest: {gap:.2e}", flush=True)
    assert gap <= 1e-10


def test_forest_samples(forest):
    torch, plan = forest["torch"], forest["plan"]
    n = plan.num_nodes
    ws = thmm.Workspace()
    ws.reserve(plan, K, 1, precision="f32",
               flags=thmm._abi.THMM_RESERVE_SAMPLING | thmm._abi.THMM_RESERVE_DEVICE_RESIDENT)
    log_z = torch.empty(plan.roots.size, dtype=torch.float64, device="cuda")
    states = torch.empty(n, dtype=torch.int32, device="cuda")
    failure = torch.ones(1, dtype=torch.int32, device="cuda")
    ws.sample_device(forest["nodes"].data_ptr(), forest["edges"].data_ptr(), forest["variates"].data_ptr(), 1,
                     log_z.data_ptr(), states.data_ptr(), failure.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    ws.close()
    assert int(failure.item()) == 0
    assert bool(torch.isfinite(log_z).all())
    assert bool(((states >= 0) & (states < K)).all())
    compared = boundary = 0
    for t in forest["picks"]:
        nn, ee, gs, u = host_tree(forest, t, lambda lo, hi, elo, ehi: states[lo:hi].cpu().numpy().astype(np.uint32),
                                  lambda lo, hi, elo, ehi: forest["variates"][lo:hi].double().cpu().numpy())
        expected, dist = oracle.sample(thmm.Plan(forest["parts"][t]), K, nn, ee, u)
        near = dist < 1e-12
        boundary += int(near.sum())
        if near.any():
            continue  # reported separately, not compared
        compared += 1
        assert np.array_equal(gs, expected), f"draws of tree {t} ({nn.shape[0]} nodes) differ"
    print(f"sampled trees compared bit-exactly: {compared}; variates within 1e-12 of a boundary: {boundary}", flush=True)
    assert compared >= len(forest["picks"]) - 1
