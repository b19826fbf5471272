"""FP32 kernels on potentials of wide dynamic range. Log potentials drawn
from N(0, sigma^2) with sigma 12 and 16 put row maxima of composed paths and
of gradient ratios far below 2^-100 and ratios adj/S above 1e30: the FP32
fast paths then hand over (wave-uniformly) to their FP64 forms, and
normalisers below 2^-100 are taken after an exact 2^100 scaling. Without
those, the K = 16 grouped compressions and the K = 32 / 64 wide compression
adjoint returned NaN gradients here.

Checked against the FP64 oracle on the same (FP32-rounded) inputs: every
output finite, |dlogZ| <= 1e-5 * max(1, |logZ|), gradients (posterior
marginals times the seed 1) within 1e-5 absolute, through the log-potential
adjoint pair at compress_depth 1, 2 and 3 (K = 16) and the wide kernels
(K = 32, 64), and through thmm_marginals_device on linear potentials (edge
entries within e^+-40: linear edges enter FP32 products unscaled).
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
import parallel_tree_hmm_amd as thmm
from conftest import caterpillar, forest, path, random_recursive

pytestmark = pytest.mark.gpu


def make_plan(rng):
    return thmm.Plan(forest([caterpillar(401), path(300), random_recursive(300, rng)]))


def check(got_lz, got_nodes, got_edges, olz, onm, oem):
    assert np.isfinite(got_lz).all() and np.isfinite(got_nodes).all() and np.isfinite(got_edges).all()
    assert np.all(np.abs(got_lz - olz) <= 1e-5 * np.maximum(1.0, np.abs(olz)))
    assert np.max(np.abs(got_nodes - onm)) <= 1e-5
    assert np.max(np.abs(got_edges - oem)) <= 1e-5


@pytest.mark.parametrize("k,depth", [(16, 1), (16, 2), (16, 3), (32, 2), (64, 2)])
@pytest.mark.parametrize("sigma", [12.0, 16.0])
def test_log_potential_pair_on_wide_ranges(require_gpu, k, depth, sigma):
    import torch

    rng = np.random.default_rng(int(100 * sigma) + k + depth)
    plan = make_plan(rng)
    log_nodes = rng.normal(scale=sigma, size=(1, plan.num_nodes, k)).astype(np.float32)
    log_edges = rng.normal(scale=sigma, size=(plan.num_edges, k, k)).astype(np.float32)
    olz, onm, oem = oracle.marginals_batched(plan, k, np.exp(log_nodes.astype(np.float64)),
                                             np.exp(log_edges.astype(np.float64)), threads=8)
    ws = thmm.Workspace()
    ws.reserve_marginals(plan, k, 1, precision="f32", device_resident=True)
    ws.set_tuning(compress_depth=depth)
    d_nodes = torch.tensor(log_nodes, device="cuda")
    d_edges = torch.tensor(log_edges, device="cuda")
    lz = torch.empty(plan.roots.size, dtype=torch.float64, device="cuda")
    seed = torch.ones(plan.roots.size, dtype=torch.float64, device="cuda")
    gn = torch.empty((1, plan.num_nodes, k), device="cuda")
    ge = torch.empty((1, plan.num_edges, k, k), device="cuda")
    ws.logz_forward_device(d_nodes.data_ptr(), d_edges.data_ptr(), 1, lz.data_ptr(), 0)
    ws.logz_backward_device(seed.data_ptr(), gn.data_ptr(), ge.data_ptr(), False, 0)
    torch.cuda.synchronize()
    got = (lz.cpu().numpy(), gn.cpu().double().numpy(), ge.cpu().double().numpy())
    print(f"K {k} depth {depth} sigma {sigma}: |dnode| {np.max(np.abs(got[1] - onm)):.2e} "
          f"|dedge| {np.max(np.abs(got[2] - oem)):.2e}")
    check(got[0], got[1], got[2], olz[0], onm, oem)
    ws.close()


@pytest.mark.parametrize("k", [16, 64])
def test_linear_marginals_on_wide_ranges(require_gpu, k):
    import torch

    rng = np.random.default_rng(k)
    plan = make_plan(rng)
    # Linear edge entries enter the FP32 products unscaled: K times a product
    # of two of them must stay below the FP32 maximum, so |log| <= 40 here
    # (nodes move to the log domain first and may span more).
    nodes = np.exp(np.clip(rng.normal(scale=12.0, size=(1, plan.num_nodes, k)), -80, 80)).astype(np.float32)
    edges = np.exp(np.clip(rng.normal(scale=12.0, size=(plan.num_edges, k, k)), -40, 40)).astype(np.float32)
    olz, onm, oem = oracle.marginals_batched(plan, k, nodes.astype(np.float64), edges.astype(np.float64), threads=8)
    ws = thmm.Workspace()
    ws.reserve_marginals(plan, k, 1, precision="f32", device_resident=True)
    d_nodes = torch.tensor(nodes, device="cuda")
    d_edges = torch.tensor(edges, device="cuda")
    lz = torch.empty(plan.roots.size, dtype=torch.float64, device="cuda")
    nm = torch.empty((1, plan.num_nodes, k), device="cuda")
    em = torch.empty((1, plan.num_edges, k, k), device="cuda")
    ws.marginals_device(d_nodes.data_ptr(), d_edges.data_ptr(), 1, lz.data_ptr(), nm.data_ptr(), em.data_ptr(), 0)
    torch.cuda.synchronize()
    check(lz.cpu().numpy(), nm.cpu().double().numpy(), em.cpu().double().numpy(), olz[0], onm, oem)
    ws.close()
