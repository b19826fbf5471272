"""torch.autograd.Functions over the forward/backward C-ABI pairs.

LogPartition (thmm_logz_forward_device / thmm_logz_backward_device): log Z
of log node and edge potentials for a batch of items sharing one plan; the
backward pass is the hand-written adjoint kernel, whose outputs are the
posterior marginals scaled by the incoming gradient. Everything stays on the
device; torch provides memory and the stream.

BranchLengthLogPartition (thmm_categorical_logz_forward /
thmm_branch_length_backward): log Z per site of a categorical model whose
edge factors are P_e = expm(Q t_e), differentiated with respect to the
branch lengths t.

A workspace holds the state of its latest forward call only: a backward
whose forward was followed by another forward on the same workspace raises
instead of returning the gradient of the later call.
"""

from __future__ import annotations

import numpy as np
import torch

from .api import BatchedCategoricalModelView, CategoricalInputUpdate, LogicError, Plan, Workspace


def _claim(workspace: Workspace) -> int:
    generation = getattr(workspace, "_forward_generation", 0) + 1
    workspace._forward_generation = generation
    return generation


def _require_latest(workspace: Workspace, generation: int) -> None:
    if getattr(workspace, "_forward_generation", 0) != generation:
        raise LogicError("the workspace ran another forward call after this one; its contraction state is gone")


class LogPartition(torch.autograd.Function):
    """log_z[b, r] = log Z of batch item b, root r."""

    @staticmethod
    def forward(ctx, log_nodes: torch.Tensor, log_edges: torch.Tensor, workspace: Workspace, reduce_edges: bool):
        if not (log_nodes.is_cuda and log_edges.is_cuda):
            raise ValueError("LogPartition runs on the GPU: pass device tensors")
        log_nodes = log_nodes.contiguous()
        log_edges = log_edges.contiguous()
        batch = log_nodes.shape[0]
        roots = workspace._plan.roots.size
        log_z = torch.empty((batch, roots), dtype=torch.float64, device=log_nodes.device)
        stream = torch.cuda.current_stream().cuda_stream
        workspace.logz_forward_device(log_nodes.data_ptr(), log_edges.data_ptr(), batch, log_z.data_ptr(), stream)
        # The backward pass reads the log edges again (the rake kernels take
        # them in place): keep this exact tensor alive and unmodified.
        ctx.save_for_backward(log_edges)
        ctx.generation = _claim(workspace)
        ctx.workspace = workspace
        ctx.reduce_edges = reduce_edges
        ctx.shapes = (log_nodes.shape, log_edges.shape, log_nodes.dtype)
        return log_z

    @staticmethod
    def backward(ctx, grad_log_z: torch.Tensor):
        _require_latest(ctx.workspace, ctx.generation)
        (_log_edges,) = ctx.saved_tensors  # checks that it was not modified in place
        node_shape, edge_shape, dtype = ctx.shapes
        device = grad_log_z.device
        grad = grad_log_z.to(torch.float64).contiguous()
        grad_nodes = torch.empty(node_shape, dtype=dtype, device=device)
        if ctx.reduce_edges:
            grad_edges = torch.empty(edge_shape, dtype=dtype, device=device)
        else:
            grad_edges = torch.empty((node_shape[0],) + tuple(edge_shape), dtype=dtype, device=device)
        stream = torch.cuda.current_stream().cuda_stream
        ctx.workspace.logz_backward_device(
            grad.data_ptr(), grad_nodes.data_ptr(), grad_edges.data_ptr(), ctx.reduce_edges, stream
        )
        if not ctx.reduce_edges:
            grad_edges = grad_edges.sum(0)
        return grad_nodes, grad_edges, None, None


def log_partition(log_nodes: torch.Tensor, log_edges: torch.Tensor, workspace: Workspace) -> torch.Tensor:
    """Differentiable log Z ([batch, roots], FP64) from log potentials
    ([batch, node, state] and [edge, parent, child], edges broadcast over
    the batch). The workspace must be reserved with reserve_marginals for
    the plan, state count and a capacity >= batch, in the precision of the
    inputs (float32 -> "f32", float64 -> "f64")."""
    return LogPartition.apply(log_nodes, log_edges, workspace, True)


def reserve_for(plan: Plan, states: int, batch: int, dtype: torch.dtype) -> Workspace:
    ws = Workspace()
    ws.reserve_marginals(plan, states, batch, precision="f32" if dtype == torch.float32 else "f64")
    return ws


class BranchLengthLogPartition(torch.autograd.Function):
    """log_z[b * roots + r] of a categorical model with edge factors
    P_e = expm(Q t_e), formed in FP64 on the host."""

    @staticmethod
    def forward(ctx, t: torch.Tensor, rate_matrix: torch.Tensor, model: BatchedCategoricalModelView,
                workspace: Workspace, update: CategoricalInputUpdate):
        q = rate_matrix.detach().to(torch.float64).cpu()
        p = torch.linalg.matrix_exp(q.unsqueeze(0) * t.detach().to(torch.float64).cpu().reshape(-1, 1, 1))
        staged = BatchedCategoricalModelView(
            model.plan, model.states, model.batch, model.categories, model.observation_nodes, model.observations,
            model.root_potential, model.emission_potentials, p.numpy(),
        )
        log_z = workspace.categorical_logz_forward(staged, update)
        ctx.generation = _claim(workspace)
        ctx.workspace = workspace
        ctx.rate = q.numpy()
        ctx.t_meta = (t.dtype, t.device)
        return torch.from_numpy(log_z).to(t.device)

    @staticmethod
    def backward(ctx, grad_log_z: torch.Tensor):
        _require_latest(ctx.workspace, ctx.generation)
        seeds = grad_log_z.detach().to(torch.float64).cpu().numpy()
        grad = ctx.workspace.branch_length_backward(ctx.rate, seeds)
        dtype, device = ctx.t_meta
        return torch.from_numpy(grad).to(device=device, dtype=dtype), None, None, None, None


def branch_length_log_partition(t: torch.Tensor, rate_matrix, model: BatchedCategoricalModelView,
                                workspace: Workspace, update=CategoricalInputUpdate.kAll) -> torch.Tensor:
    """Differentiable per-site log Z ([batch * roots], FP64) as a function of
    the branch lengths t ([num_edges]). `model` supplies the observations,
    root and emission factors (its edge potentials are replaced by
    expm(Q t_e)); the workspace must come from reserve_categorical_marginals.
    Use update=kFactors to reuse observations staged by an earlier call."""
    q = rate_matrix if isinstance(rate_matrix, torch.Tensor) else torch.as_tensor(np.asarray(rate_matrix))
    return BranchLengthLogPartition.apply(t, q, model, workspace, update)
