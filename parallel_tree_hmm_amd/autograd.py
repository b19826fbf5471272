"""torch.autograd.Function over the forward/backward C-ABI pair
(thmm_logz_forward_device / thmm_logz_backward_device).

log Z(log node potentials, log edge potentials) for a batch of items sharing
one plan; the backward pass is the hand-written adjoint kernel, whose
outputs are the posterior marginals scaled by the incoming gradient.
Everything stays on the device; torch provides memory and the stream.
"""

from __future__ import annotations

import torch

from .api import Plan, Workspace


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
        ctx.workspace = workspace
        ctx.reduce_edges = reduce_edges
        ctx.shapes = (log_nodes.shape, log_edges.shape, log_nodes.dtype)
        return log_z

    @staticmethod
    def backward(ctx, grad_log_z: torch.Tensor):
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
