"""The inviscid residual as a differentiable function for PyTorch callers:
forward through fest_residual, backward through fest_vjp (both in
libfest_hip.so, on the block's device).  The tensors are float64 host
tensors of the interior cells; the ghost cells follow from the block's
boundary rules at every evaluation, so they are not inputs (a block with
interface faces would need its neighbours' cells: not supported here).

    block = GpuBlock(case); block.set_grid(nodes)
    r = inviscid_residual(q_interior, block)    # (5, kmx-1, jmx-1, imx-1)
    r.backward(w)                               # q_interior.grad = (dR/dq)^T w
"""
import numpy as np

INNER = (slice(None),) + (slice(3, -3),) * 3


def _torch():
    import torch  # only the callers of this module need torch
    return torch


def _function():
    torch = _torch()

    class InviscidResidual(torch.autograd.Function):
        @staticmethod
        def forward(ctx, q_interior, block):
            if block.case.mu_ref != 0.0:
                # the ghost cells are uploaded as zeros: the viscous gradients
                # would read their temperatures
                raise ValueError("inviscid_residual takes inviscid blocks (mu_ref 0)")
            q = np.zeros((5,) + block.case.cell_shape())
            q[INNER] = q_interior.detach().cpu().numpy()
            block.upload_state(q)
            block.residual(1)
            ctx.block, ctx.q = block, q
            return torch.from_numpy(block.download("residue"))

        @staticmethod
        def backward(ctx, grad_residue):
            block = ctx.block
            block.upload_state(ctx.q)  # the state this residual was taken at
            bar = block.vjp(grad_residue.detach().cpu().numpy())
            return torch.from_numpy(np.ascontiguousarray(bar[INNER])), None

    return InviscidResidual


_fn = None


def inviscid_residual(q_interior, block):
    """R(q) of the interior state q (float64 tensor, (5, kmx-1, jmx-1,
    imx-1)) on `block` (festgpu.native.GpuBlock with its grid set)."""
    global _fn
    if _fn is None:
        _fn = _function()
    return _fn.apply(q_interior, block)
