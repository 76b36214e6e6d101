"""Differentiable warp: ``WarpingOpFunction`` / ``warp`` with the contract of
svs/src/warpingOps.py:13-29 (returns grad_image, grad_flow, None), backed by
the hand-written adjoint kernels.  Works with Warping3D (image [Z,Y,X], flow
[Z,Y,X,3]) and with WarpingCNN3D (image [B,C,Z,Y,X], flow [B,Z,Y,X,3]).
"""
import torch


class WarpingOpFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, image, flow, warpingOp):
        image = image.contiguous()
        flow = flow.contiguous()
        ctx.save_for_backward(image, flow)
        ctx.warpingOp = warpingOp
        return warpingOp.forward(image, flow)

    @staticmethod
    def backward(ctx, grad_out):
        image, flow = ctx.saved_tensors
        grad_image, grad_flow = ctx.warpingOp.backward(image, flow, grad_out.contiguous())
        return grad_image, grad_flow, None


def warp(image: torch.Tensor, flow: torch.Tensor, warpingOp) -> torch.Tensor:
    return WarpingOpFunction.apply(image, flow, warpingOp)
