"""Test harness: the trilinear raycast as a torch.autograd.Function.

forward(tsdf, xi) renders the camera depth map at T = exp(xi^) T0 through
the C ABI (kf_raycast in KF_RAYCAST_TRILINEAR mode); backward calls
kf_raycast_adjoint, which (the inputs are uploaded again first, which drops
the forward's saved crossings) runs the trilinear raycast into its own maps,
then the adjoint from those crossings, and returns dL/dTSDF and dL/dxi.  dL/dxi is the left-perturbation gradient at T, i.e. the gradient
with respect to xi at xi = 0.  torch only carries the arrays here; every
computation runs in libkinfu_amd.so.
"""
import numpy as np
import torch

from se3 import se3_exp


class TrilinearDepth(torch.autograd.Function):
    @staticmethod
    def forward(ctx, tsdf, xi, kctx, K, T0, weight, guide):
        T = (se3_exp(xi.detach().cpu().numpy().astype(np.float64)) @ T0).astype(np.float32)
        kctx.upload_volume(tsdf.detach().cpu().numpy(), weight)
        kctx.upload_depth_f32(guide)
        kctx.raycast(K, T)
        ctx.kctx, ctx.K, ctx.T, ctx.guide = kctx, K, T, guide
        ctx.weight = weight
        ctx.tsdf = tsdf.detach().cpu().numpy()
        return torch.from_numpy(kctx.rendered_depth())

    @staticmethod
    def backward(ctx, grad_depth):
        k = ctx.kctx
        # the context may have rendered other poses since: restore this call's inputs
        k.upload_volume(ctx.tsdf, ctx.weight)
        k.upload_depth_f32(ctx.guide)
        gF, gxi = k.raycast_adjoint(ctx.K, ctx.T, grad_depth.detach().cpu().numpy())
        return torch.from_numpy(gF), torch.from_numpy(gxi), None, None, None, None, None


def render_depth(tsdf, xi, kctx, K, T0, weight, guide):
    return TrilinearDepth.apply(tsdf, xi, kctx, K, T0, weight, guide)
