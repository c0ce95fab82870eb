"""torch.autograd.Function over the gfx950 forward / backward kernels.

``rasterize(means, log_scales, quaternions, opacity_logits, sh, camera, options)`` returns the
rendered colour [3,H,W] and alpha [H,W]; its backward is the hand-written HIP backward (the
transmittance is recomputed back to front from the stored final T and contributor count).
The torch tensors are the storage: nothing is copied on the way in or out.
"""
from __future__ import annotations

from typing import Optional

import torch

from .rasterizer import Camera, GaussianModel, RasterizeOptions, Rasterizer, RenderResult


class _Rasterize(torch.autograd.Function):
    @staticmethod
    def forward(ctx, means, log_scales, quaternions, opacity_logits, sh, camera, options, rasterizer, sh_degree):
        model = GaussianModel(means=means.detach().contiguous(), log_scales=log_scales.detach().contiguous(),
                              quaternions=quaternions.detach().contiguous(),
                              opacity_logits=opacity_logits.detach().contiguous(), sh=sh.detach().contiguous(),
                              sh_degree=sh_degree)
        rendered = rasterizer.forward(model, camera, options)
        # The backward needs the render's context and image shapes, not its images: keeping the
        # returned colour tensor here would close a reference cycle (output -> graph node -> ctx ->
        # output), and the forward's workspace would stay reserved until Python's cycle collector
        # ran, so a training loop would reserve a new workspace every step.
        meta = RenderResult(color=torch.empty(rendered.color.shape, device="meta"),
                            alpha=torch.empty(rendered.alpha.shape, device="meta"), context=rendered.context,
                            rendered_instances=rendered.rendered_instances)
        ctx.model, ctx.rendered, ctx.rasterizer = model, meta, rasterizer
        ctx.opacity_shape = opacity_logits.shape
        return rendered.color, rendered.alpha

    @staticmethod
    def backward(ctx, grad_color, grad_alpha):
        if grad_color is None:
            grad_color = torch.zeros(ctx.rendered.color.shape, device=ctx.model.means.device)
        g = ctx.rasterizer.backward(ctx.model, ctx.rendered, grad_color.contiguous(),
                                    None if grad_alpha is None else grad_alpha.contiguous())
        return (g.means, g.log_scales, g.quaternions, g.opacity_logits.view(ctx.opacity_shape), g.sh,
                None, None, None, None)


_default = None


def rasterize(means: torch.Tensor, log_scales: torch.Tensor, quaternions: torch.Tensor, opacity_logits: torch.Tensor,
              sh: torch.Tensor, camera: Camera, options: Optional[RasterizeOptions] = None,
              rasterizer: Optional[Rasterizer] = None, sh_degree: int = 3):
    """Differentiable render of one pinhole view; returns (color [3,H,W], alpha [H,W])."""
    global _default
    if rasterizer is None:
        _default = _default or Rasterizer()
        rasterizer = _default
    return _Rasterize.apply(means, log_scales, quaternions, opacity_logits, sh, camera,
                            options or RasterizeOptions(), rasterizer, sh_degree)
