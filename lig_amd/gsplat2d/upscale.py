"""Gradient-aware bicubic (Hermite) upscaling from the analytic image
derivatives the rasteriser emits.

Mirrors gsplat2d/gsplat2d/upscale.py:10-82 of the reference.
"""
from typing import Optional, Tuple

from torch import Tensor
from torch.autograd import Function

from . import cuda as _C


class _GradientAwareSplineUpscale(Function):
    @staticmethod
    def forward(ctx, render, dx, dy, dxy, dst_h, dst_w, roi):
        ctx.save_for_backward(render, dx, dy, dxy)
        ctx.dst_h = dst_h
        ctx.dst_w = dst_w
        ctx.roi = roi
        return _C.gradient_aware_upscale_forward(render, dx, dy, dxy, dst_h, dst_w, roi)

    @staticmethod
    def backward(ctx, grad_output):
        render, dx, dy, dxy = ctx.saved_tensors
        grad_render, grad_dx, grad_dy, grad_dxy = _C.gradient_aware_upscale_backward(
            grad_output.contiguous(), render, dx, dy, dxy, ctx.dst_h, ctx.dst_w, ctx.roi)
        return grad_render, grad_dx, grad_dy, grad_dxy, None, None, None


def gradient_aware_upscale(
    render: Tensor,
    dx: Tensor,
    dy: Tensor,
    dxy: Tensor,
    dst_h: int,
    dst_w: int,
    roi: Optional[Tuple[float, float, float, float]] = None,
) -> Tensor:
    """Upscale an [H, W, 3] render to [dst_h, dst_w, 3] from (f, df/dx, df/dy,
    d2f/dxdy); roi = (x1, y1, x2, y2) in source pixels, default the full image."""
    h, w, c = render.shape
    if roi is None:
        roi = (0.0, 0.0, float(w), float(h))
    return _GradientAwareSplineUpscale.apply(render.contiguous(), dx.contiguous(),
                                             dy.contiguous(), dxy.contiguous(), dst_h, dst_w,
                                             roi)
