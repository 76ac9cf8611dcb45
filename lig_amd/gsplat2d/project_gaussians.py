"""Projection of 2-D Gaussians from a covariance (upper triangle a, b, c).

Mirrors gsplat2d/gsplat2d/project_gaussians.py:13-104 of the reference: same
signature, asserts, saved tensors and gradient positions.
"""
from typing import Optional, Tuple

from torch import Tensor
from torch.autograd import Function

from . import cuda as _C


def project_gaussians(
    cov2d: Tensor,
    means2d: Tensor,
    img_height: int,
    img_width: int,
    block_width: int,
    opacities: Optional[Tensor] = None,
) -> Tuple[Tensor, Tensor, Tensor, Tensor]:
    """Returns (xys [N,2], extents [N,2], conics [N,3], num_tiles_hit [N] int32)."""
    assert block_width > 1 and block_width <= 16, "block_width must be between 2 and 16"
    return _ProjectGaussians.apply(
        cov2d.contiguous(),
        means2d.contiguous(),
        opacities.contiguous() if opacities is not None else None,
        img_height,
        img_width,
        block_width,
    )


class _ProjectGaussians(Function):
    @staticmethod
    def forward(ctx, cov2d, means2d, opacities, img_height, img_width, block_width):
        num_points = cov2d.shape[-2]
        if num_points < 1 or cov2d.shape[-1] != 3:
            raise ValueError(f"Invalid shape for cov2d: {cov2d.shape}")
        xys, extents, conics, num_tiles_hit = _C.project_gaussians_forward(
            num_points, cov2d, means2d, opacities, img_height, img_width, block_width)
        ctx.img_height = img_height
        ctx.img_width = img_width
        ctx.num_points = num_points
        ctx.save_for_backward(extents, conics)
        ctx.mark_non_differentiable(num_tiles_hit)
        return xys, extents, conics, num_tiles_hit

    @staticmethod
    def backward(ctx, v_xys, v_extents, v_conics, v_num_tiles_hit):
        extents, conics = ctx.saved_tensors
        if v_xys is None:
            v_xys = extents.new_zeros(extents.shape)
        if v_conics is None:
            v_conics = conics.new_zeros(conics.shape)
        v_cov2d, v_mean2d = _C.project_gaussians_backward(
            ctx.num_points, extents, conics, v_xys, v_conics)
        return v_cov2d, v_mean2d, None, None, None, None
