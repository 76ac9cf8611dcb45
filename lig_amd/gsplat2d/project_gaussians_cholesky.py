"""Projection of 2-D Gaussians from a Cholesky factor (l11, l21, l22).

Mirrors gsplat2d/gsplat2d/project_gaussians_cholesky.py:10-112 of the reference.
"""
from typing import Optional, Tuple

from torch import Tensor
from torch.autograd import Function

from . import cuda as _C


def project_gaussians_cholesky(
    cholesky: Tensor,
    means2d: Tensor,
    img_height: int,
    img_width: int,
    block_width: int,
    opacities: Optional[Tensor] = None,
) -> Tuple[Tensor, Tensor, Tensor, Tensor]:
    """Sigma = L L^T with L = [[l11, 0], [l21, l22]].  Returns (xys, extents,
    conics, num_tiles_hit)."""
    assert block_width > 1 and block_width <= 16, "block_width must be between 2 and 16"
    return _ProjectGaussiansCholesky.apply(
        cholesky.contiguous(),
        means2d.contiguous(),
        opacities.contiguous() if opacities is not None else None,
        img_height,
        img_width,
        block_width,
    )


class _ProjectGaussiansCholesky(Function):
    @staticmethod
    def forward(ctx, cholesky, means2d, opacities, img_height, img_width, block_width):
        num_points = cholesky.shape[-2]
        if num_points < 1 or cholesky.shape[-1] != 3:
            raise ValueError(f"Invalid shape for cholesky: {cholesky.shape}")
        xys, extents, conics, num_tiles_hit = _C.project_gaussians_forward_cholesky(
            num_points, cholesky, means2d, opacities, img_height, img_width, block_width)
        ctx.img_height = img_height
        ctx.img_width = img_width
        ctx.num_points = num_points
        ctx.save_for_backward(extents, cholesky, conics)
        ctx.mark_non_differentiable(num_tiles_hit)
        return xys, extents, conics, num_tiles_hit

    @staticmethod
    def backward(ctx, v_xys, v_extents, v_conics, v_num_tiles_hit):
        extents, cholesky, conics = ctx.saved_tensors
        if v_xys is None:
            v_xys = extents.new_zeros(extents.shape)
        if v_conics is None:
            v_conics = conics.new_zeros(conics.shape)
        v_cholesky, v_mean2d = _C.project_gaussians_backward_cholesky(
            ctx.num_points, extents, cholesky, conics, v_xys, v_conics)
        return v_cholesky, v_mean2d, None, None, None, None
