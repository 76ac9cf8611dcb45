"""Sum rasterisation of 2-D Gaussians (pixel = sum_i c_i * alpha_i).

Mirrors gsplat2d/gsplat2d/rasterize.py:13-197 of the reference: same signature,
asserts and errors, the ones-image branch when nothing intersects, the saved
tensors, `xys.absgrad`, and the gradient positions.  Binning inside the forward
uses the fused tile-key path (every depth is zero here, rasterize.py:158), which
yields the same bins as bin_and_sort_gaussians.
"""
from typing import Optional, Tuple

import torch
from torch import Tensor
from torch.autograd import Function

from . import cuda as _C
from .utils import compute_cumulative_intersects


def rasterize_gaussians(
    xys: Tensor,
    extents: Tensor,
    conics: Tensor,
    num_tiles_hit: Tensor,
    colors: Tensor,
    opacities: Optional[Tensor],
    img_height: int,
    img_width: int,
    block_width: int,
    compute_upscale_gradients: bool = False,
) -> Tuple[Tensor, Tensor, Tensor, Tensor, Tensor]:
    """Returns (out_img [H,W,3], out_wsum [H,W], out_dx, out_dy, out_dxy); the
    last three are [H,W,3] when compute_upscale_gradients, else None."""
    assert block_width > 1 and block_width <= 16, "block_width must be between 2 and 16"
    if colors.dtype == torch.uint8:
        colors = colors.float() / 255
    if xys.ndimension() != 2 or xys.size(1) != 2:
        raise ValueError("xys must have dimensions (N, 2)")
    if colors.ndimension() != 2:
        raise ValueError("colors must have dimensions (N, D)")
    if colors.size(1) != 3:
        raise ValueError("colors must have 3 channels (N, 3)")
    return _RasterizeGaussians.apply(
        xys.contiguous(),
        extents.contiguous(),
        conics.contiguous(),
        num_tiles_hit.contiguous(),
        colors.contiguous(),
        opacities.contiguous() if opacities is not None else None,
        img_height,
        img_width,
        block_width,
        compute_upscale_gradients,
    )


class _RasterizeGaussians(Function):
    """Rasterizes 2D gaussians"""

    @staticmethod
    def forward(ctx, xys, extents, conics, num_tiles_hit, colors, opacities, img_height,
                img_width, block_width, compute_upscale_gradients):
        ctx.set_materialize_grads(False)
        num_points = xys.size(0)
        dev = xys.device
        num_intersects, cum_tiles_hit = compute_cumulative_intersects(num_tiles_hit)
        packed = emission_ids = None
        if num_intersects < 1:
            out_img = torch.ones(img_height, img_width, colors.shape[-1], device=dev)
            out_wsum = torch.zeros(img_height, img_width, device=dev)
            out_dx = torch.zeros(img_height, img_width, colors.shape[-1], device=dev)
            out_dy = torch.zeros(img_height, img_width, colors.shape[-1], device=dev)
            out_dxy = torch.zeros(img_height, img_width, colors.shape[-1], device=dev)
            gaussian_ids_sorted = torch.zeros(0, 1, device=dev)
            tile_bins = torch.zeros(0, 2, device=dev)
            final_idx = torch.zeros(img_height, img_width, device=dev)
        else:
            # 16 x 16 tiles without opacities (no zero-extent phantom slots): the
            # sort carries gids and the pack addresses the backward's records;
            # otherwise the sort carries emission slots
            slots_in_pack = block_width == 16 and opacities is None
            if slots_in_pack:
                gaussian_ids_sorted, tile_bins = _C.bin_and_sort_fast(
                    num_points, num_intersects, xys, extents, cum_tiles_hit, img_height,
                    img_width, block_width)
            else:
                gaussian_ids_sorted, tile_bins, emission_ids = _C.bin_and_sort_fast(
                    num_points, num_intersects, xys, extents, cum_tiles_hit, img_height,
                    img_width, block_width, with_emission_ids=True)
            # 16 x 16 plain: no final index (the backward re-derives the
            # forward's accepted pairs with the same test and groups each
            # tile's Gaussians by the rows they reach)
            (out_img, out_wsum, out_dx, out_dy, out_dxy, final_idx,
             packed) = _C.rasterize_forward_packed(
                img_height, img_width, block_width, gaussian_ids_sorted, tile_bins, xys, conics,
                colors, opacities, compute_upscale_gradients,
                extents=extents if slots_in_pack else None,
                cum_tiles_hit=cum_tiles_hit if slots_in_pack else None,
                final_index=not (block_width == 16 and not compute_upscale_gradients))

        ctx.img_width = img_width
        ctx.img_height = img_height
        ctx.num_intersects = num_intersects
        ctx.block_width = block_width
        ctx.compute_upscale_gradients = compute_upscale_gradients
        ctx.packed = packed
        ctx.emission_ids = emission_ids
        ctx.slots_in_pack = packed is not None and emission_ids is None
        ctx.cum_tiles_hit = cum_tiles_hit
        ctx.save_for_backward(gaussian_ids_sorted, tile_bins, xys, conics, colors, opacities,
                              final_idx)
        return out_img, out_wsum, out_dx, out_dy, out_dxy

    @staticmethod
    def backward(ctx, v_out_img, v_out_wsum, v_out_dx, v_out_dy, v_out_dxy):
        (gaussian_ids_sorted, tile_bins, xys, conics, colors, opacities,
         final_idx) = ctx.saved_tensors
        if ctx.num_intersects < 1:
            v_xy = torch.zeros_like(xys)
            v_xy_abs = torch.zeros_like(xys)
            v_conic = torch.zeros_like(conics)
            v_colors = torch.zeros_like(colors)
            v_opacity = torch.zeros_like(opacities) if opacities is not None else None
        else:
            H, W = ctx.img_height, ctx.img_width
            if v_out_img is None:
                v_out_img = torch.zeros(H, W, 3, device=xys.device)
            ga = ctx.compute_upscale_gradients
            if ga:
                zeros = None
                if v_out_dx is None or v_out_dy is None or v_out_dxy is None:
                    zeros = torch.zeros(H, W, 3, device=xys.device)
                v_out_dx = zeros if v_out_dx is None else v_out_dx
                v_out_dy = zeros if v_out_dy is None else v_out_dy
                v_out_dxy = zeros if v_out_dxy is None else v_out_dxy
            packed = ctx.packed
            if packed is None:
                packed = _C.pack_gaussians(xys, conics, colors, opacities)
            v_xy, v_xy_abs, v_conic, v_colors, v_opacity = _C.rasterize_backward_packed(
                H, W, ctx.block_width, gaussian_ids_sorted, tile_bins, packed, xys.size(0),
                final_idx, v_out_img, v_out_wsum, opacities is not None,
                v_out_dx if ga else None, v_out_dy if ga else None, v_out_dxy if ga else None,
                emission_ids_sorted=ctx.emission_ids, cum_tiles_hit=ctx.cum_tiles_hit,
                slots_in_pack=ctx.slots_in_pack and packed is ctx.packed)
        ctx.packed = ctx.emission_ids = ctx.cum_tiles_hit = None
        xys.absgrad = v_xy_abs
        return (v_xy, None, v_conic, None, v_colors, v_opacity, None, None, None, None)
