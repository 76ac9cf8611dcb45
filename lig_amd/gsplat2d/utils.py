"""Binning utilities of the gsplat2d API.

Mirrors gsplat2d/gsplat2d/utils.py:12-100 of the reference.  The cumulative sum
and the sort run on this package's kernels (device int32 scan, stable radix
sort) instead of torch.cumsum / torch.sort; the sort is stable, so ties keep
the gid-major emission order.
"""
from typing import Optional, Tuple

import torch
from torch import Tensor

from . import cuda as _C


def map_gaussian_to_intersects(num_points: int, num_intersects: int, xys: Tensor,
                               depths: Tensor, extents: Tensor, cum_tiles_hit: Tensor,
                               tile_bounds: Tuple[int, int, int],
                               block_size: int) -> Tuple[Tensor, Tensor]:
    return _C.map_gaussian_to_intersects(num_points, num_intersects, xys.contiguous(),
                                         depths.contiguous(), extents.contiguous(),
                                         cum_tiles_hit.contiguous(), tile_bounds, block_size)


def get_tile_bin_edges(num_intersects: int, isect_ids_sorted: Tensor,
                       tile_bounds: Tuple[int, int, int]) -> Tensor:
    return _C.get_tile_bin_edges(num_intersects, isect_ids_sorted.contiguous(), tile_bounds)


def compute_cov2d_bounds(cov2d: Tensor,
                         opacities: Optional[Tensor] = None) -> Tuple[Tensor, Tensor]:
    assert cov2d.shape[-1] == 3, (
        "Expected input cov2d to be of shape (*batch, 3) (upper triangular values), "
        f"but got {tuple(cov2d.shape)}")
    num_pts = cov2d.shape[0]
    assert num_pts > 0
    return _C.compute_cov2d_bounds(num_pts, cov2d.contiguous(),
                                   opacities.contiguous() if opacities is not None else None)


def compute_cumulative_intersects(num_tiles_hit: Tensor) -> Tuple[int, Tensor]:
    """Inclusive int32 scan and its last element (a device->host sync, as the
    reference's .item())."""
    cum_tiles_hit = _C.inclusive_cumsum_i32(num_tiles_hit.contiguous().to(torch.int32))
    num_intersects = int(cum_tiles_hit[-1].item()) if cum_tiles_hit.numel() else 0
    return num_intersects, cum_tiles_hit


def bin_and_sort_gaussians(num_points: int, num_intersects: int, xys: Tensor, depths: Tensor,
                           extents: Tensor, cum_tiles_hit: Tensor,
                           tile_bounds: Tuple[int, int, int], block_size: int):
    """Returns (isect_ids, gaussian_ids, isect_ids_sorted, gaussian_ids_sorted, tile_bins)."""
    isect_ids, gaussian_ids = map_gaussian_to_intersects(num_points, num_intersects, xys, depths,
                                                         extents, cum_tiles_hit, tile_bounds,
                                                         block_size)
    isect_ids_sorted, gaussian_ids_sorted = _C.sort_isect_pairs(isect_ids, gaussian_ids)
    tile_bins = get_tile_bin_edges(num_intersects, isect_ids_sorted, tile_bounds)
    return isect_ids, gaussian_ids, isect_ids_sorted, gaussian_ids_sorted, tile_bins
