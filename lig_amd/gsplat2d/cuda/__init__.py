"""Native function table of the gsplat2d API, backed by liblig_hip.so.

The first eleven functions keep the names, argument lists and return tuples of
the reference's pybind11 module (gsplat2d/gsplat2d/cuda/csrc/ext.cpp:5-17,
declared in bindings.h:19-159): tensors in, new tensors out, `None` where the
reference returns an undefined tensor.  Each allocates its outputs with torch
on the input's device and calls the C-ABI on the current stream.  The functions
after them have no reference counterpart; the autograd wrappers use them to
skip redundant work (the 64-bit key sort, the re-pack in the backward).

Inputs must be HIP tensors: there is no CPU path.
"""
from __future__ import annotations

import torch

from ... import _lib as L

_f32 = torch.float32
_i32 = torch.int32


def _check(name, t, dtype=_f32):
    L.check_device(name, t, dtype)


def _empty(shape, like, dtype=_f32):
    return torch.empty(shape, dtype=dtype, device=like.device)


def _workspace(nbytes, like):
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=like.device)


# ---------------------------------------------------------------------------
# the reference's eleven entry points
# ---------------------------------------------------------------------------

def compute_cov2d_bounds(num_pts, covs2d, opacities):
    """bindings.cu:41-66 -> (conics [n,3], extents [n,2])."""
    _check("covs2d", covs2d)
    _check("opacities", opacities)
    conics = _empty((num_pts, 3), covs2d)
    extents = _empty((num_pts, 2), covs2d)
    with torch.cuda.device(covs2d.device):
        L.call("lig_compute_cov2d_bounds", num_pts, L.ptr(covs2d), L.ptr(opacities),
               L.ptr(conics), L.ptr(extents), L.stream_of(covs2d))
    return conics, extents


def _project_forward(fn, num_points, par, means2d, opacities, img_height, img_width,
                     block_width):
    _check("cov2d", par)
    _check("means2d", means2d)
    _check("opacities", opacities)
    xys = _empty((num_points, 2), par)
    extents = _empty((num_points, 2), par)
    conics = _empty((num_points, 3), par)
    ntiles = _empty((num_points,), par, _i32)
    with torch.cuda.device(par.device):
        L.call(fn, num_points, L.ptr(par), L.ptr(means2d), L.ptr(opacities), int(img_height),
               int(img_width), int(block_width), L.ptr(xys), L.ptr(extents), L.ptr(conics),
               L.ptr(ntiles), L.stream_of(par))
    return xys, extents, conics, ntiles


def project_gaussians_forward(num_points, cov2d, means2d, opacities, img_height, img_width,
                              block_width):
    """bindings.cu:68-117 -> (xys, extents, conics, num_tiles_hit)."""
    return _project_forward("lig_project_gaussians_forward", num_points, cov2d, means2d,
                            opacities, img_height, img_width, block_width)


def project_gaussians_forward_cholesky(num_points, cholesky, means2d, opacities, img_height,
                                       img_width, block_width):
    """bindings.cu:153-202 -> (xys, extents, conics, num_tiles_hit)."""
    return _project_forward("lig_project_gaussians_forward_cholesky", num_points, cholesky,
                            means2d, opacities, img_height, img_width, block_width)


def project_gaussians_view(num_points, par, means2d, opacities, cholesky, origin, scale,
                           img_height, img_width, block_width):
    """Projection into a window (see lig_project_gaussians_view): origin =
    (x, y) on the source canvas, scale = (sx, sy) output pixels per source
    pixel.  -> (xys, extents, conics, num_tiles_hit) in view coordinates."""
    _check("par", par)
    _check("means2d", means2d)
    _check("opacities", opacities)
    xys = _empty((num_points, 2), par)
    extents = _empty((num_points, 2), par)
    conics = _empty((num_points, 3), par)
    ntiles = _empty((num_points,), par, _i32)
    with torch.cuda.device(par.device):
        L.call("lig_project_gaussians_view", int(num_points), L.ptr(par), L.ptr(means2d),
               L.ptr(opacities), int(bool(cholesky)), float(origin[0]), float(origin[1]),
               float(scale[0]), float(scale[1]), int(img_height), int(img_width),
               int(block_width), L.ptr(xys), L.ptr(extents), L.ptr(conics), L.ptr(ntiles),
               L.stream_of(par))
    return xys, extents, conics, ntiles


def rasterize_image(img_height, img_width, gaussian_ids_sorted, tile_bins, xys, conics, colors,
                    opacities=None):
    """Image-only forward (16 x 16 tiles, no wsum / final_idx): out_img [H,W,3]."""
    for n_, t, dt in (("gaussian_ids_sorted", gaussian_ids_sorted, _i32),
                      ("tile_bins", tile_bins, _i32), ("xys", xys, _f32),
                      ("conics", conics, _f32), ("colors", colors, _f32),
                      ("opacities", opacities, _f32)):
        _check(n_, t, dt)
    colors = _contig(colors)
    n = int(xys.size(0))
    H, W = int(img_height), int(img_width)
    out_img = _empty((H, W, 3), xys)
    packed = _workspace(L.size_query("lig_packed_gaussians_bytes", n), xys)
    ows = _order_workspace(H, W, 16, xys)
    with torch.cuda.device(xys.device):
        L.call("lig_rasterize_forward_ws", n, H, W, 16, L.ptr(gaussian_ids_sorted),
               L.ptr(tile_bins), L.ptr(xys), L.ptr(conics), L.ptr(colors), L.ptr(opacities),
               None, None, L.ptr(packed), L.ptr(out_img), None, None, None, None, None,
               L.ptr(ows), ows.numel(), L.stream_of(xys))
    return out_img


def _contig(t):
    return t.contiguous() if t is not None else None


def _num_tiles(img_height, img_width, block_width):
    bw = int(block_width)
    return ((int(img_height) + bw - 1) // bw) * ((int(img_width) + bw - 1) // bw)


def _order_workspace(img_height, img_width, block_width, like):
    """Workspace in which a 16 x 16 raster launch orders its tiles
    (lig_tile_order_workspace_size; speed only, no result depends on it)."""
    return _workspace(L.size_query("lig_tile_order_workspace_size",
                                   _num_tiles(img_height, img_width, block_width)), like)


def project_gaussians_backward(num_points, extent, conics, v_xy, v_conic):
    """bindings.cu:119-151 -> (v_cov2d [n,3], v_mean2d [n,2])."""
    extent, conics, v_xy, v_conic = map(_contig, (extent, conics, v_xy, v_conic))
    for n_, t in (("extent", extent), ("conics", conics), ("v_xy", v_xy), ("v_conic", v_conic)):
        _check(n_, t)
    v_cov = _empty((num_points, 3), conics)
    v_mean = _empty((num_points, 2), conics)
    with torch.cuda.device(conics.device):
        L.call("lig_project_gaussians_backward", num_points, L.ptr(extent), L.ptr(conics),
               L.ptr(v_xy), L.ptr(v_conic), L.ptr(v_cov), L.ptr(v_mean), L.stream_of(conics))
    return v_cov, v_mean


def project_gaussians_backward_cholesky(num_points, extent, cholesky, conics, v_xy, v_conic):
    """bindings.cu:204-237 -> (v_cholesky [n,3], v_mean2d [n,2])."""
    extent, cholesky, conics, v_xy, v_conic = map(_contig, (extent, cholesky, conics, v_xy,
                                                            v_conic))
    for n_, t in (("extent", extent), ("cholesky", cholesky), ("conics", conics),
                  ("v_xy", v_xy), ("v_conic", v_conic)):
        _check(n_, t)
    v_chol = _empty((num_points, 3), conics)
    v_mean = _empty((num_points, 2), conics)
    with torch.cuda.device(conics.device):
        L.call("lig_project_gaussians_backward_cholesky", num_points, L.ptr(extent),
               L.ptr(cholesky), L.ptr(conics), L.ptr(v_xy), L.ptr(v_conic), L.ptr(v_chol),
               L.ptr(v_mean), L.stream_of(conics))
    return v_chol, v_mean


def map_gaussian_to_intersects(num_points, num_intersects, xys, depths, extent, cum_tiles_hit,
                               tile_bounds, block_width):
    """bindings.cu:239-281 -> (isect_ids int64, gaussian_ids int32), unsorted."""
    _check("xys", xys)
    _check("depths", depths)
    _check("extent", extent)
    _check("cum_tiles_hit", cum_tiles_hit, _i32)
    isect = torch.zeros((num_intersects,), dtype=torch.int64, device=xys.device)
    gids = torch.zeros((num_intersects,), dtype=_i32, device=xys.device)
    with torch.cuda.device(xys.device):
        L.call("lig_map_gaussian_to_intersects", num_points, num_intersects, L.ptr(xys),
               L.ptr(depths), L.ptr(extent), L.ptr(cum_tiles_hit), int(tile_bounds[0]),
               int(tile_bounds[1]), int(block_width), L.ptr(isect), L.ptr(gids),
               L.stream_of(xys))
    return isect, gids


def get_tile_bin_edges(num_intersects, isect_ids_sorted, tile_bounds):
    """bindings.cu:283-301 -> tile_bins [tiles, 2] int32."""
    _check("isect_ids_sorted", isect_ids_sorted, torch.int64)
    nt = int(tile_bounds[0]) * int(tile_bounds[1])
    bins = _empty((nt, 2), isect_ids_sorted, _i32)
    with torch.cuda.device(isect_ids_sorted.device):
        L.call("lig_get_tile_bin_edges", int(num_intersects), L.ptr(isect_ids_sorted),
               int(tile_bounds[0]), int(tile_bounds[1]), L.ptr(bins),
               L.stream_of(isect_ids_sorted))
    return bins


def rasterize_forward(tile_bounds, block, img_size, gaussian_ids_sorted, tile_bins, xys, conics,
                      colors, opacities, compute_upscale_gradients=True):
    """bindings.cu:303-401 -> (out_img, out_wsum, out_dx, out_dy, out_dxy, final_idx);
    out_dx/dy/dxy are None unless compute_upscale_gradients."""
    out = rasterize_forward_packed(int(img_size[1]), int(img_size[0]), int(block[0]),
                                   gaussian_ids_sorted, tile_bins, xys, conics, colors,
                                   opacities, compute_upscale_gradients)
    return out[:6]


def rasterize_backward(img_height, img_width, block_width, gaussians_ids_sorted, tile_bins, xys,
                       conics, colors, final_idx, v_output, v_render_wsum, opacities,
                       v_output_dx=None, v_output_dy=None, v_output_dxy=None):
    """bindings.cu:403-506 -> (v_xy, v_xy_abs, v_conic, v_colors, v_opacity|None)."""
    if xys.ndimension() != 2 or xys.size(1) != 2:
        raise RuntimeError("xys must have dimensions (num_points, 2)")
    if colors.ndimension() != 2 or colors.size(1) != 3:
        raise RuntimeError("colors must have 2 dimensions")
    packed = pack_gaussians(xys, conics, colors, opacities)
    return rasterize_backward_packed(img_height, img_width, block_width, gaussians_ids_sorted,
                                     tile_bins, packed, xys.size(0), final_idx, v_output,
                                     v_render_wsum, opacities is not None, v_output_dx,
                                     v_output_dy, v_output_dxy)


def _roi4(roi):
    return tuple(float(r) for r in roi)


def bilinear_resize(src, full_h, full_w, window=None):
    """src [h, w, 3] resized bilinearly to full_h x full_w with the half-pixel
    grid of torch's interpolate(mode="bilinear", align_corners=False); with
    window = (x0, y0, w, h) only that window of the result is computed
    (lig_bilinear_resize).  -> [h, w, 3] (the window) or [full_h, full_w, 3]."""
    _check("src", src)
    if src.dim() != 3 or src.shape[2] != 3:
        raise ValueError("src must be [H, W, 3]")
    x0, y0, w, h = window if window is not None else (0, 0, full_w, full_h)
    out = _empty((h, w, 3), src)
    with torch.cuda.device(src.device):
        L.call("lig_bilinear_resize", int(src.shape[0]), int(src.shape[1]), int(full_h),
               int(full_w), int(y0), int(x0), int(h), int(w), L.ptr(src), L.ptr(out),
               L.stream_of(src))
    return out


def gradient_aware_upscale_forward(render, dx, dy, dxy, dst_h, dst_w, roi):
    """bindings.cu:508-559 -> output [dst_h, dst_w, 3]."""
    for n_, t in (("render", render), ("dx", dx), ("dy", dy), ("dxy", dxy)):
        _check(n_, t)
    src_h, src_w = int(render.size(0)), int(render.size(1))
    out = _empty((int(dst_h), int(dst_w), 3), render)
    r = _roi4(roi)
    with torch.cuda.device(render.device):
        L.call("lig_gradient_aware_upscale_forward", src_h, src_w, int(dst_h), int(dst_w), *r,
               L.ptr(render), L.ptr(dx), L.ptr(dy), L.ptr(dxy), L.ptr(out),
               L.stream_of(render))
    return out


def gradient_aware_upscale_backward(grad_output, render, dx, dy, dxy, dst_h, dst_w, roi):
    """bindings.cu:561-605 -> (grad_render, grad_dx, grad_dy, grad_dxy)."""
    _check("grad_output", grad_output)
    src_h, src_w = int(render.size(0)), int(render.size(1))
    outs = [_empty((src_h, src_w, 3), grad_output) for _ in range(4)]
    ws = _workspace(L.size_query("lig_upscale_backward_workspace_size", src_h, src_w,
                                 int(dst_h), int(dst_w)), grad_output)
    r = _roi4(roi)
    with torch.cuda.device(grad_output.device):
        L.call("lig_gradient_aware_upscale_backward", src_h, src_w, int(dst_h), int(dst_w), *r,
               L.ptr(grad_output), *[L.ptr(o) for o in outs], L.ptr(ws), ws.numel(),
               L.stream_of(grad_output))
    return tuple(outs)


# ---------------------------------------------------------------------------
# entry points without a reference counterpart
# ---------------------------------------------------------------------------

def inclusive_cumsum_i32(x):
    """Inclusive int32 scan on the device (replaces torch.cumsum, utils.py:65)."""
    _check("num_tiles_hit", x, _i32)
    n = int(x.numel())
    out = _empty((n,), x, _i32)
    ws = _workspace(L.size_query("lig_scan_workspace_size", n), x)
    with torch.cuda.device(x.device):
        L.call("lig_inclusive_scan_i32", n, L.ptr(x), L.ptr(out), L.ptr(ws), ws.numel(),
               L.stream_of(x))
    return out


def sort_isect_pairs(isect_ids, gaussian_ids):
    """Stable sort by signed 64-bit key carrying int32 values (replaces
    torch.sort + torch.gather, utils.py:97-98)."""
    _check("isect_ids", isect_ids, torch.int64)
    _check("gaussian_ids", gaussian_ids, _i32)
    m = int(isect_ids.numel())
    ks = torch.empty_like(isect_ids)
    vs = torch.empty_like(gaussian_ids)
    if m == 0:
        return ks, vs
    ws = _workspace(L.size_query("lig_sort_pairs_workspace_size", m, 64), isect_ids)
    with torch.cuda.device(isect_ids.device):
        L.call("lig_sort_isect_pairs", m, L.ptr(isect_ids), L.ptr(gaussian_ids), L.ptr(ks),
               L.ptr(vs), L.ptr(ws), ws.numel(), L.stream_of(isect_ids))
    return ks, vs


def bin_workspace_bytes(num_points, num_intersects, img_height, img_width, block_width,
                        path="auto"):
    """Workspace of lig_bin_and_sort_gaussians: the row path's when it applies
    (path "auto"), the radix path's (lig_bin_workspace_size) otherwise or when
    path is "radix" -- the library takes the row path whenever the workspace
    holds it, so "radix" only selects the radix path where its workspace is the
    smaller of the two (both give the same outputs)."""
    tx = (img_width + block_width - 1) // block_width
    ty = (img_height + block_width - 1) // block_width
    radix = L.size_query("lig_bin_workspace_size", int(num_intersects), tx * ty)
    if path == "radix":
        return radix
    rows = L.size_query("lig_bin_rows_workspace_size", int(num_points), int(num_intersects),
                        int(img_height), int(img_width), int(block_width))
    return max(radix, rows)


def bin_and_sort_fast(num_points, num_intersects, xys, extents, cum_tiles_hit, img_height,
                      img_width, block_width, with_emission_ids=False, path="auto"):
    """Fused binning for the all-zero-depth path (emission and sort by tile
    rows, or the tile-key radix sort; the same outputs).
    -> (gaussian_ids_sorted [I] int32, tile_bins [T,2] int32), plus the emission
    slot of each sorted intersection [I] int32 when with_emission_ids."""
    _check("xys", xys)
    _check("extents", extents)
    _check("cum_tiles_hit", cum_tiles_hit, _i32)
    tx = (img_width + block_width - 1) // block_width
    ty = (img_height + block_width - 1) // block_width
    gids = _empty((num_intersects,), xys, _i32)
    bins = _empty((tx * ty, 2), xys, _i32)
    eids = _empty((num_intersects,), xys, _i32) if with_emission_ids else None
    ws = _workspace(bin_workspace_bytes(num_points, num_intersects, img_height, img_width,
                                        block_width, path), xys)
    with torch.cuda.device(xys.device):
        L.call("lig_bin_and_sort_gaussians", int(num_points), int(num_intersects), L.ptr(xys),
               L.ptr(extents), L.ptr(cum_tiles_hit), int(img_height), int(img_width),
               int(block_width), L.ptr(gids), L.ptr(bins), L.ptr(eids), L.ptr(ws), ws.numel(),
               L.stream_of(xys))
    return (gids, bins, eids) if with_emission_ids else (gids, bins)


# Fixed scratch of lig_bin_and_sort_gaussians_noalloc (LIG_BIN_NOALLOC_SCRATCH_BYTES).
BIN_NOALLOC_SCRATCH_BYTES = 64 << 20


def bin_and_sort_noalloc(num_points, num_intersects, xys, extents, cum_tiles_hit, img_height,
                         img_width, block_width, gaussian_ids_sorted, tile_bins,
                         emission_ids_sorted=None):
    """bin_and_sort_fast into caller-provided outputs, with the library's fixed
    per-device scratch as workspace (lig_bin_and_sort_gaussians_noalloc): no
    device memory is allocated.  Calls that overlap on one GPU race."""
    for name, t in (("xys", xys), ("extents", extents)):
        _check(name, t)
    for name, t in (("cum_tiles_hit", cum_tiles_hit), ("gaussian_ids_sorted",
                    gaussian_ids_sorted), ("tile_bins", tile_bins),
                    ("emission_ids_sorted", emission_ids_sorted)):
        _check(name, t, _i32)
    with torch.cuda.device(xys.device):
        L.call("lig_bin_and_sort_gaussians_noalloc", int(num_points), int(num_intersects),
               L.ptr(xys), L.ptr(extents), L.ptr(cum_tiles_hit), int(img_height),
               int(img_width), int(block_width), L.ptr(gaussian_ids_sorted), L.ptr(tile_bins),
               L.ptr(emission_ids_sorted), L.stream_of(xys))


def pack_gaussians(xys, conics, colors, opacities):
    """Per-Gaussian 48-byte records read by the rasteriser kernels."""
    for n_, t in (("xys", xys), ("conics", conics), ("colors", colors),
                  ("opacities", opacities)):
        _check(n_, t)
    n = int(xys.size(0))
    packed = _workspace(L.size_query("lig_packed_gaussians_bytes", n), xys)
    with torch.cuda.device(xys.device):
        L.call("lig_pack_gaussians", n, L.ptr(xys), L.ptr(conics), L.ptr(colors),
               L.ptr(opacities), L.ptr(packed), L.stream_of(xys))
    return packed


def rasterize_forward_packed(img_height, img_width, block_width, gaussian_ids_sorted, tile_bins,
                             xys, conics, colors, opacities, compute_upscale_gradients,
                             extents=None, cum_tiles_hit=None, ordered=True, final_index=True):
    """Forward that also returns the packed records for the backward:
    (out_img, out_wsum, out_dx, out_dy, out_dxy, final_idx, packed).  With
    extents and cum_tiles_hit (16 x 16 tiles, no opacities) the pack also
    addresses the backward's records, which then need no emission ids.
    ordered=False launches the tiles in the fixed per-XCD order (same
    results; the launch order only changes speed).  final_index=False (16 x
    16, no upscale gradients) skips the final index (final_idx is None): the
    backward given None re-derives the accepted pairs with the same test."""
    for n_, t, dt in (("gaussian_ids_sorted", gaussian_ids_sorted, _i32),
                      ("tile_bins", tile_bins, _i32), ("xys", xys, _f32),
                      ("conics", conics, _f32), ("colors", colors, _f32),
                      ("opacities", opacities, _f32)):
        _check(n_, t, dt)
    if colors.ndimension() != 2 or colors.size(1) != 3:
        raise RuntimeError("colors must have shape (num_points, 3)")
    n = int(xys.size(0))
    H, W = int(img_height), int(img_width)
    out_img = _empty((H, W, 3), xys)
    out_wsum = _empty((H, W), xys)
    if not final_index and (int(block_width) != 16 or compute_upscale_gradients):
        raise ValueError("final_index=False needs 16 x 16 tiles without upscale gradients")
    final_idx = _empty((H, W), xys, _i32) if final_index else None
    dx = dy = dxy = None
    if compute_upscale_gradients:
        dx, dy, dxy = (_empty((H, W, 3), xys) for _ in range(3))
    packed = _workspace(L.size_query("lig_packed_gaussians_bytes", n), xys)
    ows = _order_workspace(H, W, block_width, xys) if ordered else None
    with torch.cuda.device(xys.device):
        L.call("lig_rasterize_forward_ws", n, H, W, int(block_width),
               L.ptr(gaussian_ids_sorted), L.ptr(tile_bins), L.ptr(xys), L.ptr(conics),
               L.ptr(colors), L.ptr(opacities), L.ptr(extents), L.ptr(cum_tiles_hit),
               L.ptr(packed), L.ptr(out_img), L.ptr(out_wsum), L.ptr(final_idx), L.ptr(dx),
               L.ptr(dy), L.ptr(dxy), L.ptr(ows), ows.numel() if ordered else 0,
               L.stream_of(xys))
    return out_img, out_wsum, dx, dy, dxy, final_idx, packed


def rasterize_backward_packed(img_height, img_width, block_width, gaussian_ids_sorted, tile_bins,
                              packed, num_points, final_idx, v_output, v_render_wsum,
                              has_opacities, v_output_dx=None, v_output_dy=None,
                              v_output_dxy=None, emission_ids_sorted=None, cum_tiles_hit=None,
                              slots_in_pack=False, ordered=True):
    """Backward over packed records; v_render_wsum may be None (zeros).  With
    cum_tiles_hit and either emission_ids_sorted (from bin_and_sort_fast) or
    slots_in_pack (packed by a forward given extents and cum_tiles_hit) the
    per-Gaussian sums are reduced deterministically; otherwise by atomics.
    The deterministic form orders its tiles' launch in spare workspace unless
    ordered=False (same results).  final_idx None (16 x 16, no upscale
    gradients, after a forward with final_index=False): no pixel is gated by
    a final index and each tile's Gaussians are grouped by the rows they
    reach; the results equal those with the forward's final index."""
    v_output, v_render_wsum, v_output_dx, v_output_dy, v_output_dxy = map(
        _contig, (v_output, v_render_wsum, v_output_dx, v_output_dy, v_output_dxy))
    for n_, t, dt in (("gaussian_ids_sorted", gaussian_ids_sorted, _i32),
                      ("tile_bins", tile_bins, _i32), ("final_idx", final_idx, _i32),
                      ("v_output", v_output, _f32), ("v_render_wsum", v_render_wsum, _f32),
                      ("v_output_dx", v_output_dx, _f32), ("v_output_dy", v_output_dy, _f32),
                      ("v_output_dxy", v_output_dxy, _f32),
                      ("emission_ids_sorted", emission_ids_sorted, _i32),
                      ("cum_tiles_hit", cum_tiles_hit, _i32)):
        _check(n_, t, dt)
    ga = v_output_dx is not None
    if ga and (v_output_dy is None or v_output_dxy is None):
        raise RuntimeError("v_output_dx, v_output_dy and v_output_dxy go together")
    det = emission_ids_sorted is not None or slots_in_pack
    if det and cum_tiles_hit is None:
        raise RuntimeError("the deterministic backward needs cum_tiles_hit")
    n = int(num_points)
    like = v_output
    v_xy = _empty((n, 2), like)
    v_xy_abs = _empty((n, 2), like)
    v_conic = _empty((n, 3), like)
    v_colors = _empty((n, 3), like)
    v_opacity = _empty((n,), like) if has_opacities else None
    hop = int(bool(has_opacities))
    common = (n, int(img_height), int(img_width), int(block_width), L.ptr(gaussian_ids_sorted),
              L.ptr(tile_bins), L.ptr(packed), L.ptr(final_idx), L.ptr(v_output),
              L.ptr(v_render_wsum), L.ptr(v_output_dx), L.ptr(v_output_dy),
              L.ptr(v_output_dxy), hop)
    outs = (L.ptr(v_xy), L.ptr(v_xy_abs), L.ptr(v_conic), L.ptr(v_colors), L.ptr(v_opacity))
    with torch.cuda.device(like.device):
        if det:
            m = int(gaussian_ids_sorted.numel())
            nbytes = L.size_query("lig_rasterize_backward_workspace_size", m)
            if ordered:
                nbytes += 256 + L.size_query("lig_tile_order_workspace_size",
                                             _num_tiles(img_height, img_width, block_width))
            ws = _workspace(nbytes, like)
            L.call("lig_rasterize_backward_records", *common, L.ptr(emission_ids_sorted), m,
                   L.ptr(ws), ws.numel(), L.stream_of(like))
            L.call("lig_reduce_gradient_records", n, m, L.ptr(cum_tiles_hit), hop, L.ptr(ws),
                   ws.numel(), *outs, L.stream_of(like))
        else:
            L.call("lig_rasterize_backward", *common, None, None, 0, None, 0, *outs,
                   L.stream_of(like))
    return v_xy, v_xy_abs, v_conic, v_colors, v_opacity


def fit_render_l2(img_height, img_width, gaussian_ids_sorted, tile_bins, xys, extents,
                  cum_tiles_hit, conics, colors, target, batch_depth=0):
    """Fused fit-step render (16 x 16 tiles, no opacities): (out_img [H,W,3],
    loss [] tensor, packed).  Bins from bin_and_sort_fast without emission
    ids.  See lig_fit_render_l2 in include/lig_abi.h."""
    colors, target = _contig(colors), _contig(target)
    for n_, t, dt in (("gaussian_ids_sorted", gaussian_ids_sorted, _i32),
                      ("tile_bins", tile_bins, _i32), ("xys", xys, _f32),
                      ("extents", extents, _f32), ("cum_tiles_hit", cum_tiles_hit, _i32),
                      ("conics", conics, _f32), ("colors", colors, _f32),
                      ("target", target, _f32)):
        _check(n_, t, dt)
    H, W = int(img_height), int(img_width)
    if tuple(target.shape) != (H, W, 3):
        raise ValueError(f"target must be [{H}, {W}, 3], got {tuple(target.shape)}")
    n = int(xys.size(0))
    out_img = _empty((H, W, 3), xys)
    loss = _empty((), xys)
    packed = _workspace(L.size_query("lig_packed_gaussians_bytes", n), xys)
    ws = _workspace(L.size_query("lig_fit_render_workspace_size", H, W), xys)
    with torch.cuda.device(xys.device):
        L.call("lig_fit_render_l2", n, H, W, L.ptr(gaussian_ids_sorted), L.ptr(tile_bins),
               L.ptr(xys), L.ptr(extents), L.ptr(cum_tiles_hit), L.ptr(conics), L.ptr(colors),
               L.ptr(packed), L.ptr(target), L.ptr(out_img), L.ptr(loss), int(batch_depth),
               L.ptr(ws), ws.numel(), L.stream_of(xys))
    return out_img, loss, packed


def fit_backward_l2(img_height, img_width, gaussian_ids_sorted, tile_bins, packed, num_points,
                    num_intersects, img, target, grad_loss, cum_tiles_hit):
    """Backward of fit_render_l2's loss: (v_xy, v_conic, v_colors),
    deterministic (records at emission slots + slot-order reduction)."""
    img, target, grad_loss = _contig(img), _contig(target), _contig(grad_loss)
    for n_, t, dt in (("gaussian_ids_sorted", gaussian_ids_sorted, _i32),
                      ("tile_bins", tile_bins, _i32), ("img", img, _f32),
                      ("target", target, _f32), ("grad_loss", grad_loss, _f32),
                      ("cum_tiles_hit", cum_tiles_hit, _i32)):
        _check(n_, t, dt)
    n, m = int(num_points), int(num_intersects)
    like = img
    v_xy = _empty((n, 2), like)
    v_conic = _empty((n, 3), like)
    v_colors = _empty((n, 3), like)
    ws = _workspace(L.size_query("lig_rasterize_backward_workspace_size", m), like)
    with torch.cuda.device(like.device):
        L.call("lig_fit_backward_records_l2", n, int(img_height), int(img_width),
               L.ptr(gaussian_ids_sorted), L.ptr(tile_bins), L.ptr(packed), L.ptr(img),
               L.ptr(target), L.ptr(grad_loss), m, L.ptr(ws), ws.numel(), L.stream_of(like))
        L.call("lig_fit_reduce_records", n, m, L.ptr(cum_tiles_hit), L.ptr(ws), ws.numel(),
               L.ptr(v_xy), L.ptr(v_conic), L.ptr(v_colors), L.stream_of(like))
    return v_xy, v_conic, v_colors


def count_pairs(xys, extents, conics, opacities, num_tiles_hit, img_height, img_width):
    """(E, C) workload counters (see lig_count_pairs in include/lig_abi.h)."""
    for n_, t in (("xys", xys), ("extents", extents), ("conics", conics),
                  ("opacities", opacities)):
        _check(n_, t)
    _check("num_tiles_hit", num_tiles_hit, _i32)
    out = torch.empty(2, dtype=torch.int64, device=xys.device)
    with torch.cuda.device(xys.device):
        L.call("lig_count_pairs", int(xys.size(0)), int(img_height), int(img_width), L.ptr(xys),
               L.ptr(extents), L.ptr(conics), L.ptr(opacities), L.ptr(num_tiles_hit),
               L.ptr(out), L.stream_of(xys))
    e, c = out.tolist()
    return int(e), int(c)
