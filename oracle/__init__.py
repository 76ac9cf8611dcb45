"""CPU oracle for the gsplat2d hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this package, and only as the checker.  The product
(``lig_amd``) never imports it and has no CPU fallback.

The numerics live in ``lig_oracle.c`` (built by ``oracle/Makefile`` into
``liblig_oracle.so``); this module wraps it with numpy arrays.  Each wrapper
names the reference file:line it restates (paths relative to the reference
repository).

Parity pins (what checks each function; see tests/test_oracle_*.py):
  * projection: closed forms (conic = inverse covariance, extent =
    ceil(sqrt(t * Sigma_ii))), special cases (det == 0, det < 0, t <= 0,
    off-screen), and an independent numpy FP32 restatement;
  * binning: brute-force enumeration of (tile, gid) pairs on tiny scenes,
    permutation / sortedness / bin-size invariants;
  * rasteriser forward: single-Gaussian analytic scenes and a tile-free
    brute-force sum over every Gaussian for every pixel;
  * rasteriser backward (plain and gradient-aware): FP64 central finite
    differences of the forward, with the pair set held fixed;
  * upscaler: reproduction of bicubic polynomials from exact derivatives,
    corner values at integer-aligned samples, and the adjoint identity for
    the backward.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_DIR = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_DIR, "liblig_oracle.so")
_lib = None

NG = 11  # gradient slots per Gaussian: rgb(3) conic(3) xy(2) xy_abs(2) opacity(1)


def build() -> str:
    """Compile the C oracle with its Makefile (gcc, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _DIR], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.orc_num_threads.restype = ctypes.c_int
    return _lib


def _p(a):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.c_void_p)


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


def _i32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


def set_num_threads(n: int) -> None:
    lib().orc_set_num_threads(ctypes.c_int(int(n)))


def num_threads() -> int:
    return int(lib().orc_num_threads())


# --------------------------------------------------------------------------
# projection (helpers.cuh:47-93, forward.cu:13-105, backward.cu:185-247,
# bindings.cu:19-66)
# --------------------------------------------------------------------------

def compute_cov2d_bounds(cov2d, opacities=None):
    cov2d = _f32(cov2d)
    n = cov2d.shape[0]
    conics = np.zeros((n, 3), np.float32)
    extents = np.zeros((n, 2), np.float32)
    lib().orc_compute_cov2d_bounds(ctypes.c_int(n), _p(cov2d), _p(_f32(opacities)),
                                   _p(conics), _p(extents))
    return conics, extents


def project_gaussians_forward(cov2d, means2d, img_height, img_width, block_width,
                              opacities=None, cholesky=False):
    cov2d = _f32(cov2d)
    means2d = _f32(means2d)
    opac = _f32(opacities)
    n = cov2d.shape[0]
    xys = np.zeros((n, 2), np.float32)
    extents = np.zeros((n, 2), np.float32)
    conics = np.zeros((n, 3), np.float32)
    ntiles = np.zeros((n,), np.int32)
    fn = lib().orc_project_fwd_chol if cholesky else lib().orc_project_fwd
    fn(ctypes.c_int(n), _p(cov2d), _p(means2d), _p(opac), ctypes.c_int(img_height),
       ctypes.c_int(img_width), ctypes.c_int(block_width), _p(xys), _p(extents),
       _p(conics), _p(ntiles))
    return xys, extents, conics, ntiles


def project_gaussians_backward(extents, conics, v_xy, v_conic, cholesky=None):
    """``cholesky`` given -> the Cholesky chain (backward.cu:206-247)."""
    extents, conics, v_xy, v_conic = map(_f32, (extents, conics, v_xy, v_conic))
    n = extents.shape[0]
    v_par = np.zeros((n, 3), np.float32)
    v_mean = np.zeros((n, 2), np.float32)
    if cholesky is None:
        lib().orc_project_bwd(ctypes.c_int(n), _p(extents), _p(conics), _p(v_xy),
                              _p(v_conic), _p(v_par), _p(v_mean))
    else:
        chol = _f32(cholesky)
        lib().orc_project_bwd_chol(ctypes.c_int(n), _p(extents), _p(chol), _p(conics),
                                   _p(v_xy), _p(v_conic), _p(v_par), _p(v_mean))
    return v_par, v_mean


# --------------------------------------------------------------------------
# binning (utils.py:61-100, forward.cu:109-170)
# --------------------------------------------------------------------------

def cumsum_i32(x):
    x = _i32(x)
    out = np.zeros_like(x)
    lib().orc_cumsum_i32(ctypes.c_int(x.shape[0]), _p(x), _p(out))
    return out


def map_gaussian_to_intersects(num_intersects, xys, depths, extents, cum_tiles_hit,
                               tile_bounds, block_width):
    xys, depths, extents = map(_f32, (xys, depths, extents))
    cum = _i32(cum_tiles_hit)
    n = xys.shape[0]
    isect = np.zeros((num_intersects,), np.int64)
    gids = np.zeros((num_intersects,), np.int32)
    lib().orc_map_intersects(ctypes.c_int(n), _p(xys), _p(depths), _p(extents), _p(cum),
                             ctypes.c_int(tile_bounds[0]), ctypes.c_int(tile_bounds[1]),
                             ctypes.c_int(block_width), _p(isect), _p(gids))
    return isect, gids


def sort_pairs(keys, vals):
    keys = np.ascontiguousarray(keys, dtype=np.int64)
    vals = _i32(vals)
    ko = np.zeros_like(keys)
    vo = np.zeros_like(vals)
    lib().orc_sort_pairs(ctypes.c_int(keys.shape[0]), _p(keys), _p(vals), _p(ko), _p(vo))
    return ko, vo


def tile_bin_edges(isect_sorted, num_tiles):
    isect_sorted = np.ascontiguousarray(isect_sorted, dtype=np.int64)
    bins = np.zeros((num_tiles, 2), np.int32)
    lib().orc_tile_bin_edges(ctypes.c_int(isect_sorted.shape[0]), _p(isect_sorted), _p(bins))
    return bins


def bin_and_sort(xys, extents, num_tiles_hit, img_height, img_width, block_width,
                 depths=None):
    """utils.py:61-100 end to end: returns (num_intersects, cum, isect_ids,
    gids, isect_sorted, gids_sorted, tile_bins)."""
    n = xys.shape[0]
    tx = (img_width + block_width - 1) // block_width
    ty = (img_height + block_width - 1) // block_width
    cum = cumsum_i32(num_tiles_hit)
    m = int(cum[-1]) if n else 0
    if depths is None:
        depths = np.zeros((n,), np.float32)
    isect, gids = map_gaussian_to_intersects(m, xys, depths, extents, cum, (tx, ty, 1),
                                             block_width)
    ks, gs = sort_pairs(isect, gids)
    bins = tile_bin_edges(ks, tx * ty)
    return m, cum, isect, gids, ks, gs, bins


# --------------------------------------------------------------------------
# rasteriser (forward.cu:173-279, backward.cu:8-183,
# gradient_aware_rasterizer.cu:118-495)
# --------------------------------------------------------------------------

def rasterize_forward(gids_sorted, tile_bins, xys, conics, colors, opacities, img_height,
                      img_width, block_width, gradient_aware=False):
    """FP64 forward.  Returns dict with img, wsum, (dx, dy, dxy), final_idx and
    ambig (ambiguous pairs per pixel)."""
    gids_sorted = _i32(gids_sorted)
    tile_bins = _i32(tile_bins)
    xys, conics, colors, opac = map(_f32, (xys, conics, colors, opacities))
    H, W = img_height, img_width
    img = np.zeros((H, W, 3), np.float64)
    wsum = np.zeros((H, W), np.float64)
    fin = np.zeros((H, W), np.int32)
    amb = np.zeros((H, W), np.int32)
    dx = dy = dxy = None
    if gradient_aware:
        dx = np.zeros((H, W, 3), np.float64)
        dy = np.zeros((H, W, 3), np.float64)
        dxy = np.zeros((H, W, 3), np.float64)
    lib().orc_rasterize_fwd(ctypes.c_int(H), ctypes.c_int(W), ctypes.c_int(block_width),
                            _p(gids_sorted), _p(tile_bins), _p(xys), _p(conics), _p(colors),
                            _p(opac), _p(img), _p(wsum), _p(dx), _p(dy), _p(dxy), _p(fin),
                            _p(amb))
    out = {"img": img, "wsum": wsum, "final_idx": fin, "ambig": amb}
    if gradient_aware:
        out.update(dx=dx, dy=dy, dxy=dxy)
    return out


def rasterize_backward(gids_sorted, tile_bins, xys, conics, colors, opacities, final_idx,
                       v_out, v_wsum, img_height, img_width, block_width, v_dx=None,
                       v_dy=None, v_dxy=None, mode="fp64"):
    """Per-Gaussian gradients.  mode "fp64": exact math; "ref32": every pair in
    FP32 with the reference's expression order (the reference gradient).
    Returns dict v_rgb, v_conic, v_xy, v_xy_abs, v_opacity (FP64) and ambig
    (ambiguous pairs per Gaussian, fp64 mode only)."""
    gids_sorted = _i32(gids_sorted)
    tile_bins = _i32(tile_bins)
    xys, conics, colors, opac = map(_f32, (xys, conics, colors, opacities))
    v_out, v_wsum, v_dx, v_dy, v_dxy = map(_f32, (v_out, v_wsum, v_dx, v_dy, v_dxy))
    final_idx = _i32(final_idx)
    n = xys.shape[0]
    grads = np.zeros((n, NG), np.float64)
    amb = np.zeros((n,), np.int32)
    lib().orc_rasterize_bwd(ctypes.c_int(n), ctypes.c_int(img_height), ctypes.c_int(img_width),
                            ctypes.c_int(block_width), _p(gids_sorted), _p(tile_bins), _p(xys),
                            _p(conics), _p(colors), _p(opac), _p(final_idx), _p(v_out),
                            _p(v_wsum), _p(v_dx), _p(v_dy), _p(v_dxy),
                            ctypes.c_int(0 if mode == "fp64" else 1), _p(grads), _p(amb))
    return {"v_rgb": grads[:, 0:3], "v_conic": grads[:, 3:6], "v_xy": grads[:, 6:8],
            "v_xy_abs": grads[:, 8:10], "v_opacity": grads[:, 10], "ambig": amb}


# --------------------------------------------------------------------------
# upscaler (gradient_aware_upscaler.cu:137-371, upscale.py:50-82)
# --------------------------------------------------------------------------

def _roi(roi, src_h, src_w):
    if roi is None:
        roi = (0.0, 0.0, float(src_w), float(src_h))
    return np.asarray(roi, dtype=np.float32)


def upscale_forward(f, fx, fy, fxy, dst_h, dst_w, roi=None):
    f, fx, fy, fxy = map(_f32, (f, fx, fy, fxy))
    src_h, src_w = f.shape[0], f.shape[1]
    r = _roi(roi, src_h, src_w)
    out = np.zeros((dst_h, dst_w, 3), np.float64)
    lib().orc_upscale_fwd(ctypes.c_int(src_h), ctypes.c_int(src_w), ctypes.c_int(dst_h),
                          ctypes.c_int(dst_w), _p(r), _p(f), _p(fx), _p(fy), _p(fxy), _p(out))
    return out


def upscale_backward(grad_out, src_h, src_w, roi=None):
    g = _f32(grad_out)
    dst_h, dst_w = g.shape[0], g.shape[1]
    r = _roi(roi, src_h, src_w)
    outs = [np.zeros((src_h, src_w, 3), np.float64) for _ in range(4)]
    lib().orc_upscale_bwd(ctypes.c_int(src_h), ctypes.c_int(src_w), ctypes.c_int(dst_h),
                          ctypes.c_int(dst_w), _p(r), _p(g), *[_p(o) for o in outs])
    return tuple(outs)


# --------------------------------------------------------------------------
# roofline workload counters
# --------------------------------------------------------------------------

def count_pairs(xys, extents, conics, opacities, num_tiles_hit, img_height, img_width):
    """(E, C): pairs inside each Gaussian's clipped SnugBox pixel box, and the
    subset with alpha >= 1/255.  Independent of the tile size."""
    xys, extents, conics, opac = map(_f32, (xys, extents, conics, opacities))
    nt = _i32(num_tiles_hit)
    E = ctypes.c_int64(0)
    C = ctypes.c_int64(0)
    lib().orc_count_pairs(ctypes.c_int(xys.shape[0]), ctypes.c_int(img_height),
                          ctypes.c_int(img_width), _p(xys), _p(extents), _p(conics), _p(opac),
                          _p(nt), ctypes.byref(E), ctypes.byref(C))
    return int(E.value), int(C.value)
