"""ctypes binding of ``liblig_hip.so`` (the C-ABI declared in include/lig_abi.h).

The library is loaded after ``torch`` so that it shares torch's HIP runtime
(both name ``libamdhip64.so.7``).  There is no fallback: a missing library or a
tensor that is not on a HIP device raises.
"""
from __future__ import annotations

import ctypes
import os

import torch

_PKG_DIR = os.path.dirname(os.path.abspath(__file__))
# LIG_LIB_VARIANT=checked loads the bounds-checked build (liblig_hip_checked.so,
# lig_check_status reports its failed checks); anything else the normal one
VARIANT = os.environ.get("LIG_LIB_VARIANT", "")
LIB_PATH = os.path.join(_PKG_DIR, "liblig_hip_checked.so" if VARIANT == "checked"
                        else "liblig_hip.so")
CHECKED_LIB_PATH = os.path.join(_PKG_DIR, "liblig_hip_checked.so")

_c_int = ctypes.c_int
_c_float = ctypes.c_float
_c_size = ctypes.c_size_t
_vp = ctypes.c_void_p

# name -> (restype, argtypes)
_SIGNATURES = {
    "lig_strerror": (ctypes.c_char_p, [_c_int]),
    "lig_abi_version": (_c_int, []),
    "lig_compute_cov2d_bounds": (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp]),
    "lig_project_gaussians_forward": (
        _c_int, [_c_int, _vp, _vp, _vp, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp]),
    "lig_project_gaussians_forward_cholesky": (
        _c_int, [_c_int, _vp, _vp, _vp, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp]),
    "lig_project_gaussians_backward": (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "lig_project_gaussians_backward_cholesky": (
        _c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "lig_project_gaussians_view": (
        _c_int, [_c_int, _vp, _vp, _vp, _c_int, _c_float, _c_float, _c_float, _c_float, _c_int,
                 _c_int, _c_int, _vp, _vp, _vp, _vp, _vp]),
    "lig_scan_workspace_size": (_c_size, [_c_int]),
    "lig_inclusive_scan_i32": (_c_int, [_c_int, _vp, _vp, _vp, _c_size, _vp]),
    "lig_map_gaussian_to_intersects": (
        _c_int, [_c_int, _c_int, _vp, _vp, _vp, _vp, _c_int, _c_int, _c_int, _vp, _vp, _vp]),
    "lig_sort_pairs_workspace_size": (_c_size, [_c_int, _c_int]),
    "lig_sort_isect_pairs": (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _c_size, _vp]),
    "lig_get_tile_bin_edges": (_c_int, [_c_int, _vp, _c_int, _c_int, _vp, _vp]),
    "lig_bin_workspace_size": (_c_size, [_c_int, _c_int]),
    "lig_bin_and_sort_gaussians": (
        _c_int, [_c_int, _c_int, _vp, _vp, _vp, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp,
                 _c_size, _vp]),
    "lig_bin_rows_workspace_size": (_c_size, [_c_int, _c_int, _c_int, _c_int, _c_int]),
    "lig_bin_and_sort_gaussians_noalloc": (
        _c_int, [_c_int, _c_int, _vp, _vp, _vp, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp]),
    "lig_packed_gaussians_bytes": (_c_size, [_c_int]),
    "lig_pack_gaussians": (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    "lig_rasterize_forward": (
        _c_int, [_c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                 _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "lig_rasterize_forward_ws": (
        _c_int, [_c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                 _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_size, _vp]),
    "lig_tile_order_workspace_size": (_c_size, [_c_int]),
    "lig_rasterize_backward_workspace_size": (_c_size, [_c_int]),
    "lig_rasterize_backward": (
        _c_int, [_c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                 _c_int, _vp, _vp, _c_int, _vp, _c_size, _vp, _vp, _vp, _vp, _vp, _vp]),
    "lig_l2_loss_workspace_size": (_c_size, [ctypes.c_longlong]),
    "lig_l2_clamped_loss_forward": (
        _c_int, [ctypes.c_longlong, _vp, _vp, _vp, _vp, _c_size, _vp]),
    "lig_l2_clamped_loss_backward": (_c_int, [ctypes.c_longlong, _vp, _vp, _vp, _vp, _vp]),
    "lig_rasterize_backward_records": (
        _c_int, [_c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                 _c_int, _vp, _c_int, _vp, _c_size, _vp]),
    "lig_reduce_gradient_records": (
        _c_int, [_c_int, _c_int, _vp, _c_int, _vp, _c_size, _vp, _vp, _vp, _vp, _vp, _vp]),
    "lig_fit_render_workspace_size": (_c_size, [_c_int, _c_int]),
    "lig_fit_render_l2": (
        _c_int, [_c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                 _c_int, _vp, _c_size, _vp]),
    "lig_fit_backward_records_l2": (
        _c_int, [_c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _c_int, _vp, _c_size,
                 _vp]),
    "lig_fit_reduce_records": (
        _c_int, [_c_int, _c_int, _vp, _vp, _c_size, _vp, _vp, _vp, _vp]),
    "lig_set_backward_shares": (_c_int, [_c_int]),
    "lig_adam_step": (
        _c_int, [ctypes.c_longlong, _vp, _vp, _vp, _vp, _c_float, _c_float, _c_float, _c_float,
                 _c_float, _c_int, _vp]),
    "lig_adan_step": (
        _c_int, [ctypes.c_longlong, _vp, _vp, _vp, _vp, _vp, _vp, _c_float, _c_float, _c_float,
                 _c_float, _c_float, _c_float, _c_int, _c_float, _c_int, _vp]),
    "lig_project_batched_cholesky": (
        _c_int, [_c_int, _vp, _vp, _vp, _c_int, _vp, _vp, _vp, _vp, _vp]),
    "lig_bin_batched": (
        _c_int, [_c_int, _c_int, _vp, _vp, _vp, _vp, _c_int, _c_int, _vp, _vp, _vp, _c_size,
                 _vp]),
    "lig_fit_batched_workspace_size": (_c_size, [_c_int]),
    "lig_fit_render_l2_batched": (
        _c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _vp,
                 _vp, _vp, _c_int, _vp, _c_size, _vp]),
    "lig_fit_backward_records_l2_batched": (
        _c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _c_int, _vp,
                 _c_size, _vp]),
    "lig_count_pairs": (_c_int, [_c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "lig_project_views_batched": (
        _c_int, [_c_int, _vp, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    "lig_render_batched": (
        _c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _vp, _c_int,
                 _vp, _c_size, _vp]),
    "lig_compose_crops": (_c_int, [_c_int, _vp, _c_int, _vp]),
    "lig_check_status": (_c_int, [_vp, _vp, _vp]),
    "lig_bilinear_resize": (
        _c_int, [_c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp]),
    "lig_gradient_aware_upscale_forward": (
        _c_int, [_c_int, _c_int, _c_int, _c_int, _c_float, _c_float, _c_float, _c_float, _vp,
                 _vp, _vp, _vp, _vp, _vp]),
    "lig_upscale_backward_workspace_size": (_c_size, [_c_int, _c_int, _c_int, _c_int]),
    "lig_gradient_aware_upscale_backward": (
        _c_int, [_c_int, _c_int, _c_int, _c_int, _c_float, _c_float, _c_float, _c_float, _vp,
                 _vp, _vp, _vp, _vp, _vp, _c_size, _vp]),
    "lig_fit_render_l2_batched_persistent": (
        _c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _vp,
                 _vp, _vp, _c_int, _vp, _c_size, _vp]),
    "lig_fit_backward_records_l2_batched_persistent": (
        _c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _c_int, _vp,
                 _c_size, _vp]),
    "lig_bin_batched_capacity": (
        _c_int, [_c_int, _c_int, _vp, _vp, _vp, _vp, _c_int, _c_int, _vp, _vp, _vp, _c_size,
                 _vp]),
    "lig_fit_step_begin": (_c_int, [_vp, _c_int, _c_int, _vp, _vp]),
    "lig_adam_step_device": (
        _c_int, [ctypes.c_longlong, _vp, _vp, _vp, _vp, _c_float, _c_float, _c_float, _c_float,
                 _c_float, _vp, _vp, _vp]),
    "lig_adam_bias_corrections": (_c_int, [_c_float, _c_float, _c_int, _c_int, _vp]),
    "lig_ssim_workspace_size": (_c_size, [_c_int, _c_int, _c_int, _c_int, _c_int, _c_int]),
    "lig_ssim_multiscale": (
        _c_int, [_c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _c_float, _c_int, _vp,
                 _c_float, _c_float, _c_int, _vp, _vp, _vp, _vp, _vp, _c_size, _vp]),
}

EXPORTED = tuple(_SIGNATURES)

_lib = None


class LigError(RuntimeError):
    """A C-ABI call returned a non-zero status."""


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m lig_amd.build` "
                "(there is no CPU fallback)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


# Optional per-entry-point timing: name -> callable(name, phase) invoked with
# phase "begin" / "end" around the call on the calling thread (see
# lig_amd.profiling).  Empty in normal use.
hooks = {}


def call(name: str, *args) -> None:
    """Call an int-returning entry point and raise LigError on failure."""
    hook = hooks.get(name)
    if hook is not None:
        hook(name, "begin")
    rc = getattr(load(), name)(*args)
    if hook is not None:
        hook(name, "end")
    if rc != 0:
        msg = load().lig_strerror(rc).decode()
        raise LigError(f"{name} failed: {msg} (status {rc})")


def size_query(name: str, *args) -> int:
    return int(getattr(load(), name)(*args))


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def stream_of(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def check_device(name: str, t, dtype=None):
    """Reference's CHECK_INPUT (bindings.h:10-15): a HIP tensor, contiguous."""
    if t is None:
        return
    if not t.is_cuda:
        raise RuntimeError(f"{name} must be a CUDA tensor")
    if not t.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous")
    if dtype is not None and t.dtype != dtype:
        raise RuntimeError(f"{name} must have dtype {dtype}, got {t.dtype}")


def check_status():
    """(failures, first line, first file) of the bounds-checked build's failed
    checks since the last call (then cleared); None for a normal build."""
    lib = load()
    failures = ctypes.c_ulonglong(0)
    line = ctypes.c_int(0)
    path = ctypes.c_char_p(None)
    rc = lib.lig_check_status(ctypes.byref(failures), ctypes.byref(line), ctypes.byref(path))
    if rc == -4:   # LIG_ERR_UNSUPPORTED: built without checks
        return None
    if rc != 0:
        raise LigError(f"lig_check_status: {lib.lig_strerror(rc).decode()}")
    return (int(failures.value), int(line.value),
            path.value.decode() if path.value is not None else None)
