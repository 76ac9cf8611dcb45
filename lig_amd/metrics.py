"""SSIM and MS-SSIM on the HIP path (lig_ssim_multiscale).

The reference scores a fit with ``ms_ssim(out, gt, data_range=1,
size_average=True)`` (train.py:200) and lists single-scale ``ssim`` among its
losses (utils.py:26-34); both come from the package pytorch-msssim
(requirements.txt:5), which is not part of this build.  `ms_ssim` and `ssim`
here take that package's arguments and give its values: Gaussian window
(`win_size` taps, `win_sigma`) applied with no padding, C1 = (K1 R)^2,
C2 = (K2 R)^2, per-plane means, 2 x 2 zero-padded average pools between
scales, relu on the factors, `weights` per scale, and the average over images
and channels (`size_average`) or over channels only.

Inputs are [N, C, H, W] device tensors; any strides are accepted (an NHWC
image viewed as NCHW is read in place).  Values are computed in FP32 with FP64
sums; they are metrics only: no gradient flows (the SSIM losses of
utils.py:26-34 are not on the fit path), and inputs that require a gradient
while autograd is recording raise.  Planes smaller than the window raise
instead of skipping the filter along that axis.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import torch
from torch import nn

from . import _lib as L

MS_WEIGHTS = (0.0448, 0.2856, 0.3001, 0.2363, 0.1333)


def gauss_window(win_size: int = 11, win_sigma: float = 1.5) -> torch.Tensor:
    """Normalised 1-D Gaussian taps as the package builds them
    (_fspecial_gauss_1d): coordinates k - win_size // 2, FP32."""
    k = torch.arange(win_size, dtype=torch.float32) - win_size // 2
    g = torch.exp(-(k ** 2) / (2 * win_sigma ** 2))
    return g / g.sum()


def _check(X: torch.Tensor, Y: torch.Tensor):
    if X.shape != Y.shape:
        raise ValueError(f"Input images should have the same dimensions, but got {X.shape} "
                         f"and {Y.shape}.")
    if X.dim() != 4:
        raise ValueError(f"Input images should be 4-d tensors [N, C, H, W], got {X.shape}")
    if X.dtype != Y.dtype:
        raise ValueError(f"Input images should have the same dtype, got {X.dtype} and {Y.dtype}.")
    if torch.is_grad_enabled() and (X.requires_grad or Y.requires_grad):
        raise NotImplementedError("ssim / ms_ssim are evaluation metrics here: no backward")


def _taps(win: Optional[torch.Tensor], win_size: int, win_sigma: float) -> torch.Tensor:
    if win is not None:
        win_size = win.shape[-1]
    if win_size % 2 != 1:
        raise ValueError("Window size should be odd.")
    if win is None:
        return gauss_window(win_size, win_sigma)
    return win.detach().reshape(-1, win_size)[0].to("cpu", torch.float32)


def _run(X, Y, data_range, taps, K, weights):
    for name, t in (("X", X), ("Y", Y)):
        if not t.is_cuda:
            raise RuntimeError(f"{name} must be a CUDA tensor")
    X = X.float() if X.dtype != torch.float32 else X
    Y = Y.float() if Y.dtype != torch.float32 else Y
    n, c, h, w = X.shape
    levels = len(weights)
    ws_bytes = L.size_query("lig_ssim_workspace_size", n, c, h, w, len(taps), levels)
    if ws_bytes == 0:
        raise ValueError(f"images {tuple(X.shape)} do not fit a {len(taps)}-tap window at "
                         f"{levels} scale(s) (window 3..11 taps, odd; every scale at least "
                         f"the window on both sides; n*c <= 65535)")
    dev = X.device
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    planes = n * c
    value = torch.empty(planes, dtype=torch.float32, device=dev)
    per_ssim = torch.empty(levels, planes, dtype=torch.float32, device=dev)
    LL = ctypes.c_longlong * 4
    xs, ys = LL(*X.stride()), LL(*Y.stride())
    tap_arr = (ctypes.c_float * len(taps))(*[float(v) for v in taps])
    w_arr = (ctypes.c_float * levels)(*[float(v) for v in weights])
    with torch.cuda.device(dev):
        L.call("lig_ssim_multiscale", n, c, h, w, L.ptr(X), xs, L.ptr(Y), ys, float(data_range),
               len(taps), tap_arr, float(K[0]), float(K[1]), levels, w_arr, L.ptr(value),
               L.ptr(per_ssim), None, L.ptr(ws), ws_bytes, L.stream_of(X))
    return value.view(n, c), per_ssim.view(levels, n, c)


def ms_ssim(X: torch.Tensor, Y: torch.Tensor, data_range: float = 255, size_average: bool = True,
            win_size: int = 11, win_sigma: float = 1.5, win: Optional[torch.Tensor] = None,
            weights: Optional[Sequence[float]] = None, K=(0.01, 0.03)) -> torch.Tensor:
    """Multi-scale SSIM (pytorch_msssim.ms_ssim arguments and values)."""
    _check(X, Y)
    taps = _taps(win, win_size, win_sigma)
    win_size = len(taps)
    smaller_side = min(X.shape[-2:])
    limit = (win_size - 1) * (2 ** 4)
    if not smaller_side > limit:
        raise AssertionError(
            f"Image size should be larger than {limit} due to the 4 downsamplings in ms-ssim")
    weights = MS_WEIGHTS if weights is None else [float(v) for v in
                                                  torch.as_tensor(weights).reshape(-1)]
    value, _ = _run(X, Y, data_range, taps, K, weights)
    return value.mean() if size_average else value.mean(1)


def ssim(X: torch.Tensor, Y: torch.Tensor, data_range: float = 255, size_average: bool = True,
         win_size: int = 11, win_sigma: float = 1.5, win: Optional[torch.Tensor] = None,
         K=(0.01, 0.03), nonnegative_ssim: bool = False) -> torch.Tensor:
    """Single-scale SSIM (pytorch_msssim.ssim arguments and values)."""
    _check(X, Y)
    taps = _taps(win, win_size, win_sigma)
    _, per = _run(X, Y, data_range, taps, K, (1.0,))
    v = per[0]
    if nonnegative_ssim:
        v = torch.relu(v)
    return v.mean() if size_average else v.mean(1)


class SSIM(nn.Module):
    """Module form of `ssim` (pytorch_msssim.SSIM)."""

    def __init__(self, data_range=255, size_average=True, win_size=11, win_sigma=1.5,
                 channel=3, spatial_dims=2, K=(0.01, 0.03), nonnegative_ssim=False):
        super().__init__()
        if spatial_dims != 2:
            raise NotImplementedError("only 2-d images")
        self.win_size, self.win_sigma = win_size, win_sigma
        self.data_range, self.size_average, self.K = data_range, size_average, K
        self.nonnegative_ssim = nonnegative_ssim

    def forward(self, X, Y):
        return ssim(X, Y, self.data_range, self.size_average, self.win_size, self.win_sigma,
                    K=self.K, nonnegative_ssim=self.nonnegative_ssim)


class MS_SSIM(nn.Module):
    """Module form of `ms_ssim` (pytorch_msssim.MS_SSIM)."""

    def __init__(self, data_range=255, size_average=True, win_size=11, win_sigma=1.5,
                 channel=3, spatial_dims=2, weights=None, K=(0.01, 0.03)):
        super().__init__()
        if spatial_dims != 2:
            raise NotImplementedError("only 2-d images")
        self.win_size, self.win_sigma = win_size, win_sigma
        self.data_range, self.size_average = data_range, size_average
        self.weights, self.K = weights, K

    def forward(self, X, Y):
        return ms_ssim(X, Y, self.data_range, self.size_average, self.win_size, self.win_sigma,
                       weights=self.weights, K=self.K)
