"""Crop-native rendering of fitted levels and of two-level LIG representations.

`render_view` renders the window [x0, x0 + w/s) x [y0, y0 + h/s) of a level's
canvas at s output pixels per canvas pixel into an h x w image, projecting each
Gaussian straight into the view (lig_project_gaussians_view: the reference's
scale_to, minimal_trainer_upscale.py:139-159, plus an origin) instead of
shifting and rescaling the parameters, then binning and rasterising only the
window (image-only forward: no wsum, no final index).

`TwoLevel` holds a fitted two-level representation (train.py:92-158): level 1
on an H/f x W/f canvas, level 2 on H x W fitted to the normalised residual, and
the normalisation bounds.  `render_crop` composes a window of the full image as
train.py:150-154 does for the whole image: the level-1 render of the window's
source region with its gradients, upscaled by the gradient-aware upscaler's ROI
path (upscale.py, gradient_aware_upscaler.cu:201-206), plus the level-2 window.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import torch

from . import gsplat2d as g
from .fit import compose
from .gsplat2d import cuda as _C


@dataclass
class LevelParams:
    """Cholesky-parameterised Gaussians of one level on an H x W canvas."""
    means: torch.Tensor
    cholesky: torch.Tensor
    rgbs: torch.Tensor
    H: int
    W: int

    @classmethod
    def of(cls, level) -> "LevelParams":
        return cls(level.means.detach(), level.cholesky.detach(), level.rgbs.detach(), level.H,
                   level.W)


def _project(p: LevelParams, x0: float, y0: float, scale: float, h: int, w: int):
    n = p.means.shape[0]
    return _C.project_gaussians_view(n, p.cholesky, p.means, None, True, (x0, y0),
                                     (scale, scale), h, w, 16)


@torch.no_grad()
def render_view(p: LevelParams, x0: float, y0: float, w: int, h: int,
                scale: float = 1.0) -> torch.Tensor:
    """h x w x 3 raw sum image of the window (x0, y0) at `scale`."""
    xys, ext, con, nt = _project(p, x0, y0, scale, h, w)
    num, cum = g.compute_cumulative_intersects(nt)
    if num < 1:
        return torch.ones(h, w, 3, device=p.means.device)   # rasterize.py:80-90
    gids, bins = _C.bin_and_sort_fast(p.means.shape[0], num, xys, ext, cum, h, w, 16)
    return _C.rasterize_image(h, w, gids, bins, xys, con, p.rgbs)


@torch.no_grad()
def render_view_ga(p: LevelParams, x0: float, y0: float, w: int, h: int):
    """(img, dx, dy, dxy) of the window at scale 1 (gradient-aware forward)."""
    xys, ext, con, nt = _project(p, x0, y0, 1.0, h, w)
    out = g.rasterize_gaussians(xys, ext, con, nt, p.rgbs, None, h, w, 16,
                                compute_upscale_gradients=True)
    return out[0], out[2], out[3], out[4]


@dataclass
class TwoLevel:
    l1: LevelParams
    l2: LevelParams
    lo: float
    hi: float

    @property
    def factor(self) -> int:
        return self.l2.W // self.l1.W

    @torch.no_grad()
    def render_full(self) -> torch.Tensor:
        """The whole image, exactly the composition of fit.two_level_fit."""
        img1, dx, dy, dxy = render_view_ga(self.l1, 0.0, 0.0, self.l1.W, self.l1.H)
        up = torch.clamp(g.gradient_aware_upscale(img1, dx, dy, dxy, self.l2.H, self.l2.W), 0, 1)
        img2 = torch.clamp(render_view(self.l2, 0.0, 0.0, self.l2.W, self.l2.H), 0, 1)
        return compose(img2, up, self.lo, self.hi)

    def l1_region(self, x0: float, y0: float, w: int, h: int, scale: float):
        """Level-1 pixel region [rx0, rx1) x [ry0, ry1) whose render feeds the
        upscaled window, with the upscaler ROI relative to it.  The window's
        source span w / (f s) must be whole level-1 pixels (the upscaler
        truncates the ROI size to int, upscale.py / gradient_aware_upscaler.cu:201-206)."""
        f = self.factor
        span_x, span_y = w / (f * scale), h / (f * scale)
        if span_x != int(span_x) or span_y != int(span_y):
            raise ValueError(f"window {w}x{h} at scale {scale} spans a fractional number of "
                             f"level-1 pixels ({span_x} x {span_y})")
        sx0, sy0 = x0 / f, y0 / f
        # source samples lie in [s0 - 0.5, s0 + span - 0.5); the Hermite patch
        # reads the two corners around each, so one pixel of margin suffices
        rx0 = max(0, math.floor(sx0 - 0.5) - 1)
        ry0 = max(0, math.floor(sy0 - 0.5) - 1)
        rx1 = min(self.l1.W, math.floor(sx0 + span_x - 0.5) + 3)
        ry1 = min(self.l1.H, math.floor(sy0 + span_y - 0.5) + 3)
        roi = (sx0 - rx0, sy0 - ry0, sx0 - rx0 + span_x, sy0 - ry0 + span_y)
        return (rx0, ry0, rx1, ry1), roi

    @torch.no_grad()
    def render_crop(self, x0: float, y0: float, w: int, h: int,
                    scale: float = 1.0) -> torch.Tensor:
        """h x w x 3 window of the composed image at `scale` output pixels per
        full-resolution pixel, in [0, 1]."""
        (rx0, ry0, rx1, ry1), roi = self.l1_region(x0, y0, w, h, scale)
        img1, dx, dy, dxy = render_view_ga(self.l1, float(rx0), float(ry0), rx1 - rx0,
                                           ry1 - ry0)
        up = torch.clamp(g.gradient_aware_upscale(img1, dx, dy, dxy, h, w, roi), 0, 1)
        img2 = torch.clamp(render_view(self.l2, x0, y0, w, h, scale), 0, 1)
        return compose(img2, up, self.lo, self.hi)
