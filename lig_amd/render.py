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
import threading
from dataclasses import dataclass, field
from typing import Optional

import torch

from . import _lib as L
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


@dataclass
class Projected:
    """A level projected into a window, with its intersection count."""
    xys: torch.Tensor
    extents: torch.Tensor
    conics: torch.Tensor
    num_tiles_hit: torch.Tensor
    cum: torch.Tensor
    num_intersects: int
    h: int
    w: int


def _project_scan(p: LevelParams, x0: float, y0: float, w: int, h: int, scale: float):
    n = p.means.shape[0]
    xys, ext, con, nt = _C.project_gaussians_view(n, p.cholesky, p.means, None, True, (x0, y0),
                                                  (scale, scale), h, w, 16)
    cum = _C.inclusive_cumsum_i32(nt)
    return xys, ext, con, nt, cum


def project_windows(views) -> list:
    """Projections of several (params, x0, y0, w, h, scale) views, with one
    device-to-host read for all their intersection counts."""
    parts = [_project_scan(*v) for v in views]
    lasts = [c[-1:] if c.numel() else torch.zeros(1, dtype=torch.int32, device=c.device)
             for *_, c in parts]
    counts = torch.cat(lasts).cpu().tolist() if parts else []
    return [Projected(xys, ext, con, nt, cum, int(m), v[4], v[3])
            for (xys, ext, con, nt, cum), m, v in zip(parts, counts, views)]


def project_window(p: LevelParams, x0: float, y0: float, w: int, h: int,
                   scale: float = 1.0) -> Projected:
    """View projection and the inclusive scan of tiles hit (one device-to-host
    read of the intersection count, on the current stream)."""
    return project_windows([(p, x0, y0, w, h, scale)])[0]


def raster_window(p: LevelParams, q: Projected) -> torch.Tensor:
    """Image-only render of a projected window."""
    if q.num_intersects < 1:
        return torch.ones(q.h, q.w, 3, device=p.means.device)   # rasterize.py:80-90
    gids, bins = _C.bin_and_sort_fast(p.means.shape[0], q.num_intersects, q.xys, q.extents,
                                      q.cum, q.h, q.w, 16)
    return _C.rasterize_image(q.h, q.w, gids, bins, q.xys, q.conics, p.rgbs)


def raster_window_ga(p: LevelParams, q: Projected):
    """(img, dx, dy, dxy) of a projected window (gradient-aware forward)."""
    dev = p.means.device
    if q.num_intersects < 1:
        z = torch.zeros(q.h, q.w, 3, device=dev)
        return torch.ones(q.h, q.w, 3, device=dev), z, z.clone(), z.clone()
    gids, bins = _C.bin_and_sort_fast(p.means.shape[0], q.num_intersects, q.xys, q.extents,
                                      q.cum, q.h, q.w, 16)
    img, _, dx, dy, dxy, _, _ = _C.rasterize_forward_packed(q.h, q.w, 16, gids, bins, q.xys,
                                                            q.conics, p.rgbs, None, True)
    return img, dx, dy, dxy


def window_bytes(p: LevelParams, q: Projected, ga: bool) -> int:
    """HBM the rest of a window's render allocates after projection: binning
    workspace and outputs, packed records, images."""
    tiles = ((q.h + 15) // 16) * ((q.w + 15) // 16)
    ws = L.size_query("lig_bin_workspace_size", max(q.num_intersects, 0), tiles)
    pix = q.h * q.w
    per_pix = (12 + 4 + 36 + 4) if ga else 12
    return int(ws + 4 * q.num_intersects + 8 * tiles + 64 * p.means.shape[0] + per_pix * pix)


@torch.no_grad()
def render_view(p: LevelParams, x0: float, y0: float, w: int, h: int,
                scale: float = 1.0) -> torch.Tensor:
    """h x w x 3 raw sum image of the window (x0, y0) at `scale`."""
    return raster_window(p, project_window(p, x0, y0, w, h, scale))


@torch.no_grad()
def render_view_ga(p: LevelParams, x0: float, y0: float, w: int, h: int):
    """(img, dx, dy, dxy) of the window at scale 1 (gradient-aware forward)."""
    return raster_window_ga(p, project_window(p, x0, y0, w, h, 1.0))


@dataclass
class TwoLevel:
    l1: LevelParams
    l2: LevelParams
    lo: float
    hi: float
    # the level-1 canvas rendered once (image, dx, dy, dxy), shared by every
    # crop: a crop's level-1 source region is small (few tiles, thousands of
    # Gaussians each), so rendering it per request left most of the device idle
    _l1_render: Optional[tuple] = field(default=None, repr=False, compare=False)
    _l1_lock: threading.Lock = field(default_factory=threading.Lock, repr=False, compare=False)

    @property
    def factor(self) -> int:
        return self.l2.W // self.l1.W

    @torch.no_grad()
    def level1_render(self):
        """(img, dx, dy, dxy) of the whole level-1 canvas, rendered on first use
        and kept resident (48 bytes per level-1 pixel)."""
        cached = self._l1_render
        if cached is None:
            with self._l1_lock:
                if self._l1_render is None:
                    out = render_view_ga(self.l1, 0.0, 0.0, self.l1.W, self.l1.H)
                    torch.cuda.current_stream(self.l1.means.device).synchronize()
                    self._l1_render = out
                cached = self._l1_render
        return cached

    @torch.no_grad()
    def render_full(self) -> torch.Tensor:
        """The whole image, exactly the composition of fit.two_level_fit."""
        img1, dx, dy, dxy = self.level1_render()
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

    def crop_plan(self, x0: float, y0: float, w: int, h: int, scale: float = 1.0,
                  level1_cache: bool = True):
        """The level-2 window's projection (and, without the level-1 cache,
        the level-1 source region's), the upscaler ROI and the HBM the rest of
        the render needs.  With the cache the ROI is in level-1 canvas
        coordinates and level 1 is not rendered per request."""
        (rx0, ry0, rx1, ry1), roi = self.l1_region(x0, y0, w, h, scale)
        if level1_cache:
            (q2,) = project_windows([(self.l2, x0, y0, w, h, scale)])
            roi = (roi[0] + rx0, roi[1] + ry0, roi[2] + rx0, roi[3] + ry0)
            return None, q2, roi, window_bytes(self.l2, q2, False) + 3 * 12 * w * h
        q1, q2 = project_windows([(self.l1, float(rx0), float(ry0), rx1 - rx0, ry1 - ry0, 1.0),
                                  (self.l2, x0, y0, w, h, scale)])
        need = (window_bytes(self.l1, q1, True) + window_bytes(self.l2, q2, False) +
                3 * 12 * w * h)
        return q1, q2, roi, need

    @torch.no_grad()
    def finish_crop(self, plan, w: int, h: int) -> torch.Tensor:
        q1, q2, roi, _ = plan
        img1, dx, dy, dxy = self.level1_render() if q1 is None else raster_window_ga(self.l1, q1)
        up = torch.clamp(g.gradient_aware_upscale(img1, dx, dy, dxy, h, w, roi), 0, 1)
        img2 = torch.clamp(raster_window(self.l2, q2), 0, 1)
        return compose(img2, up, self.lo, self.hi)

    @torch.no_grad()
    def render_crop(self, x0: float, y0: float, w: int, h: int, scale: float = 1.0,
                    level1_cache: bool = True) -> torch.Tensor:
        """h x w x 3 window of the composed image at `scale` output pixels per
        full-resolution pixel, in [0, 1]."""
        return self.finish_crop(self.crop_plan(x0, y0, w, h, scale, level1_cache), w, h)
