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
train.py:150-154 does for the whole image.  Two ways of carrying level 1 up to
full resolution: "hermite" (the gradient-aware upscaler's ROI path, upscale.py,
gradient_aware_upscaler.cu:201-206, on the level-1 render and its analytic
derivatives) and "bilinear" (train.py's own composition, train.py:103-105 and
187-196: bilinear resampling of the clamped level-1 render).
`TwoLevel.from_checkpoint` loads the checkpoint train.py writes
(train.py:172-176: per-level {means, cov2d, rgbs} state dicts of the
reference's Gaussian2D, gaussianlig.py:52-58, and store_min / store_max).
"""
from __future__ import annotations

import math
import threading
from dataclasses import dataclass, field
from typing import Optional

import numpy as np
import torch

from . import _lib as L
from . import gsplat2d as g
from .fit import compose
from .gsplat2d import cuda as _C


KINDS = ("cholesky", "cov2d")


@dataclass
class LevelParams:
    """Gaussians of one level on an H x W canvas: means [n, 2], rgbs [n, 3]
    and `par` [n, 3], the Cholesky factors (l11, l21, l22) when kind is
    "cholesky" or the covariance (a, b, c) of the reference's Gaussian2D
    (gaussianlig.py:56, which may be indefinite) when kind is "cov2d"."""
    means: torch.Tensor
    par: torch.Tensor
    rgbs: torch.Tensor
    H: int
    W: int
    kind: str = "cholesky"

    def __post_init__(self):
        if self.kind not in KINDS:
            raise ValueError(f"kind must be one of {KINDS}, got {self.kind!r}")

    @property
    def cholesky(self) -> torch.Tensor:
        if self.kind != "cholesky":
            raise AttributeError("a covariance-parameterised level has no Cholesky factors")
        return self.par

    @property
    def cov2d(self) -> torch.Tensor:
        if self.kind != "cov2d":
            raise AttributeError("a Cholesky-parameterised level has no cov2d parameter")
        return self.par

    @classmethod
    def of(cls, level) -> "LevelParams":
        """The parameters of a fit.Level (Cholesky) or of a Gaussian2D (cov2d)."""
        if hasattr(level, "cholesky"):
            par, kind = level.cholesky, "cholesky"
        else:
            par, kind = level.cov2d, "cov2d"
        return cls(level.means.detach(), par.detach(), level.rgbs.detach(), level.H, level.W,
                   kind)


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
    xys, ext, con, nt = _C.project_gaussians_view(n, p.par, p.means, None, p.kind == "cholesky",
                                                  (x0, y0), (scale, scale), h, w, 16)
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
    ws = _C.bin_workspace_bytes(p.means.shape[0], max(q.num_intersects, 0), q.h, q.w, 16)
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


UPSAMPLE = ("hermite", "bilinear")


def _tile_span(o: int, n: int, full: int):
    """[a, b): the 16-pixel-tile-aligned span of an axis of `full` pixels that
    covers [o, o + n), cut at the image's edge when the window ends inside it."""
    a = (o // 16) * 16
    b = -(-(o + n) // 16) * 16
    if o + n <= full:
        b = min(b, full)
    return a, b


def _aligned_view(p: LevelParams, x0: float, y0: float, w: int, h: int, scale: float):
    """The view (p, x0', y0', w', h', scale) to render for the window (x0, y0,
    w, h) at `scale`, and the window's pixel offset inside it.  A Gaussian is
    drawn only inside its tile rectangle, and an indefinite covariance (which
    the reference's cov2d levels allow, gaussianlig.py:56) can reach pixels
    beyond that rectangle, so a window only holds the full image's pixels when
    its tiles are the full image's tiles: the view is widened to the 16-pixel
    grid of the H*scale x W*scale image.  Windows off that pixel grid render
    as asked."""
    ox, oy = x0 * scale, y0 * scale
    fw, fh = p.W * scale, p.H * scale
    if any(v != int(v) for v in (ox, oy, fw, fh)):
        return (p, x0, y0, w, h, scale), (0, 0)
    ax, bx = _tile_span(int(ox), w, int(fw))
    ay, by = _tile_span(int(oy), h, int(fh))
    return (p, ax / scale, ay / scale, bx - ax, by - ay, scale), (int(ox) - ax, int(oy) - ay)


@dataclass
class TwoLevel:
    l1: LevelParams
    l2: LevelParams
    lo: float
    hi: float
    upsample: str = "hermite"
    # the level-1 canvas rendered once and shared by every crop ("hermite":
    # image, dx, dy, dxy; "bilinear": the clamped image): a crop's level-1
    # source region is small (few tiles, thousands of Gaussians each), so
    # rendering it per request left most of the device idle
    _l1_render: Optional[tuple] = field(default=None, repr=False, compare=False)
    _l1_lock: threading.Lock = field(default_factory=threading.Lock, repr=False, compare=False)

    def __post_init__(self):
        self.lo, self.hi = float(self.lo), float(self.hi)
        if self.upsample not in UPSAMPLE:
            raise ValueError(f"upsample must be one of {UPSAMPLE}, got {self.upsample!r}")

    @property
    def factor(self) -> int:
        return self.l2.W // self.l1.W

    @classmethod
    def from_checkpoint(cls, path: str, H: int, W: int, device="cuda:0",
                        upsample: str = "bilinear") -> "TwoLevel":
        """A two-level representation from a checkpoint in train.py's format
        (train.py:172-176; loaded as train.py:51-67 does): 'state_dict' holds
        one state dict per level with keys {means, cov2d, rgbs} (the
        reference's Gaussian2D) or {means, cholesky, rgbs} (fit.Level), and
        'store_min' / 'store_max' the level-2 normalisation bounds.  The
        checkpoint does not store level sizes: level s of an n-level model is
        int(H * 2^(s + 1 - n)) x int(W * 2^(s + 1 - n)) (gaussianlig.py:20-22),
        so H, W are the full image's.  upsample="bilinear" composes as
        train.py does; the loader executes nothing from the file
        (torch.load(weights_only=True))."""
        ck = torch.load(path, map_location=device, weights_only=True)
        sds = ck["state_dict"]
        if len(sds) != 2:
            raise ValueError(f"a two-level representation needs 2 levels, the checkpoint has "
                             f"{len(sds)}")
        levels = []
        for s, sd in enumerate(sds):
            kinds = [k for k in KINDS if k in sd]
            if len(kinds) != 1 or set(sd) != {"means", kinds[0], "rgbs"}:
                raise KeyError(f"level {s}: expected keys means, rgbs and one of {KINDS}, got "
                               f"{sorted(sd)}")
            f = 2.0 ** (s + 1 - len(sds))
            levels.append(LevelParams(*(sd[k].detach().float().contiguous().to(device)
                                        for k in ("means", kinds[0], "rgbs")),
                                      int(H * f), int(W * f), kinds[0]))
        if not ck["store_min"] or not ck["store_max"]:
            raise KeyError("the checkpoint holds no store_min / store_max for level 2")
        return cls(levels[0], levels[1], float(ck["store_min"][0]), float(ck["store_max"][0]),
                   upsample=upsample)

    @torch.no_grad()
    def level1_render(self):
        """The whole level-1 canvas, rendered on first use and kept resident:
        (img, dx, dy, dxy) for "hermite" (48 bytes per level-1 pixel), the
        image clamped to [0, 1] (as Gaussian2D.forward returns it,
        gaussianlig.py:95) for "bilinear"."""
        cached = self._l1_render
        if cached is None:
            with self._l1_lock:
                if self._l1_render is None:
                    if self.upsample == "hermite":
                        out = render_view_ga(self.l1, 0.0, 0.0, self.l1.W, self.l1.H)
                    else:
                        out = torch.clamp(render_view(self.l1, 0.0, 0.0, self.l1.W, self.l1.H),
                                          0, 1)
                    torch.cuda.current_stream(self.l1.means.device).synchronize()
                    self._l1_render = out
                cached = self._l1_render
        return cached

    def _bilinear_window(self, x0: float, y0: float, w: int, h: int, scale: float):
        """The window of the clamped level-1 render bilinearly resized to the
        H*scale x W*scale image (the crop of train.py's resampled estimate)."""
        fh, fw = self.l2.H * scale, self.l2.W * scale
        ox, oy = x0 * scale, y0 * scale
        if fh != int(fh) or fw != int(fw) or ox != int(ox) or oy != int(oy):
            raise ValueError(f"bilinear composition needs whole output pixels: image "
                             f"{fw} x {fh}, window origin ({ox}, {oy}) at scale {scale}")
        return _C.bilinear_resize(self.level1_render(), int(fh), int(fw),
                                  (int(ox), int(oy), w, h))

    @torch.no_grad()
    def render_full(self) -> torch.Tensor:
        """The whole image: fit.two_level_fit's composition ("hermite"), or
        train.py's (train.py:187-196) for "bilinear"."""
        if self.upsample == "bilinear":
            up = self._bilinear_window(0.0, 0.0, self.l2.W, self.l2.H, 1.0)
        else:
            img1, dx, dy, dxy = self.level1_render()
            up = torch.clamp(g.gradient_aware_upscale(img1, dx, dy, dxy, self.l2.H, self.l2.W),
                             0, 1)
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
        # reads the two corners around each, so one pixel of margin suffices;
        # the region is then widened to level 1's 16-pixel tile grid (see
        # _aligned_view)
        rx0 = max(0, math.floor(sx0 - 0.5) - 1)
        ry0 = max(0, math.floor(sy0 - 0.5) - 1)
        rx1 = min(self.l1.W, math.floor(sx0 + span_x - 0.5) + 3)
        ry1 = min(self.l1.H, math.floor(sy0 + span_y - 0.5) + 3)
        rx0, rx1 = _tile_span(rx0, rx1 - rx0, self.l1.W)
        ry0, ry1 = _tile_span(ry0, ry1 - ry0, self.l1.H)
        roi = (sx0 - rx0, sy0 - ry0, sx0 - rx0 + span_x, sy0 - ry0 + span_y)
        return (rx0, ry0, rx1, ry1), roi

    def crop_plan(self, x0: float, y0: float, w: int, h: int, scale: float = 1.0,
                  level1_cache: bool = True):
        """The level-2 window's projection (and, without the level-1 cache,
        the level-1 source region's), the upscaler ROI, the HBM the rest of
        the render needs, and the window's offset in its tile-aligned level-2
        view (plan[3] is the byte count).  With the cache the ROI is in
        level-1 canvas coordinates and level 1 is not rendered per request."""
        v2, off = _aligned_view(self.l2, x0, y0, w, h, scale)
        if self.upsample == "bilinear":
            (q2,) = project_windows([v2])
            return (None, q2, (x0, y0, scale), window_bytes(self.l2, q2, False) + 3 * 12 * w * h,
                    off)
        (rx0, ry0, rx1, ry1), roi = self.l1_region(x0, y0, w, h, scale)
        if level1_cache:
            (q2,) = project_windows([v2])
            roi = (roi[0] + rx0, roi[1] + ry0, roi[2] + rx0, roi[3] + ry0)
            return None, q2, roi, window_bytes(self.l2, q2, False) + 3 * 12 * w * h, off
        q1, q2 = project_windows([(self.l1, float(rx0), float(ry0), rx1 - rx0, ry1 - ry0, 1.0),
                                  v2])
        need = (window_bytes(self.l1, q1, True) + window_bytes(self.l2, q2, False) +
                3 * 12 * w * h)
        return q1, q2, roi, need, off

    @torch.no_grad()
    def finish_crop(self, plan, w: int, h: int) -> torch.Tensor:
        q1, q2, roi, _, (ox, oy) = plan
        if self.upsample == "bilinear":
            up = self._bilinear_window(roi[0], roi[1], w, h, roi[2])
        else:
            img1, dx, dy, dxy = (self.level1_render() if q1 is None
                                 else raster_window_ga(self.l1, q1))
            up = torch.clamp(g.gradient_aware_upscale(img1, dx, dy, dxy, h, w, roi), 0, 1)
        img2 = torch.clamp(raster_window(self.l2, q2)[oy:oy + h, ox:ox + w], 0, 1)
        return compose(img2, up, self.lo, self.hi)

    @torch.no_grad()
    def render_crop(self, x0: float, y0: float, w: int, h: int, scale: float = 1.0,
                    level1_cache: bool = True) -> torch.Tensor:
        """h x w x 3 window of the composed image at `scale` output pixels per
        full-resolution pixel, in [0, 1]."""
        return self.finish_crop(self.crop_plan(x0, y0, w, h, scale, level1_cache), w, h)

    def crop_geometry(self, x0: float, y0: float, w: int, h: int, scale: float = 1.0):
        """Where a crop's inputs come from, without touching the device: its
        tile-aligned level-2 view (params, x0', y0', w', h', scale), the
        window's offset in that view, and the level-1 source -- ("hermite",
        ROI on the level-1 canvas) or ("bilinear", full_h, full_w, win_x,
        win_y).  Raises ValueError where render_crop would."""
        if int(w) != w or int(h) != h or w < 1 or h < 1 or not scale > 0:
            raise ValueError(f"crop {w} x {h} at scale {scale}: sizes must be positive "
                             f"integers and the scale positive")
        v2, off = _aligned_view(self.l2, x0, y0, int(w), int(h), scale)
        if self.upsample == "bilinear":
            fh, fw = self.l2.H * scale, self.l2.W * scale
            ox, oy = x0 * scale, y0 * scale
            if fh != int(fh) or fw != int(fw) or ox != int(ox) or oy != int(oy):
                raise ValueError(f"bilinear composition needs whole output pixels: image "
                                 f"{fw} x {fh}, window origin ({ox}, {oy}) at scale {scale}")
            if ox < 0 or oy < 0 or ox + w > fw or oy + h > fh:
                raise ValueError(f"window ({ox}, {oy}) + {w} x {h} leaves the {fw} x {fh} "
                                 f"image")
            return v2, off, ("bilinear", int(fh), int(fw), int(ox), int(oy))
        (rx0, ry0, _, _), roi = self.l1_region(x0, y0, w, h, scale)
        return v2, off, ("hermite", (roi[0] + rx0, roi[1] + ry0, roi[2] + rx0, roi[3] + ry0))


# -- batched crops (serving) ---------------------------------------------------

_VIEW_DTYPE = np.dtype([("par", "<u8"), ("means2d", "<u8"), ("rgbs", "<u8"), ("n", "<i4"),
                        ("cholesky", "<i4"), ("origin_x", "<f4"), ("origin_y", "<f4"),
                        ("scale_x", "<f4"), ("scale_y", "<f4"), ("img_height", "<i4"),
                        ("img_width", "<i4"), ("gauss_off", "<i4"), ("reserved", "<i4")])
_CROP_DTYPE = np.dtype([("level2", "<u8"), ("f", "<u8"), ("fx", "<u8"), ("fy", "<u8"),
                        ("fxy", "<u8"), ("out", "<u8"), ("view_w", "<i4"), ("off_x", "<i4"),
                        ("off_y", "<i4"), ("w", "<i4"), ("h", "<i4"), ("mode", "<i4"),
                        ("src_h", "<i4"), ("src_w", "<i4"), ("roi_x1", "<f4"),
                        ("roi_y1", "<f4"), ("roi_x2", "<f4"), ("roi_y2", "<f4"),
                        ("full_h", "<i4"), ("full_w", "<i4"), ("win_x", "<i4"),
                        ("win_y", "<i4"), ("span", "<f4"), ("lo", "<f4"),
                        ("block_off", "<i4"), ("reserved", "<i4")])
assert _VIEW_DTYPE.itemsize == 64 and _CROP_DTYPE.itemsize == 128


def _to_device(parts, dev) -> list:
    """Host tables (numpy arrays) copied to the device in one transfer on the
    current stream (pinned staging, so the copy is asynchronous): uint8
    device views, each starting on a 256-byte boundary."""
    offs, o = [], 0
    for a in parts:
        offs.append(o)
        o += -(-a.nbytes // 256) * 256
    host = torch.empty(max(o, 1), dtype=torch.uint8, pin_memory=True)
    hv = host.numpy()
    for a, off in zip(parts, offs):
        hv[off:off + a.nbytes] = np.ascontiguousarray(a).view(np.uint8).reshape(-1)
    d = host.to(dev, non_blocking=True)
    return [d[off:off + a.nbytes] for a, off in zip(parts, offs)]


class CropBatch:
    """Several crops, of any of the resident representations, rendered as one
    batch: one projection launch over every crop's level-2 view
    (lig_project_views_batched), one host read of the views' intersection
    counts, one batched binning (lig_bin_batched: each view an image of the
    batch table, as in batched fitting), one render of every view
    (lig_render_batched) and one composition launch for all crops
    (lig_compose_crops).  Each crop gets the bits TwoLevel.render_crop gives
    it alone (level 1 from the shared level-1 render).  Like crop_plan /
    finish_crop: construction projects and reads the counts, `need` is the
    HBM the rest allocates, `finish()` renders."""

    depth = 64   # LDS batch depth of the render (lig_render_batched's batch_depth)

    def __init__(self, items):
        """items: (TwoLevel, x0, y0, w, h, scale) per crop."""
        from .batch import batch_table
        if not items:
            raise ValueError("an empty crop batch")
        self.items = [(rep, float(x0), float(y0), int(w), int(h), float(scale))
                      for rep, x0, y0, w, h, scale in items]
        self.geo = [rep.crop_geometry(x0, y0, w, h, scale)
                    for rep, x0, y0, w, h, scale in self.items]
        dev = self.dev = self.items[0][0].l2.means.device
        K = len(self.items)
        views = np.zeros(K, _VIEW_DTYPE)
        sizes, counts_n, goff = [], [], 0
        for k, ((p, vx, vy, vw, vh, vs), _, _) in enumerate(self.geo):
            if p.means.device != dev:
                raise ValueError("a crop batch renders on one device")
            n = int(p.means.shape[0])
            views[k] = (p.par.data_ptr(), p.means.data_ptr(), p.rgbs.data_ptr(), n,
                        int(p.kind == "cholesky"), vx, vy, vs, vs, vh, vw, goff, 0)
            sizes.append((vh, vw))
            counts_n.append(n)
            goff += n
        if goff >= 2 ** 31:
            raise ValueError(f"a crop batch of {goff} Gaussians exceeds int32 slots")
        self.n = n_tot = goff
        self.counts_n = counts_n
        tab, self.num_tiles, self.view_pix = batch_table(sizes, counts_n)
        ends = np.cumsum(counts_n, dtype=np.int64) - 1
        self._views, self._tab, ends_b = _to_device([views, tab, np.maximum(ends, 0)], dev)
        # projection outputs in one allocation: xys, extents [n,2], conics,
        # colors [n,3], tiles hit [n] int32
        buf = torch.empty(max(n_tot, 1) * 11, device=dev)
        self.xys = buf[:2 * n_tot].view(n_tot, 2)
        self.ext = buf[2 * n_tot:4 * n_tot].view(n_tot, 2)
        self.con = buf[4 * n_tot:7 * n_tot].view(n_tot, 3)
        self.colors = buf[7 * n_tot:10 * n_tot].view(n_tot, 3)
        self.nt = buf[10 * n_tot:11 * n_tot].view(torch.int32)
        self._buf = buf
        s = L.stream_of(buf)
        with torch.cuda.device(dev):
            L.call("lig_project_views_batched", K, L.ptr(self._views), n_tot, 16,
                   L.ptr(self.xys), L.ptr(self.ext), L.ptr(self.con), L.ptr(self.nt),
                   L.ptr(self.colors), s)
        if n_tot:
            self.cum = _C.inclusive_cumsum_i32(self.nt)
            upto = self.cum.index_select(0, ends_b.view(torch.int64)).cpu().tolist()
        else:
            self.cum = torch.zeros(0, dtype=torch.int32, device=dev)
            upto = [0] * K
        before = 0
        self.counts = []
        for k in range(K):
            c = upto[k] if counts_n[k] > 0 else before
            self.counts.append(int(c) - before)
            before = int(c)
        self.num_intersects = before
        self.out_pix = sum(w * h for _, _, _, w, h, _ in self.items)
        m = self.num_intersects
        self.need = int((L.size_query("lig_bin_workspace_size", m, self.num_tiles) if m else 0) +
                        4 * m + 8 * self.num_tiles + 64 * n_tot +
                        L.size_query("lig_fit_batched_workspace_size", self.num_tiles) +
                        12 * self.view_pix + 12 * self.out_pix + 128 * K)

    @torch.no_grad()
    def finish(self) -> list:
        """The crops, h x w x 3 each, in [0, 1] (views of one output buffer).
        Afterwards `view_images` holds each crop's level-2 view render (when
        any view has intersections)."""
        dev = self.dev
        s = L.stream_of(self.xys)
        geo, K = self.geo, len(self.items)
        pix_off = np.cumsum([0] + [gm[0][4] * gm[0][3] for gm in geo])
        img = None
        if self.num_intersects > 0:
            m, tiles = self.num_intersects, self.num_tiles
            gids = torch.empty(m, dtype=torch.int32, device=dev)
            bins = torch.empty(tiles, 2, dtype=torch.int32, device=dev)
            ws = torch.empty(max(L.size_query("lig_bin_workspace_size", m, tiles), 1),
                             dtype=torch.uint8, device=dev)
            img = torch.empty(self.view_pix * 3, device=dev)
            packed = torch.empty(L.size_query("lig_packed_gaussians_bytes", self.n),
                                 dtype=torch.uint8, device=dev)
            rws = torch.empty(L.size_query("lig_fit_batched_workspace_size", tiles),
                              dtype=torch.uint8, device=dev)
            with torch.cuda.device(dev):
                L.call("lig_bin_batched", self.n, m, L.ptr(self.xys), L.ptr(self.ext),
                       L.ptr(self.cum), L.ptr(self._tab), K, tiles, L.ptr(gids), L.ptr(bins),
                       L.ptr(ws), ws.numel(), s)
                L.call("lig_render_batched", self.n, L.ptr(gids), L.ptr(bins), L.ptr(self.xys),
                       L.ptr(self.ext), L.ptr(self.cum), L.ptr(self.con), L.ptr(self.colors),
                       L.ptr(packed), L.ptr(self._tab), K, tiles, L.ptr(img), self.depth,
                       L.ptr(rws), rws.numel(), s)
            self.view_images = [img[3 * int(pix_off[k]):3 * int(pix_off[k + 1])]
                                .view(gm[0][4], gm[0][3], 3) for k, gm in enumerate(geo)]
        out = torch.empty(self.out_pix * 3, device=dev)
        crops = np.zeros(K, _CROP_DTYPE)
        results, o, blk = [], 0, 0
        base = out.data_ptr()
        for k, ((rep, x0, y0, w, h, scale), ((p, vx, vy, vw, vh, vs), (ox, oy), src)) in \
                enumerate(zip(self.items, geo)):
            l2 = 0 if img is None or self.counts[k] == 0 else \
                img.data_ptr() + 12 * int(pix_off[k])
            l1 = rep.level1_render()
            span = rep.hi - rep.lo   # in FP64, rounded once to FP32, as torch does
            if src[0] == "hermite":
                f, fx, fy, fxy = l1
                r = src[1]
                crops[k] = (l2, f.data_ptr(), fx.data_ptr(), fy.data_ptr(), fxy.data_ptr(),
                            base + 12 * o, vw, ox, oy, w, h, 0, rep.l1.H, rep.l1.W, r[0], r[1],
                            r[2], r[3], 0, 0, 0, 0, span, rep.lo, blk, 0)
            else:
                _, fh, fw, wx, wy = src
                crops[k] = (l2, l1.data_ptr(), 0, 0, 0, base + 12 * o, vw, ox, oy, w, h, 1,
                            rep.l1.H, rep.l1.W, 0.0, 0.0, 0.0, 0.0, fh, fw, wx, wy, span, rep.lo,
                            blk, 0)
            results.append(out[3 * o:3 * (o + w * h)].view(h, w, 3))
            o += w * h
            blk += -(-(w * h) // 256)
        (crops_d,) = _to_device([crops], dev)
        with torch.cuda.device(dev):
            L.call("lig_compose_crops", K, L.ptr(crops_d), blk, s)
        return results


@torch.no_grad()
def render_crops(items) -> list:
    """The crops (TwoLevel, x0, y0, w, h, scale) rendered as one batch
    (CropBatch): the same bits as render_crop of each."""
    return CropBatch(items).finish()
