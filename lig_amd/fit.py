"""Fit harness: the callers of the gsplat2d path.

`Level` is one resolution level parameterised by Cholesky factors (the form
the workloads use): project_gaussians_cholesky -> rasterize_gaussians ->
clamp(0, 1) -> MSE -> backward -> Adam.  `Gaussian2D` and `LIG` mirror the
reference's model classes (gaussianlig.py:9-111: covariance parameters,
rgbs initialised to zero, the level/point split), and `two_level_fit`
follows the two-level flow of train.py:92-158 (downsampled level-1 target,
normalised residual target for level 2, de-normalised composition).
"""
from __future__ import annotations

import math
from typing import List, Optional

import torch
import torch.nn as nn
import torch.nn.functional as F

from . import gsplat2d as g
from .data import init_gaussians
from .gsplat2d import cuda as _C
from .loss import l2_loss_clamped
from .optim import FusedAdam


def _adam(params, lr):
    """Adam with the fused HIP step (lig_adam_step), torch.optim.Adam's update."""
    return FusedAdam(params, lr=lr)


class _RenderL2(torch.autograd.Function):
    """mean((clamp(rasterize(project_cholesky(...)), 0, 1) - target)^2) as one
    autograd node, 16 x 16 tiles, sum mode, no opacities: the loss is fused
    into the rasteriser (lig_fit_render_l2 / lig_fit_backward_records_l2).
    Image and gradients carry the bits of the unfused chain
    project_gaussians_cholesky -> rasterize_gaussians -> l2_loss_clamped;
    the loss value is summed per tile instead of per 8192-float block."""

    @staticmethod
    def forward(ctx, means, cholesky, rgbs, target, H, W, depth=0):
        n = means.shape[0]
        xys, ext, conics, nt = _C.project_gaussians_forward_cholesky(n, cholesky, means, None,
                                                                     H, W, 16)
        num_int, cum = g.compute_cumulative_intersects(nt)
        ctx.n, ctx.hw, ctx.empty = n, (H, W), num_int < 1
        if num_int < 1:
            # rasterize.py:80-90: no intersection renders an image of ones,
            # which depends on no parameter
            ctx.shapes = (means.shape, cholesky.shape, rgbs.shape)
            return ((1.0 - target) ** 2).mean()
        gids, bins = _C.bin_and_sort_fast(n, num_int, xys, ext, cum, H, W, 16)
        img, loss, packed = _C.fit_render_l2(H, W, gids, bins, xys, ext, cum, conics, rgbs,
                                             target, depth)
        ctx.num_int = num_int
        ctx.save_for_backward(cholesky, ext, conics, gids, bins, cum, img, target, packed)
        return loss

    @staticmethod
    def backward(ctx, grad_loss):
        if ctx.empty:
            z = [torch.zeros(s, dtype=grad_loss.dtype, device=grad_loss.device)
                 for s in ctx.shapes]
            return z[0], z[1], z[2], None, None, None, None
        cholesky, ext, conics, gids, bins, cum, img, target, packed = ctx.saved_tensors
        H, W = ctx.hw
        v_xy, v_conic, v_rgb = _C.fit_backward_l2(H, W, gids, bins, packed, ctx.n,
                                                     ctx.num_int, img, target, grad_loss.float(),
                                                     cum)
        v_chol, v_mean = _C.project_gaussians_backward_cholesky(ctx.n, ext, cholesky, conics,
                                                                v_xy, v_conic)
        return v_mean, v_chol, v_rgb, None, None, None, None


def render_l2_loss(means, cholesky, rgbs, target, H: int, W: int,
                   batch_depth: int = 0) -> torch.Tensor:
    """Fused clamp + L2 fit loss of a Cholesky level (16 x 16 tiles); batch_depth
    is the forward's LDS batch depth (64, 128, 256; 0 = 64)."""
    return _RenderL2.apply(means, cholesky, rgbs, target.contiguous(), H, W, batch_depth)


class Level(nn.Module):
    """One level: n Gaussians on an H x W canvas, Cholesky parameterisation.
    step(): render -> clamp + L2 loss -> backward -> Adam, by default through
    the fused render + loss node (16 x 16 tiles); fused=False runs the
    gsplat2d API chain (rasterize_gaussians + l2_loss_clamped) instead."""

    def __init__(self, n: int, H: int, W: int, seed: int, scale: float = 1.0, lr: float = 1e-3,
                 block_width: int = 16, device="cuda", fused: bool = True):
        super().__init__()
        means, chol, rgbs = init_gaussians(n, H, W, seed, scale, device)
        self.means = nn.Parameter(means)
        self.cholesky = nn.Parameter(chol)
        self.rgbs = nn.Parameter(rgbs)
        self.H, self.W, self.block_width = H, W, block_width
        self.fused = fused and block_width == 16
        self.batch_depth = 0   # forward LDS batch depth (0 = default; see lig_amd.autotune)
        self.optimizer = _adam([self.rgbs, self.means, self.cholesky], lr)

    def project(self):
        return g.project_gaussians_cholesky(self.cholesky, self.means, self.H, self.W,
                                            self.block_width)

    def render(self, gradient_aware: bool = False):
        """Returns the raw [H, W, 3] sum image (and dx, dy, dxy when asked)."""
        xys, ext, con, nt = self.project()
        out = g.rasterize_gaussians(xys, ext, con, nt, self.rgbs, None, self.H, self.W,
                                    self.block_width, compute_upscale_gradients=gradient_aware)
        return out if gradient_aware else out[0]

    def loss(self, target_hwc: torch.Tensor) -> torch.Tensor:
        if self.fused:
            return render_l2_loss(self.means, self.cholesky, self.rgbs, target_hwc, self.H,
                                  self.W, self.batch_depth)
        return l2_loss_clamped(self.render(), target_hwc)

    def graph(self, target_hwc: torch.Tensor, headroom: float = 1.25):
        """The fused step against `target_hwc` recorded once as a hipGraph
        (lig_amd.batch.BatchedStepGraph over this level as a batch of one):
        graph.run(k) takes k steps with the bits of k step() calls and no host
        read of the intersection count.  Buffers hold `headroom` times the
        count at capture; a step that outgrows them is redone after a new
        capture (512^2 x 20 K Gaussians at lr 1e-2 over 500 steps: 1.05 made
        3 captures and ran 4,650 steps/s, 1.25 one capture and 8,080)."""
        from .batch import BatchedStepGraph, LevelBatch
        if not self.fused:
            raise ValueError("the captured step is the fused (16 x 16 tile) step")
        return BatchedStepGraph(LevelBatch(self, target_hwc), headroom=headroom)

    def step(self, target_hwc: torch.Tensor) -> torch.Tensor:
        """One optimiser step against an [H, W, 3] target; returns the loss
        tensor (no host synchronisation beyond the binning's count)."""
        loss = self.loss(target_hwc)
        loss.backward()
        self.optimizer.step()
        self.optimizer.zero_grad(set_to_none=True)
        return loss.detach()


def psnr(img: torch.Tensor, target: torch.Tensor) -> float:
    mse = F.mse_loss(img.float(), target.float()).item()
    return 10.0 * math.log10(1.0 / max(mse, 1e-20))


def evaluate(out_hwc: torch.Tensor, gt_hwc: torch.Tensor):
    """(PSNR, MS-SSIM) of a composed H x W x 3 image against the ground truth,
    as the reference's test() scores a fit (train.py:197-200, data range 1)."""
    from .metrics import ms_ssim
    x = out_hwc.float().permute(2, 0, 1)[None]
    y = gt_hwc.float().permute(2, 0, 1)[None]
    return psnr(out_hwc, gt_hwc), float(ms_ssim(x, y, data_range=1, size_average=True))


def downsample(img_hwc: torch.Tensor, factor: int, mode: str = "bicubic") -> torch.Tensor:
    """H x W x 3 -> H/factor x W/factor x 3: antialiased bicubic, or "area"
    (the mean of each factor x factor block, train.py:95-98's
    interpolate(scale_factor=2^-k, mode='area'))."""
    x = img_hwc.permute(2, 0, 1)[None]
    H, W = img_hwc.shape[0] // factor, img_hwc.shape[1] // factor
    if mode == "area":
        y = F.interpolate(x, scale_factor=1.0 / factor, mode="area")
    else:
        kw = dict(antialias=True, align_corners=False) if mode == "bicubic" else {}
        y = F.interpolate(x, size=(H, W), mode=mode, **kw)
    return y[0].permute(1, 2, 0).contiguous()


def upscale_render(level: Level, H: int, W: int) -> torch.Tensor:
    """Level render upscaled to H x W with the gradient-aware Hermite upscaler."""
    img, wsum, dx, dy, dxy = level.render(gradient_aware=True)
    return g.gradient_aware_upscale(img, dx, dy, dxy, H, W)


def residual_target(gt_hwc: torch.Tensor, prev_up: torch.Tensor):
    """train.py:113-135: residual + 0.5, min-max normalised; returns
    (target, store_min, store_max)."""
    t = gt_hwc - prev_up + 0.5
    lo, hi = t.min(), t.max()
    return (t - lo) / (hi - lo), lo, hi


def compose(level2_img: torch.Tensor, prev_up: torch.Tensor, lo, hi) -> torch.Tensor:
    """train.py:150-154 / 187-196: de-normalise the level-2 render and add it
    to the upscaled level-1 estimate."""
    return torch.clamp(level2_img * (hi - lo) + prev_up - 0.5 + lo, 0, 1)


def fit_level(level: Level, target_hwc: torch.Tensor, steps: int):
    """`steps` optimiser steps of `level`: replayed as a captured step
    (Level.graph, the same bits; small images are otherwise bound by the
    host's launch work) when the level is fused, else eager."""
    if steps <= 0:
        return
    if level.fused:
        level.graph(target_hwc).run(steps)
    else:
        for _ in range(steps):
            level.step(target_hwc)


def bilinear_render(level: Level, H: int, W: int) -> torch.Tensor:
    """train.py:103-105: the clamped level render resized to H x W by bilinear
    interpolation (half-pixel grid, lig_bilinear_resize)."""
    return _C.bilinear_resize(torch.clamp(level.render(), 0, 1).contiguous(), H, W)


FLOWS = ("lig", "train")


def two_level_fit(gt_hwc: torch.Tensor, n1: int, n2: int, steps: int, seed: int = 0,
                  scale: float = 1.0, lr: float = 1e-3, factor: int = 4, device="cuda",
                  flow: str = "lig"):
    """Plain two-level fit; returns (level1, level2, final PSNR).

    flow "lig": level 1 fits the antialiased bicubic downsample and is carried
    up by the gradient-aware Hermite upscaler from its render and analytic
    derivatives (the upscaler the path provides, upscale.py:50-82).  flow
    "train": train.py's own resampling, level 1 fits the area downsample
    (train.py:95-98) and the previous estimate is the bilinear resize of its
    clamped render (train.py:103-105).  Both then fit level 2 to the
    normalised residual (train.py:113-135) and compose (train.py:150-154)."""
    if flow not in FLOWS:
        raise ValueError(f"flow must be one of {FLOWS}, got {flow!r}")
    H, W = gt_hwc.shape[:2]
    l1 = Level(n1, H // factor, W // factor, seed, scale, lr, device=device)
    t1 = downsample(gt_hwc, factor, "bicubic" if flow == "lig" else "area")
    fit_level(l1, t1, steps)
    with torch.no_grad():
        if flow == "lig":
            prev = torch.clamp(upscale_render(l1, H, W), 0, 1)
        else:
            prev = bilinear_render(l1, H, W)
        t2, lo, hi = residual_target(gt_hwc, prev)
    l2 = Level(n2, H, W, seed + 1, scale, lr, device=device)
    fit_level(l2, t2, steps)
    with torch.no_grad():
        out = compose(torch.clamp(l2.render(), 0, 1), prev, lo, hi)
    return l1, l2, psnr(out, gt_hwc)


# ---------------------------------------------------------------------------
# Mirrors of the reference's model classes (gaussianlig.py:9-111)
# ---------------------------------------------------------------------------

class Gaussian2D(nn.Module):
    """Covariance-parameterised level exactly as gaussianlig.py:39-111: means
    uniform, cov2d ~U(0,1)^3 (may be indefinite), rgbs = 0, Adam + StepLR."""

    def __init__(self, loss_type="L2", **kwargs):
        super().__init__()
        self.loss_type = loss_type
        self.init_num_points = kwargs["num_points"]
        self.H, self.W = kwargs["H"], kwargs["W"]
        self.B_SIZE = 16
        self.device = kwargs["device"]
        self.last_size = (self.H, self.W)
        w_init = torch.rand(self.init_num_points, 1, device=self.device) * self.W
        h_init = torch.rand(self.init_num_points, 1, device=self.device) * self.H
        self.means = nn.Parameter(torch.cat((w_init, h_init), dim=1))
        self.cov2d = nn.Parameter(torch.rand(self.init_num_points, 3, device=self.device))
        self.rgbs = nn.Parameter(torch.zeros(self.init_num_points, 3, device=self.device))
        self.optimizer = torch.optim.Adam([self.rgbs, self.means, self.cov2d], lr=kwargs["lr"])
        self.scheduler = torch.optim.lr_scheduler.StepLR(self.optimizer, step_size=70000,
                                                         gamma=0.7)

    def forward(self):
        xys, extents, conics, num_tiles_hit = g.project_gaussians(self.cov2d, self.means,
                                                                  self.H, self.W, self.B_SIZE)
        out_img = g.rasterize_gaussians(xys, extents, conics, num_tiles_hit, self.rgbs, None,
                                        self.H, self.W, self.B_SIZE)[0][..., :3]
        out_img = torch.clamp(out_img, 0, 1)
        out_img = out_img.view(-1, self.H, self.W, 3).permute(0, 3, 1, 2).contiguous()
        return {"render": out_img}

    def train_iter(self, gt_image):
        image = self.forward()["render"]
        loss = F.mse_loss(image, gt_image)
        loss.backward()
        with torch.no_grad():
            mse_loss = F.mse_loss(image, gt_image)
            p = 10 * math.log10(1.0 / mse_loss.item())
        self.optimizer.step()
        self.optimizer.zero_grad(set_to_none=True)
        self.scheduler.step()
        return loss, p


class LIG(nn.Module):
    """Level split of gaussianlig.py:9-37: level s is H * 2^(s+1-n) and gets
    num_points * 4^(s+1-n) * allo_ratio points; the last level takes the rest."""

    def __init__(self, loss_type="L2", **kwargs):
        super().__init__()
        self.loss_type = loss_type
        self.init_num_points = kwargs["num_points"]
        self.H, self.W = kwargs["H"], kwargs["W"]
        self.n_scales = kwargs["n_scales"]
        self.allo_ratio = kwargs["allo_ratio"]
        self.level_models: List[Gaussian2D] = []
        self.store_min: List[torch.Tensor] = []
        self.store_max: List[torch.Tensor] = []
        for s in range(self.n_scales):
            H = int(self.H * pow(2.0, -self.n_scales + s + 1))
            W = int(self.W * pow(2.0, -self.n_scales + s + 1))
            if s != self.n_scales - 1:
                num_points = int(kwargs["num_points"] * pow(2.0, (-self.n_scales + s + 1) * 2)
                                 * self.allo_ratio)
            else:
                num_points = kwargs["num_points"]
                for i in range(s):
                    num_points -= int(kwargs["num_points"] *
                                      pow(2.0, (-self.n_scales + i + 1) * 2) * self.allo_ratio)
            self.level_models.append(Gaussian2D(loss_type="L2", num_points=num_points, H=H, W=W,
                                                device=kwargs["device"], lr=kwargs["lr"]))


def save_checkpoint(path: str, levels: List[nn.Module], store_min: Optional[list] = None,
                    store_max: Optional[list] = None) -> None:
    """The dict layout of train.py:172-176, {'state_dict': [one state dict per
    level], 'store_max': [...], 'store_min': [...]}.  Each level's entries are
    its own parameters: {means, cov2d, rgbs} for the Gaussian2D mirror (the
    reference's keys, gaussianlig.py:52-58), {means, cholesky, rgbs} for a
    Level; render.TwoLevel.from_checkpoint serves either."""
    torch.save({"state_dict": [m.state_dict() for m in levels],
                "store_max": store_max or [], "store_min": store_min or []}, path)


def load_checkpoint(path: str, levels: List[nn.Module], map_location=None):
    """train.py:51-67: per-level state dicts by key; returns (store_min,
    store_max).  Where train.py keeps a level's own values for keys the
    checkpoint lacks (so a checkpoint of another parameterisation loads as a
    silent no-op), this raises KeyError naming the missing and unexpected
    keys, and ValueError when the level counts differ."""
    ck = torch.load(path, map_location=map_location, weights_only=True)
    sds = ck["state_dict"]
    if len(sds) != len(levels):
        raise ValueError(f"checkpoint has {len(sds)} levels, the model {len(levels)}")
    for i, (level, sd) in enumerate(zip(levels, sds)):
        want = set(level.state_dict())
        missing, unexpected = sorted(want - set(sd)), sorted(set(sd) - want)
        if missing or unexpected:
            raise KeyError(f"level {i}: checkpoint keys do not match the model "
                           f"(missing {missing}, unexpected {unexpected})")
        level.load_state_dict(sd)
    return ck["store_min"], ck["store_max"]

def _main(argv=None):
    """One-level fit for timing: python -m lig_amd.fit [--size 512] [--gaussians
    20000] [--steps 500] [--eager].  Prints one JSON line with the steps/s of
    the timed steps (captured, or eager with --eager) and the final PSNR of
    the clamped render against the seeded target."""
    import argparse
    import json
    import time
    ap = argparse.ArgumentParser(description=_main.__doc__)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--gaussians", type=int, default=20_000)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--lr", type=float, default=1e-2)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--eager", action="store_true", help="eager steps instead of a captured step")
    ap.add_argument("--headroom", type=float, default=1.25,
                    help="captured step: intersection capacity over the count at capture")
    a = ap.parse_args(argv)
    from .data import make_image
    dev = torch.device("cuda:0")
    gt = make_image(a.size, a.size, a.seed, device=dev)
    lv = Level(a.gaussians, a.size, a.size, a.seed + 1, a.scale, a.lr, device=dev)
    if a.eager:
        for _ in range(2):
            lv.step(gt)
    else:
        gph = lv.graph(gt, headroom=a.headroom)
        gph.run(2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if a.eager:
        for _ in range(a.steps):
            lv.step(gt)
    else:
        gph.run(a.steps)
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    with torch.no_grad():
        p = psnr(torch.clamp(lv.render(), 0, 1), gt)
    print(json.dumps({"workload": f"{a.size}x{a.size} x {a.gaussians} Gaussians, scale {a.scale}, "
                                  f"lr {a.lr}, Adam, L2",
                      "mode": "eager" if a.eager else "captured", "steps": a.steps + 2,
                      "timed_steps": a.steps, "steps_s": round(a.steps / secs, 1),
                      "captures": None if a.eager else gph.captures,
                      "psnr_db": round(p, 3)}), flush=True)


if __name__ == "__main__":
    _main()
