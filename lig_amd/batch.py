"""Batched fitting: many images of uneven size fitted together.

All images' Gaussians live in one set of parameters and every stage runs once
for the whole batch: projection, binning into one global tile space, the fused
render + clamp/L2 loss, the deterministic backward and the Adam step.  A
per-image table (lig_abi.h, batched fit) maps each tile and each Gaussian to
its image, and coordinates inside an image are the image's own, so every image
gets the bits it would get fitted alone with lig_amd.fit.Level (the plain
form, one image per launch sequence).  The objective is the sum of the images'
mean clamp + L2 losses, so each image's gradient is that of its own loss.
"""
from __future__ import annotations

import math
from typing import List, Sequence

import numpy as np
import torch
import torch.nn as nn

from . import _lib as L
from . import gsplat2d as g
from .data import init_gaussians, make_image
from .gsplat2d import cuda as _C
from .optim import FusedAdam

_TABLE_DTYPE = np.dtype([("tile_off", "<i4"), ("tiles_x", "<i4"), ("tiles_y", "<i4"),
                         ("w", "<i4"), ("h", "<i4"), ("gauss_off", "<i4"), ("pix_off", "<i8")],
                        align=True)
assert _TABLE_DTYPE.itemsize == 32


def batch_table(sizes: Sequence[tuple], counts: Sequence[int]):
    """(table [B+1] records as a uint8 numpy array, total tiles, total pixels)."""
    tab = np.zeros(len(sizes) + 1, _TABLE_DTYPE)
    t = p = gg = 0
    for i, ((h, w), n) in enumerate(zip(sizes, counts)):
        tx, ty = (w + 15) // 16, (h + 15) // 16
        tab[i] = (t, tx, ty, w, h, gg, p)
        t += tx * ty
        p += h * w
        gg += n
    tab[-1] = (t, 0, 0, 0, 0, gg, p)
    return tab.view(np.uint8), t, p


class _BatchedRenderL2(torch.autograd.Function):
    @staticmethod
    def forward(ctx, means, cholesky, rgbs, target, table, num_images, num_tiles, depth):
        n = means.shape[0]
        dev = means.device
        xys = torch.empty(n, 2, device=dev)
        ext = torch.empty(n, 2, device=dev)
        con = torch.empty(n, 3, device=dev)
        nt = torch.empty(n, dtype=torch.int32, device=dev)
        with torch.cuda.device(dev):
            L.call("lig_project_batched_cholesky", n, L.ptr(cholesky), L.ptr(means), L.ptr(table),
                   num_images, L.ptr(xys), L.ptr(ext), L.ptr(con), L.ptr(nt), L.stream_of(means))
        num_int, cum = g.compute_cumulative_intersects(nt)
        loss_img = torch.empty(num_images, device=dev)
        loss = torch.empty((), device=dev)
        ctx.n, ctx.num_int, ctx.nimg, ctx.ntiles = n, int(num_int), num_images, num_tiles
        if num_int < 1:
            raise RuntimeError("batched fit: no Gaussian intersects any tile")
        gids = torch.empty(num_int, dtype=torch.int32, device=dev)
        bins = torch.empty(num_tiles, 2, dtype=torch.int32, device=dev)
        ws = torch.empty(L.size_query("lig_bin_workspace_size", num_int, num_tiles),
                         dtype=torch.uint8, device=dev)
        img = torch.empty_like(target)
        packed = torch.empty(L.size_query("lig_packed_gaussians_bytes", n), dtype=torch.uint8,
                             device=dev)
        rws = torch.empty(L.size_query("lig_fit_batched_workspace_size", num_tiles),
                          dtype=torch.uint8, device=dev)
        rgbs = rgbs.contiguous()
        with torch.cuda.device(dev):
            s = L.stream_of(means)
            L.call("lig_bin_batched", n, num_int, L.ptr(xys), L.ptr(ext), L.ptr(cum),
                   L.ptr(table), num_images, num_tiles, L.ptr(gids), L.ptr(bins), L.ptr(ws),
                   ws.numel(), s)
            del ws
            L.call("lig_fit_render_l2_batched", n, L.ptr(gids), L.ptr(bins), L.ptr(xys),
                   L.ptr(ext), L.ptr(cum), L.ptr(con), L.ptr(rgbs), L.ptr(packed),
                   L.ptr(target), L.ptr(table), num_images, num_tiles, L.ptr(img),
                   L.ptr(loss_img), L.ptr(loss), int(depth), L.ptr(rws), rws.numel(), s)
        ctx.save_for_backward(cholesky, ext, con, gids, bins, cum, img, target, packed, table)
        ctx.mark_non_differentiable(loss_img)
        return loss, loss_img

    @staticmethod
    def backward(ctx, grad_loss, _grad_img):
        cholesky, ext, con, gids, bins, cum, img, target, packed, table = ctx.saved_tensors
        n, m = ctx.n, ctx.num_int
        dev = img.device
        grad_loss = grad_loss.float().contiguous()
        ws = torch.empty(L.size_query("lig_rasterize_backward_workspace_size", m),
                         dtype=torch.uint8, device=dev)
        v_xy = torch.empty(n, 2, device=dev)
        v_conic = torch.empty(n, 3, device=dev)
        v_rgb = torch.empty(n, 3, device=dev)
        with torch.cuda.device(dev):
            s = L.stream_of(img)
            L.call("lig_fit_backward_records_l2_batched", n, L.ptr(gids), L.ptr(bins),
                   L.ptr(packed), L.ptr(img), L.ptr(target), L.ptr(grad_loss), L.ptr(table),
                   ctx.nimg, ctx.ntiles, m, L.ptr(ws), ws.numel(), s)
            L.call("lig_fit_reduce_records", n, m, L.ptr(cum), L.ptr(ws), ws.numel(),
                   L.ptr(v_xy), L.ptr(v_conic), L.ptr(v_rgb), s)
        v_chol, v_mean = _C.project_gaussians_backward_cholesky(n, ext, cholesky, con, v_xy,
                                                                v_conic)
        return v_mean, v_chol, v_rgb, None, None, None, None, None


class BatchedFit(nn.Module):
    """Images `targets` ([H_i, W_i, 3] each) fitted together; image i gets
    counts[i] Gaussians drawn by data.init_gaussians(counts[i], H_i, W_i,
    seeds[i], scale) — the draws lig_amd.fit.Level makes for it."""

    def __init__(self, targets: List[torch.Tensor], counts: Sequence[int], seeds: Sequence[int],
                 scale: float = 1.0, lr: float = 1e-3):
        super().__init__()
        dev = targets[0].device
        self.sizes = [(int(t.shape[0]), int(t.shape[1])) for t in targets]
        self.counts = [int(c) for c in counts]
        parts = [init_gaussians(c, h, w, s, scale, dev)
                 for (h, w), c, s in zip(self.sizes, self.counts, seeds)]
        self.means = nn.Parameter(torch.cat([p[0] for p in parts]))
        self.cholesky = nn.Parameter(torch.cat([p[1] for p in parts]))
        self.rgbs = nn.Parameter(torch.cat([p[2] for p in parts]))
        tab, self.num_tiles, self.num_pixels = batch_table(self.sizes, self.counts)
        self.table = torch.from_numpy(tab.copy()).to(dev)
        self.target = torch.cat([t.reshape(-1, 3) for t in targets]).contiguous()
        self.optimizer = FusedAdam([self.rgbs, self.means, self.cholesky], lr=lr)
        self.last_losses = None
        self.batch_depth = 0   # LDS batch depth of the forward (0 = default; see autotune)

    def loss(self):
        loss, per_image = _BatchedRenderL2.apply(self.means, self.cholesky, self.rgbs,
                                                 self.target, self.table, len(self.sizes),
                                                 self.num_tiles, self.batch_depth)
        self.last_losses = per_image
        return loss

    def step(self):
        loss = self.loss()
        loss.backward()
        self.optimizer.step()
        self.optimizer.zero_grad(set_to_none=True)
        return loss.detach()

    def psnr(self) -> np.ndarray:
        """Per-image PSNR of the last step's render (clamped, vs the target)."""
        mse = self.last_losses.detach().double().cpu().numpy()
        return 10.0 * np.log10(1.0 / np.maximum(mse, 1e-20))


def batch_workload(count: int = 64, seed: int = 1000, gauss_per_px: float = 0.05,
                   min_hw=(256, 384), max_hw=(2048, 1536)):
    """Image sizes of the batched workload: height and width each log-uniform
    between the smallest and the largest image's (so the aspect ratio is
    log-uniform too), multiples of 8; Gaussians at gauss_per_px; per-image
    seeds seed + i."""
    rng = np.random.default_rng(seed)
    sizes = []
    for _ in range(count):
        h = int(math.exp(rng.uniform(math.log(min_hw[0]), math.log(max_hw[0])))) // 8 * 8
        w = int(math.exp(rng.uniform(math.log(min_hw[1]), math.log(max_hw[1])))) // 8 * 8
        sizes.append((h, w))
    counts = [max(1, int(round(gauss_per_px * h * w))) for h, w in sizes]
    seeds = [seed + i for i in range(count)]
    return sizes, counts, seeds


def make_targets(sizes, seeds, device):
    return [make_image(h, w, s, device=device) for (h, w), s in zip(sizes, seeds)]
