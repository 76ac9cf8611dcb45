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

import ctypes
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
    def forward(ctx, means, cholesky, rgbs, target, table, num_images, num_tiles, depth,
                persistent=False):
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
        ctx.persistent = bool(persistent)
        sfx = "_persistent" if persistent else ""
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
            L.call("lig_fit_render_l2_batched" + sfx, n, L.ptr(gids), L.ptr(bins), L.ptr(xys),
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
            sfx = "_persistent" if ctx.persistent else ""
            L.call("lig_fit_backward_records_l2_batched" + sfx, n, L.ptr(gids), L.ptr(bins),
                   L.ptr(packed), L.ptr(img), L.ptr(target), L.ptr(grad_loss), L.ptr(table),
                   ctx.nimg, ctx.ntiles, m, L.ptr(ws), ws.numel(), s)
            L.call("lig_fit_reduce_records", n, m, L.ptr(cum), L.ptr(ws), ws.numel(),
                   L.ptr(v_xy), L.ptr(v_conic), L.ptr(v_rgb), s)
        v_chol, v_mean = _C.project_gaussians_backward_cholesky(n, ext, cholesky, con, v_xy,
                                                                v_conic)
        return v_mean, v_chol, v_rgb, None, None, None, None, None, None


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
        # raster stages as persistent grids fed by the per-XCD tile queue
        # (lig_fit_render_l2_batched_persistent); same bits either way
        self.persistent = False

    def loss(self):
        loss, per_image = _BatchedRenderL2.apply(self.means, self.cholesky, self.rgbs,
                                                 self.target, self.table, len(self.sizes),
                                                 self.num_tiles, self.batch_depth,
                                                 self.persistent)
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


class LevelBatch:
    """A single-image fit (lig_amd.fit.Level and its target) seen as a batch of
    one image: it shares the Level's parameter tensors and optimiser, so a
    BatchedStepGraph over it steps the Level itself.  A batch of one computes
    exactly what the Level's own fused step computes."""

    def __init__(self, level, target_hwc: torch.Tensor):
        dev = level.means.device
        if tuple(target_hwc.shape) != (level.H, level.W, 3):
            raise ValueError(f"target {tuple(target_hwc.shape)} does not match the level "
                             f"({level.H}, {level.W}, 3)")
        self.means, self.cholesky, self.rgbs = level.means, level.cholesky, level.rgbs
        self.optimizer = level.optimizer
        self.sizes = [(level.H, level.W)]
        self.counts = [int(level.means.shape[0])]
        tab, self.num_tiles, self.num_pixels = batch_table(self.sizes, self.counts)
        self.table = torch.from_numpy(tab.copy()).to(dev)
        self.target = target_hwc.reshape(-1, 3).contiguous()
        self.batch_depth = level.batch_depth
        self.persistent = False
        self.last_losses = None


class BatchedStepGraph:
    """BatchedFit's fit step recorded once as a hipGraph and replayed.

    Every buffer of the step is allocated up front: intersection buffers for
    `capacity` intersections (the count measured at set-up times `headroom`,
    unless given), and no stage reads the intersection count on the host.  The
    graph holds projection, scan, lig_fit_step_begin, binning with the count
    read on the device, the fused render + loss, the backward records and their
    reduction, the projection backward and one device-step Adam update per
    parameter tensor.  A step whose count exceeds the capacity changes nothing
    (empty bins, no update) and is counted as skipped; `run` then grows the
    capacity to the largest count seen times `headroom`, records the graph
    again and takes the remaining steps, so the parameters after `run(k)` are
    bit-identical to k calls of BatchedFit.step.

    The optimiser's lr, betas, eps and weight decay, and the model's
    batch_depth and persistent settings, are recorded in the graph: build a new
    BatchedStepGraph after changing them.  Eager BatchedFit.step calls may be
    interleaved with `run` (the step counter is kept in sync).
    """

    def __init__(self, bf: BatchedFit, capacity: int = None, headroom: float = 1.25,
                 table_steps: int = 4096):
        self.bf = bf
        self.headroom = float(headroom)
        self.table_steps = int(table_steps)
        self.dev = bf.means.device
        opt = bf.optimizer
        if len(opt.param_groups) != 1:
            raise ValueError("one parameter group expected")
        grp = opt.param_groups[0]
        self.lr, (self.b1, self.b2) = grp["lr"], grp["betas"]
        self.eps, self.wd = grp["eps"], grp["weight_decay"]
        self.params = [bf.rgbs, bf.means, bf.cholesky]
        for p in self.params:
            st = opt.state[p]
            if not st:   # the state FusedAdam creates at its first step
                st["step"] = 0
                st["exp_avg"] = torch.zeros_like(p, memory_format=torch.contiguous_format)
                st["exp_avg_sq"] = torch.zeros_like(p, memory_format=torch.contiguous_format)
        if len({int(opt.state[p]["step"]) for p in self.params}) != 1:
            raise ValueError("the parameters' optimiser steps differ")
        n, dev = bf.means.shape[0], self.dev
        self.n = n
        f32 = dict(device=dev, dtype=torch.float32)
        self.xys = torch.empty(n, 2, **f32)
        self.ext = torch.empty(n, 2, **f32)
        self.con = torch.empty(n, 3, **f32)
        self.nt = torch.empty(n, dtype=torch.int32, device=dev)
        self.cum = torch.empty(n, dtype=torch.int32, device=dev)
        self.scan_ws = torch.empty(max(L.size_query("lig_scan_workspace_size", n), 1),
                                   dtype=torch.uint8, device=dev)
        self.bins = torch.empty(bf.num_tiles, 2, dtype=torch.int32, device=dev)
        self.img = torch.empty_like(bf.target)
        self.packed = torch.empty(L.size_query("lig_packed_gaussians_bytes", n),
                                  dtype=torch.uint8, device=dev)
        self.rws = torch.empty(L.size_query("lig_fit_batched_workspace_size", bf.num_tiles),
                               dtype=torch.uint8, device=dev)
        self.loss_img = torch.empty(len(bf.sizes), **f32)
        self.loss = torch.empty((), **f32)
        self.grad_loss = torch.ones((), **f32)
        self.v_xy = torch.empty(n, 2, **f32)
        self.v_conic = torch.empty(n, 3, **f32)
        self.v_rgb = torch.empty(n, 3, **f32)
        self.v_chol = torch.empty(n, 3, **f32)
        self.v_mean = torch.empty(n, 2, **f32)
        self.grads = [self.v_rgb, self.v_mean, self.v_chol]
        self.ctl = torch.zeros(4, dtype=torch.int32, device=dev)
        self.captures = 0
        step = self._host_step()
        self._make_table(step + self.table_steps)
        if capacity is None:
            with torch.cuda.device(dev):
                self._project_and_scan(L.stream_of(bf.means))
            capacity = max(1, math.ceil(int(self.cum[-1].item()) * self.headroom))
        self._alloc(int(capacity))
        self._capture()

    # -- set-up --------------------------------------------------------------
    def _host_step(self) -> int:
        return int(self.bf.optimizer.state[self.params[0]]["step"])

    def _make_table(self, steps: int):
        host = np.empty(2 * steps, np.float32)
        L.call("lig_adam_bias_corrections", float(self.b1), float(self.b2), 1, steps,
               host.ctypes.data_as(ctypes.c_void_p))
        self.table = torch.from_numpy(host).to(self.dev)
        self.table_len = steps

    def _alloc(self, capacity: int):
        dev, ntiles = self.dev, self.bf.num_tiles
        self.capacity = capacity
        self.gids = torch.empty(capacity, dtype=torch.int32, device=dev)
        self.bin_ws = torch.empty(L.size_query("lig_bin_workspace_size", capacity, ntiles),
                                  dtype=torch.uint8, device=dev)
        self.rec_ws = torch.empty(L.size_query("lig_rasterize_backward_workspace_size", capacity),
                                  dtype=torch.uint8, device=dev)

    def _project_and_scan(self, s):
        bf, n = self.bf, self.n
        L.call("lig_project_batched_cholesky", n, L.ptr(bf.cholesky), L.ptr(bf.means),
               L.ptr(bf.table), len(bf.sizes), L.ptr(self.xys), L.ptr(self.ext),
               L.ptr(self.con), L.ptr(self.nt), s)
        L.call("lig_inclusive_scan_i32", n, L.ptr(self.nt), L.ptr(self.cum),
               L.ptr(self.scan_ws), self.scan_ws.numel(), s)

    def _launch(self):
        bf, n, cap = self.bf, self.n, self.capacity
        nimg, ntiles = len(bf.sizes), bf.num_tiles
        s = L.stream_of(bf.means)
        self._project_and_scan(s)
        count = ctypes.c_void_p(self.cum.data_ptr() + 4 * (n - 1))
        L.call("lig_fit_step_begin", count, cap, self.table_len, L.ptr(self.ctl), s)
        L.call("lig_bin_batched_capacity", n, cap, L.ptr(self.xys), L.ptr(self.ext),
               L.ptr(self.cum), L.ptr(bf.table), nimg, ntiles, L.ptr(self.gids),
               L.ptr(self.bins), L.ptr(self.bin_ws), self.bin_ws.numel(), s)
        sfx = "_persistent" if bf.persistent else ""
        L.call("lig_fit_render_l2_batched" + sfx, n, L.ptr(self.gids), L.ptr(self.bins),
               L.ptr(self.xys), L.ptr(self.ext), L.ptr(self.cum), L.ptr(self.con),
               L.ptr(bf.rgbs), L.ptr(self.packed), L.ptr(bf.target), L.ptr(bf.table), nimg,
               ntiles, L.ptr(self.img), L.ptr(self.loss_img), L.ptr(self.loss),
               int(bf.batch_depth), L.ptr(self.rws), self.rws.numel(), s)
        L.call("lig_fit_backward_records_l2_batched" + sfx, n, L.ptr(self.gids),
               L.ptr(self.bins), L.ptr(self.packed), L.ptr(self.img), L.ptr(bf.target),
               L.ptr(self.grad_loss),
               L.ptr(bf.table), nimg, ntiles, cap, L.ptr(self.rec_ws), self.rec_ws.numel(), s)
        L.call("lig_fit_reduce_records", n, cap, L.ptr(self.cum), L.ptr(self.rec_ws),
               self.rec_ws.numel(), L.ptr(self.v_xy), L.ptr(self.v_conic), L.ptr(self.v_rgb), s)
        L.call("lig_project_gaussians_backward_cholesky", n, L.ptr(self.ext),
               L.ptr(bf.cholesky), L.ptr(self.con), L.ptr(self.v_xy), L.ptr(self.v_conic),
               L.ptr(self.v_chol), L.ptr(self.v_mean), s)
        opt = bf.optimizer
        for p, g in zip(self.params, self.grads):
            st = opt.state[p]
            L.call("lig_adam_step_device", p.numel(), L.ptr(p), L.ptr(g), L.ptr(st["exp_avg"]),
                   L.ptr(st["exp_avg_sq"]), self.lr, self.b1, self.b2, self.eps, self.wd,
                   L.ptr(self.ctl), L.ptr(self.table), s)

    def _capture(self):
        for name, t in (("means", self.bf.means), ("cholesky", self.bf.cholesky),
                        ("rgbs", self.bf.rgbs)):
            L.check_device(name, t, torch.float32)
        torch.cuda.synchronize(self.dev)
        self.graph = torch.cuda.CUDAGraph()
        stream = torch.cuda.Stream(device=self.dev)   # default priority
        with torch.no_grad(), torch.cuda.device(self.dev):
            with torch.cuda.graph(self.graph, stream=stream):
                self._launch()
        self.captures += 1

    # -- stepping ------------------------------------------------------------
    def _read_ctl(self):
        return [int(v) for v in self.ctl.cpu()]

    def run(self, steps: int) -> torch.Tensor:
        """Takes `steps` fit steps by replaying the graph (one host
        synchronisation at the end, more only when a step was skipped);
        returns the last step's loss."""
        if steps < 1:
            raise ValueError("steps must be >= 1")
        ctl = self._read_ctl()
        host_step = self._host_step()
        if ctl[0] != host_step:          # eager steps were taken in between
            self.ctl[0] = host_step
        remaining = steps
        while remaining > 0:
            before = self._host_step()
            for _ in range(remaining):
                self.graph.replay()
            ctl = self._read_ctl()
            taken = ctl[0] - before
            for p in self.params:
                self.bf.optimizer.state[p]["step"] = ctl[0]
            remaining -= taken
            if remaining == 0:
                break
            if ctl[2] <= self.capacity and ctl[0] < self.table_len:
                raise RuntimeError(f"captured fit step skipped for no known reason: {ctl}")
            if ctl[2] > self.capacity:
                self._alloc(max(ctl[2], math.ceil(ctl[2] * self.headroom)))
            if ctl[0] >= self.table_len:
                self._make_table(ctl[0] + self.table_steps)
            self._capture()
        self.bf.last_losses = self.loss_img.clone()
        return self.loss.clone()

    @property
    def skipped(self) -> int:
        return self._read_ctl()[1]


def _main(argv=None):
    """Steps the 64-image batched workload (batch_workload) in one launch form,
    for timing and profiling: python -m lig_amd.batch [--skew-factor F]
    [--persistent 0|1] [--steps K]."""
    import argparse
    import json
    import time
    ap = argparse.ArgumentParser(description=_main.__doc__)
    ap.add_argument("--images", type=int, default=64)
    ap.add_argument("--skew-factor", type=int, default=1,
                    help="Gaussian density multiplier of every second one of the first "
                         "--skew-images images")
    ap.add_argument("--skew-images", type=int, default=16)
    ap.add_argument("--persistent", type=int, default=0)
    ap.add_argument("--depth", type=int, default=128)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    a = ap.parse_args(argv)
    dev = torch.device("cuda:0")
    sizes, counts, seeds = batch_workload(a.images, seed=1000, gauss_per_px=0.05)
    counts = [c * a.skew_factor if i < a.skew_images and i % 2 == 0 else c
              for i, c in enumerate(counts)]
    bf = BatchedFit(make_targets(sizes, seeds, dev), counts, seeds, scale=2.0)
    bf.batch_depth, bf.persistent = a.depth, bool(a.persistent)
    for _ in range(a.warmup):
        bf.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        bf.step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / a.steps
    print(json.dumps({"images": a.images, "gaussians": int(sum(counts)),
                      "persistent": bf.persistent, "depth": bf.batch_depth,
                      "ms_per_step": round(dt * 1e3, 3), "steps_s": round(1 / dt, 3)}))


if __name__ == "__main__":
    _main()
