"""Seeded synthetic target images for the fit workloads.

Every random draw comes from numpy's PCG64 generator with the given seed;
fields are evaluated with torch on the requested device (CPU or HIP), so an
image depends only on (seed, H, W, options) up to FP32 rounding of the device.

Image recipe (RGB, FP32 in [0, 1]):
  * 64 anisotropic Gaussian blobs: centre uniform, per-axis sigma
    U(0.01, 0.08) * min(H, W), angle U[0, pi), RGB amplitude U(-1, 1);
  * 1/f band-limited noise: white N(0,1) per channel, spectrum scaled by
    1 / max(f, 1/min(H, W)), zeroed above 0.25 cycles/px, scaled to std 0.15;
  * 8 hard linear edges: random point and normal, RGB step U(-0.5, 0.5);
  * optional cell texture: 5000 ellipses (axes U(4, 24) px, colour U(0,1),
    50 % blend) drawn in a 1024 x 1024 cell that is tiled over the image;
  * per-image min-max normalisation to [0, 1].
"""
from __future__ import annotations

import math

import numpy as np
import torch


def _blobs(rng, H, W, dev, out):
    m = min(H, W)
    ys = torch.arange(H, device=dev, dtype=torch.float32) + 0.5
    xs = torch.arange(W, device=dev, dtype=torch.float32) + 0.5
    for _ in range(64):
        cx, cy = rng.uniform(0, W), rng.uniform(0, H)
        su, sv = rng.uniform(0.01, 0.08, 2) * m
        ang = rng.uniform(0, math.pi)
        amp = torch.tensor(rng.uniform(-1, 1, 3), device=dev, dtype=torch.float32)
        r = 4.0 * max(su, sv)
        x0, x1 = max(0, int(cx - r)), min(W, int(cx + r) + 1)
        y0, y1 = max(0, int(cy - r)), min(H, int(cy + r) + 1)
        if x1 <= x0 or y1 <= y0:
            continue
        dx = (xs[x0:x1] - cx)[None, :]
        dy = (ys[y0:y1] - cy)[:, None]
        c, s = math.cos(ang), math.sin(ang)
        u = c * dx + s * dy
        v = -s * dx + c * dy
        g = torch.exp(-0.5 * (u * u / (su * su) + v * v / (sv * sv)))
        out[y0:y1, x0:x1, :] += g[..., None] * amp


def _noise(rng, H, W, dev, out):
    white = torch.tensor(rng.standard_normal((3, H, W), dtype=np.float32), device=dev)
    spec = torch.fft.rfft2(white)
    fy = torch.fft.fftfreq(H, device=dev).abs()[:, None]
    fx = torch.fft.rfftfreq(W, device=dev)[None, :]
    f = torch.sqrt(fx * fx + fy * fy)
    scale = 1.0 / torch.clamp(f, min=1.0 / min(H, W))
    scale = torch.where(f > 0.25, torch.zeros_like(scale), scale)
    noise = torch.fft.irfft2(spec * scale, s=(H, W))
    noise = noise / noise.std() * 0.15
    out += noise.permute(1, 2, 0)


def _edges(rng, H, W, dev, out):
    ys = torch.arange(H, device=dev, dtype=torch.float32)[:, None] + 0.5
    xs = torch.arange(W, device=dev, dtype=torch.float32)[None, :] + 0.5
    for _ in range(8):
        px, py = rng.uniform(0, W), rng.uniform(0, H)
        ang = rng.uniform(0, 2 * math.pi)
        step = torch.tensor(rng.uniform(-0.5, 0.5, 3), device=dev, dtype=torch.float32)
        side = ((xs - px) * math.cos(ang) + (ys - py) * math.sin(ang)) > 0
        out += side[..., None].float() * step


def _cells(rng, H, W, dev, out, cell=1024, count=5000):
    """Ellipse texture drawn once in a cell x cell numpy canvas (CPU), then
    tiled over the image and blended 50 % on the device."""
    tile = np.zeros((cell, cell, 3), np.float32)
    alpha = np.zeros((cell, cell, 1), np.float32)
    params = rng.uniform(size=(count, 8))
    for k in range(count):
        cx, cy = params[k, 0] * cell, params[k, 1] * cell
        a, b = 4 + 20 * params[k, 2], 4 + 20 * params[k, 3]
        ang = params[k, 4] * math.pi
        r = int(max(a, b)) + 1
        x0, x1 = max(0, int(cx) - r), min(cell, int(cx) + r + 1)
        y0, y1 = max(0, int(cy) - r), min(cell, int(cy) + r + 1)
        dx = (np.arange(x0, x1, dtype=np.float32) + 0.5 - cx)[None, :]
        dy = (np.arange(y0, y1, dtype=np.float32) + 0.5 - cy)[:, None]
        c, s = math.cos(ang), math.sin(ang)
        u = (c * dx + s * dy) / a
        v = (-s * dx + c * dy) / b
        inside = (u * u + v * v) <= 1.0
        tile[y0:y1, x0:x1][inside] = params[k, 5:8].astype(np.float32)
        alpha[y0:y1, x0:x1][inside] = 1.0
    reps_y, reps_x = -(-H // cell), -(-W // cell)
    big = torch.tensor(tile, device=dev).repeat(reps_y, reps_x, 1)[:H, :W]
    mask = torch.tensor(alpha, device=dev).repeat(reps_y, reps_x, 1)[:H, :W]
    out.mul_(1 - 0.5 * mask).add_(0.5 * mask * big)


def make_image(H: int, W: int, seed: int, cells: bool = False,
               device="cpu") -> torch.Tensor:
    """Seeded target image, HWC float32 in [0, 1]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    dev = torch.device(device)
    img = torch.zeros(H, W, 3, device=dev, dtype=torch.float32)
    _blobs(rng, H, W, dev, img)
    _noise(rng, H, W, dev, img)
    _edges(rng, H, W, dev, img)
    if cells:
        _cells(rng, H, W, dev, img)
    lo, hi = img.min(), img.max()
    return ((img - lo) / (hi - lo)).contiguous()


def init_gaussians(n: int, H: int, W: int, seed: int, scale: float = 1.0, device="cpu"):
    """Fit initialisation: means uniform over the image, Cholesky diagonal
    ~U(0.5, 3) * scale px, off-diagonal 0, RGB ~U(0, 1).  Returns
    (means [n,2], cholesky [n,3], rgbs [n,3]) float32 tensors."""
    rng = np.random.Generator(np.random.PCG64(seed))
    means = np.stack([rng.uniform(0, W, n), rng.uniform(0, H, n)], 1)
    l11 = rng.uniform(0.5, 3.0, n) * scale
    l22 = rng.uniform(0.5, 3.0, n) * scale
    chol = np.stack([l11, np.zeros(n), l22], 1)
    rgbs = rng.uniform(0, 1, (n, 3))
    dev = torch.device(device)
    f = lambda a: torch.tensor(a, dtype=torch.float32, device=dev)  # noqa: E731
    return f(means), f(chol), f(rgbs)
