"""Test infrastructure only: FP64 numpy restatement of the image-quality
metrics the reference's evaluation calls, the checker for the HIP kernels in
lig_amd/csrc/ssim.hip.

The reference evaluates a fitted image with
``ms_ssim(out, gt, data_range=1, size_average=True)`` (train.py:200) and offers
single-scale ``ssim`` in its loss menu (utils.py:26-34), both from the
third-party package pytorch-msssim (requirements.txt:5, version unpinned; not
installed here).  This file restates that package's published algorithm
(Wang, Simoncelli & Bovik, "Multiscale structural similarity for image quality
assessment", 2003, as implemented in pytorch_msssim/ssim.py):

* window: 1-D Gaussian of `win_size` taps (odd), coordinates k - win_size//2,
  weights exp(-k^2 / (2 sigma^2)) normalised to sum 1 (default 11 taps,
  sigma 1.5), applied separably along H then W with no padding ("valid"), so
  an H x W plane gives an (H - win_size + 1) x (W - win_size + 1) map;
* per plane (image n, channel c), with C1 = (K1 R)^2, C2 = (K2 R)^2 for the
  data range R (default K = (0.01, 0.03)):
      mu_x = G*x, mu_y = G*y, s_xx = G*(x x) - mu_x^2, s_yy = G*(y y) - mu_y^2,
      s_xy = G*(x y) - mu_x mu_y,
      cs   = (2 s_xy + C2) / (s_xx + s_yy + C2),
      ssim = (2 mu_x mu_y + C1) / (mu_x^2 + mu_y^2 + C1) * cs,
  and the plane's ssim / cs values are the means of the two maps;
* MS-SSIM over L = len(weights) scales (default weights 0.0448, 0.2856,
  0.3001, 0.2363, 0.1333): at scales 0..L-2 keep relu(cs) and halve the
  images with a 2 x 2 average pool of stride 2, padded by (H mod 2, W mod 2)
  zeros on each side and divided by 4 also where the window holds padding;
  at the last scale keep relu(ssim); the plane's value is
  prod_l relu(cs_l)^w_l * relu(ssim_{L-1})^w_{L-1};
* `size_average` averages the per-plane values over images and channels, else
  over channels only; single-scale ssim optionally clamps at 0
  (`nonnegative_ssim`) before averaging.

Pinned in tests/test_oracle_ssim.py by: the separable filter against
scipy.signal.correlate2d with the outer-product window; the per-pixel maps
against a brute-force weighted sum over each 11 x 11 window; closed forms for
constant images (cs = 1, ssim = (2ab + C1)/(a^2 + b^2 + C1), and MS-SSIM =
that ssim to the last weight when every pool is unpadded); identity (= 1);
the pool against torch.nn.functional.avg_pool2d; and one-hot weights, which
single out each scale's cs.
"""
from __future__ import annotations

import numpy as np

MS_WEIGHTS = (0.0448, 0.2856, 0.3001, 0.2363, 0.1333)


def gauss_window(size: int = 11, sigma: float = 1.5) -> np.ndarray:
    """pytorch_msssim _fspecial_gauss_1d: normalised 1-D Gaussian taps."""
    if size % 2 != 1:
        raise ValueError("window size must be odd")
    k = np.arange(size, dtype=np.float64) - size // 2
    g = np.exp(-(k ** 2) / (2.0 * sigma ** 2))
    return g / g.sum()


def filter_valid(img: np.ndarray, g: np.ndarray) -> np.ndarray:
    """Separable valid correlation of the last two axes with g (H, then W)."""
    img = np.asarray(img, np.float64)
    n = len(g)
    H, W = img.shape[-2:]
    if H < n or W < n:
        raise ValueError("plane smaller than the window")
    t = sum(g[k] * img[..., k:H - n + 1 + k, :] for k in range(n))
    return sum(g[k] * t[..., :, k:W - n + 1 + k] for k in range(n))


def ssim_maps(x, y, data_range=1.0, g=None, K=(0.01, 0.03)):
    """(ssim_map, cs_map) of planes [..., H, W] (pytorch_msssim _ssim)."""
    g = gauss_window() if g is None else np.asarray(g, np.float64)
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    C1 = (K[0] * data_range) ** 2
    C2 = (K[1] * data_range) ** 2
    mx, my = filter_valid(x, g), filter_valid(y, g)
    sxx = filter_valid(x * x, g) - mx * mx
    syy = filter_valid(y * y, g) - my * my
    sxy = filter_valid(x * y, g) - mx * my
    cs = (2 * sxy + C2) / (sxx + syy + C2)
    ss = (2 * mx * my + C1) / (mx * mx + my * my + C1) * cs
    return ss, cs


def ssim_cs(x, y, data_range=1.0, g=None, K=(0.01, 0.03)):
    """Per-plane means (ssim [..], cs [..]) over the last two axes."""
    ss, cs = ssim_maps(x, y, data_range, g, K)
    return ss.mean(axis=(-2, -1)), cs.mean(axis=(-2, -1))


def avg_pool_pad(x) -> np.ndarray:
    """2 x 2 average pool, stride 2, zero padding (H mod 2, W mod 2) on each
    side, always divided by 4."""
    x = np.asarray(x, np.float64)
    H, W = x.shape[-2:]
    ph, pw = H % 2, W % 2
    pad = [(0, 0)] * (x.ndim - 2) + [(ph, ph), (pw, pw)]
    xp = np.pad(x, pad)
    Ho, Wo = (H + 2 * ph - 2) // 2 + 1, (W + 2 * pw - 2) // 2 + 1
    xp = xp[..., :2 * Ho, :2 * Wo]
    return 0.25 * (xp[..., 0::2, 0::2] + xp[..., 1::2, 0::2] + xp[..., 0::2, 1::2] +
                   xp[..., 1::2, 1::2])


def ms_ssim_planes(x, y, data_range=1.0, g=None, weights=MS_WEIGHTS, K=(0.01, 0.03)):
    """Per-plane MS-SSIM values [..] of planes [..., H, W]."""
    w = np.asarray(weights, np.float64)
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    terms = []
    for lvl in range(len(w)):
        ss, cs = ssim_cs(x, y, data_range, g, K)
        if lvl < len(w) - 1:
            terms.append(np.maximum(cs, 0.0))
            x, y = avg_pool_pad(x), avg_pool_pad(y)
        else:
            terms.append(np.maximum(ss, 0.0))
    out = np.ones_like(terms[0])
    for t, wl in zip(terms, w):
        out = out * t ** wl
    return out


def ms_ssim(x, y, data_range=1.0, size_average=True, win_size=11, win_sigma=1.5,
            weights=MS_WEIGHTS, K=(0.01, 0.03)):
    """x, y: [N, C, H, W]; the scalar mean, or [N] when size_average is False."""
    v = ms_ssim_planes(x, y, data_range, gauss_window(win_size, win_sigma), weights, K)
    return v.mean() if size_average else v.mean(axis=1)


def ssim(x, y, data_range=1.0, size_average=True, win_size=11, win_sigma=1.5,
         K=(0.01, 0.03), nonnegative_ssim=False):
    """Single-scale SSIM of [N, C, H, W] images."""
    ss, _ = ssim_cs(x, y, data_range, gauss_window(win_size, win_sigma), K)
    if nonnegative_ssim:
        ss = np.maximum(ss, 0.0)
    return ss.mean() if size_average else ss.mean(axis=1)
