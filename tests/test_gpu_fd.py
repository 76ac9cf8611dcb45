"""The HIP rasteriser backward against FP64 central finite differences.

The oracle's backward is pinned by finite differences on the CPU
(test_oracle_binning_raster.py); this test makes the same check on the GPU
kernels' own output: the deterministic backward of a small scene (plain and
gradient-aware, with opacities, never clamped) against central differences of
a tile-free FP64 forward with the set of contributing pairs held fixed.
Tolerance: |gpu - fd| <= 1e-4 * max(1, |fd|) per checked derivative (the
GPU computes in FP32).  The largest error per output is printed.
"""
import numpy as np
import pytest
import torch

import oracle
from tests.scenes import random_scene
from tests.test_oracle_binning_raster import _loss, _pairs

pytestmark = pytest.mark.gpu


def _t(a, dev, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(dev)


@pytest.mark.parametrize("ga", [False, True])
def test_gpu_backward_matches_finite_differences(device, ga):
    from lig_amd.gsplat2d import cuda as C
    H, W, n = 24, 28, 12
    s = random_scene(n, H, W, 8 + ga, opacities=True, scale=(1.5, 4.0))
    s["opac"] = np.clip(s["opac"], 0.1, 0.9).astype(np.float32)   # never clamped
    xys, ext, con, nt = oracle.project_gaussians_forward(s["par"], s["means"], H, W, 16,
                                                         s["opac"], cholesky=s["cholesky"])
    assert np.all(nt > 0)
    m, cum, isect, gids, ks, gs, bins = oracle.bin_and_sort(xys, ext, nt, H, W, 16)
    dev = device
    gxys, gcon = _t(xys, dev), _t(con, dev)
    gcol, gop = _t(s["colors"], dev), _t(s["opac"], dev)
    gcum = _t(cum, dev, torch.int32)
    g_gs, g_bins, eids = C.bin_and_sort_fast(n, m, gxys, _t(ext, dev), gcum, H, W, 16,
                                             with_emission_ids=True)
    img, wsum, dx, dy, dxy, fin, packed = C.rasterize_forward_packed(
        H, W, 16, g_gs, g_bins, gxys, gcon, gcol, gop, ga)
    rng = np.random.default_rng(3)
    w_img = rng.normal(size=(H, W, 3)).astype(np.float32)
    w_ws = rng.normal(size=(H, W)).astype(np.float32)
    ga_w = [rng.normal(size=(H, W, 3)).astype(np.float32) for _ in range(3)] if ga else None
    got = C.rasterize_backward_packed(H, W, 16, g_gs, g_bins, packed, n, fin, _t(w_img, dev),
                                      _t(w_ws, dev), True,
                                      *([_t(v, dev) for v in ga_w] if ga else [None] * 3),
                                      emission_ids_sorted=eids, cum_tiles_hit=gcum)
    names = ("v_xy", "v_xy_abs", "v_conic", "v_rgb", "v_opacity")
    grads = {k: v.detach().cpu().numpy().astype(np.float64) for k, v in zip(names, got)
             if v is not None}

    mu = xys.astype(np.float64)
    cn = con.astype(np.float64)
    col = s["colors"].astype(np.float64)
    op = s["opac"].astype(np.float64)
    base_pairs = _pairs(mu, cn, op, H, W)
    h = 1e-6
    worst = {}
    checked = 0
    for g in range(n):
        for name, arr, k in (("v_xy", mu, 0), ("v_xy", mu, 1), ("v_conic", cn, 0),
                             ("v_conic", cn, 1), ("v_conic", cn, 2), ("v_rgb", col, 0),
                             ("v_rgb", col, 1), ("v_rgb", col, 2), ("v_opacity", op, None)):
            if name == "v_opacity" and ga:
                continue   # the reference's GA backward drops d(dx, dy, dxy)/d opacity
            plus, minus = arr.copy(), arr.copy()
            if k is None:
                plus[g] += h
                minus[g] -= h
            else:
                plus[g, k] += h
                minus[g, k] -= h
            vals = []
            for a in (plus, minus):
                args = {"v_xy": (a, cn, col, op), "v_conic": (mu, a, col, op),
                        "v_rgb": (mu, cn, a, op), "v_opacity": (mu, cn, col, a)}[name]
                if not np.array_equal(_pairs(args[0], args[1], args[3], H, W), base_pairs):
                    break
                vals.append(_loss(*args, H, W, w_img, w_ws, ga_w))
            if len(vals) < 2:
                continue
            fd = (vals[0] - vals[1]) / (2 * h)
            an = grads[name][g] if k is None else grads[name][g, k]
            err = abs(an - fd) / max(1.0, abs(fd))
            worst[name] = max(worst.get(name, 0.0), err)
            assert err <= 1e-4, (g, name, k, an, fd)
            checked += 1
    print(f"ga={ga}: {checked} derivatives; worst |gpu - fd| / max(1, |fd|): " +
          ", ".join(f"{k} {v:.2e}" for k, v in sorted(worst.items())), flush=True)
    assert checked > 70
