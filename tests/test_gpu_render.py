"""GPU checks of crop-native rendering (lig_project_gaussians_view, the
image-only forward, lig_amd.render).

* view projection: bit-exact against the oracle's projection of parameters
  transformed on the host in FP32 with the same operations;
* a window at scale 1 whose origin makes mean - origin exact renders the
  same bits as the matching crop of the full render (same pixel offsets,
  same bins per pixel, same order);
* a scale-0.5 view of the whole canvas renders the same bits as the
  reference's scale_to (means and Cholesky factors times 0.5) rendered
  through the API;
* two-level crops match the composed full image within 1e-5 (the upscaler's
  ROI grid rounds differently from the full grid), at scale 1 and 0.5.
"""
import numpy as np
import pytest
import torch

from tests.scenes import random_scene

pytestmark = pytest.mark.gpu


def _t(a, dev, dtype=torch.float32):
    return None if a is None else torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(dev)


@pytest.mark.parametrize("chol", [True, False])
def test_view_projection_bit_exact(device, oracle_lib, chol):
    from lig_amd.gsplat2d import cuda as C
    n, H, W = 3000, 300, 400
    s = random_scene(n, H, W, 8, cholesky=chol, margin=30.0)
    ox, oy, sx, sy = 37.25, 81.5, 0.5, 0.75
    f32 = np.float32
    m = s["means"].astype(f32)
    par = s["par"].astype(f32)
    vm = np.stack([(m[:, 0] - f32(ox)) * f32(sx), (m[:, 1] - f32(oy)) * f32(sy)], 1)
    if chol:
        vp = np.stack([par[:, 0] * f32(sx), par[:, 1] * f32(sy), par[:, 2] * f32(sy)], 1)
    else:
        vp = np.stack([(par[:, 0] * f32(sx)) * f32(sx), (par[:, 1] * f32(sx)) * f32(sy),
                       (par[:, 2] * f32(sy)) * f32(sy)], 1)
    h, w = 150, 180
    out = C.project_gaussians_view(n, _t(par, device), _t(m, device), None, chol, (ox, oy),
                                   (sx, sy), h, w, 16)
    ref = oracle_lib.project_gaussians_forward(vp, vm, h, w, 16, None, cholesky=chol)
    for name, a, b in zip(("xys", "extents", "conics", "num_tiles_hit"), out, ref):
        np.testing.assert_array_equal(a.cpu().numpy(), b, err_msg=name)


def _level(n, H, W, seed, scale, device):
    from lig_amd import data
    from lig_amd.render import LevelParams
    means, chol, rgbs = data.init_gaussians(n, H, W, seed, scale, device)
    return LevelParams(means, chol, rgbs * 0.1, H, W)


def _api_render(p, H, W, means=None, chol=None):
    import lig_amd.gsplat2d as g
    means = p.means if means is None else means
    chol = p.cholesky if chol is None else chol
    xys, ext, con, nt = g.project_gaussians_cholesky(chol, means, H, W, 16)
    return g.rasterize_gaussians(xys, ext, con, nt, p.rgbs, None, H, W, 16)[0]


def test_window_matches_full_render(device):
    from lig_amd.render import render_view
    H, W = 512, 640
    p = _level(20000, H, W, 3, 1.5, device)
    with torch.no_grad():
        full = _api_render(p, H, W)
        assert torch.equal(render_view(p, 0.0, 0.0, W, H), full)
        for x0, y0, w, h in ((256, 192, 256, 160), (320, 256, 320, 256)):
            win = render_view(p, float(x0), float(y0), w, h)
            assert torch.equal(win, full[y0:y0 + h, x0:x0 + w]), (x0, y0)


def test_half_scale_view_is_scale_to(device):
    from lig_amd.render import render_view
    H, W = 512, 384
    p = _level(12000, H, W, 4, 2.0, device)
    with torch.no_grad():
        view = render_view(p, 0.0, 0.0, W // 2, H // 2, 0.5)
        ref = _api_render(p, H // 2, W // 2, means=p.means * 0.5, chol=p.cholesky * 0.5)
    assert torch.equal(view, ref)


@pytest.mark.parametrize("scale", [1.0, 0.5])
def test_two_level_crop_matches_full(device, scale):
    import lig_amd.gsplat2d as g
    from lig_amd.fit import compose
    from lig_amd.render import TwoLevel, render_view, render_view_ga
    H, W = 512, 512
    rep = TwoLevel(_level(3000, H // 4, W // 4, 5, 1.0, device),
                   _level(20000, H, W, 6, 1.0, device), lo=-0.2, hi=1.3)
    with torch.no_grad():
        if scale == 1.0:
            full = rep.render_full()
        else:
            img1, dx, dy, dxy = render_view_ga(rep.l1, 0.0, 0.0, W // 4, H // 4)
            up = torch.clamp(g.gradient_aware_upscale(img1, dx, dy, dxy, H // 2, W // 2), 0, 1)
            img2 = torch.clamp(render_view(rep.l2, 0.0, 0.0, W // 2, H // 2, 0.5), 0, 1)
            full = compose(img2, up, rep.lo, rep.hi)
        wins = ((0, 0, 128, 128), (96, 200, 256, 128), (384, 384, 64, 64)) if scale == 1.0 \
            else ((0, 0, 128, 128), (96, 200, 128, 64), (384, 384, 64, 64))
        for x0, y0, w, h in wins:
            i0, j0 = int(y0 * scale), int(x0 * scale)
            want = full[i0:i0 + h, j0:j0 + w]
            # resident level-1 render (default) and the per-request level-1
            # source region
            for cache in (True, False):
                crop = rep.render_crop(float(x0), float(y0), w, h, scale, level1_cache=cache)
                assert crop.shape == want.shape
                err = float((crop - want).abs().max())
                print(f"two-level crop scale={scale} cache={cache} at ({x0},{y0}) {w}x{h}: "
                      f"max err {err:.2e}")
                assert err <= 1e-5
                if cache:
                    assert torch.equal(crop, want)
