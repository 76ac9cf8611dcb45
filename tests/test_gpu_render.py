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


@pytest.mark.parametrize("sh,sw,fh,fw", [(64, 64, 256, 256), (37, 50, 100, 131),
                                         (128, 96, 256, 192), (50, 40, 20, 17)])
def test_bilinear_resize_matches_torch(device, sh, sw, fh, fw):
    """lig_bilinear_resize against torch's interpolate(mode='bilinear',
    align_corners=False) (train.py:103-105), within FP32 rounding, and every
    window equal to the same window of the full result bit for bit."""
    import torch.nn.functional as F
    from lig_amd.gsplat2d import cuda as C
    g = torch.Generator(device="cpu").manual_seed(sh * 1000 + fw)
    src = torch.rand(sh, sw, 3, generator=g).to(device)
    full = C.bilinear_resize(src, fh, fw)
    ref = F.interpolate(src.permute(2, 0, 1)[None], size=(fh, fw), mode="bilinear",
                        align_corners=False)[0].permute(1, 2, 0)
    err = float((full - ref).abs().max())
    print(f"bilinear {sh}x{sw} -> {fh}x{fw}: max |lig - torch| {err:.2e}", flush=True)
    assert err <= 1e-6
    for x0, y0, w, h in ((0, 0, fw, fh), (3, 5, fw // 2, fh // 3), (fw - 7, fh - 4, 7, 4)):
        win = C.bilinear_resize(src, fh, fw, (x0, y0, w, h))
        assert torch.equal(win, full[y0:y0 + h, x0:x0 + w]), (x0, y0, w, h)


def _reference_checkpoint(path, device, H=256, W=320):
    """Two Gaussian2D mirror levels of the reference's LIG (gaussianlig.py:9-58:
    level 1 at H/2, covariance parameters) with a fitted-looking state, saved
    in train.py's layout (train.py:172-176)."""
    from lig_amd import fit
    torch.manual_seed(5)
    model = fit.LIG(num_points=9000, H=H, W=W, n_scales=2, allo_ratio=0.5, device=device,
                    lr=1e-2)
    with torch.no_grad():
        for m in model.level_models:
            m.cov2d.mul_(6.0)           # a few pixels across; some stay indefinite
            m.rgbs.uniform_(0.0, 0.15)
    # bounds that keep the composition mid-range: out = up(r1) + r2 (r2 (hi - lo) - 0.5 + lo)
    lo, hi = torch.tensor(0.5, device=device), torch.tensor(1.5, device=device)
    fit.save_checkpoint(path, model.level_models, [lo], [hi])
    return model, float(lo), float(hi)


def test_serve_reference_checkpoint(device, tmp_path):
    """A checkpoint in train.py's layout, written from the Gaussian2D mirrors,
    is served: the full image equals train.py's test() composition computed
    through the gsplat2d API and torch (bilinear resize of the clamped level-1
    render plus the de-normalised level-2 render, clamped; train.py:187-196)
    within FP32 rounding; crops at scale 1 equal the full image's pixels bit for
    bit, crops at 0.5 those of the half-size image; the server gives every
    request the bits of rendering it alone."""
    import os
    import torch.nn.functional as F
    from lig_amd.render import TwoLevel
    from lig_amd.serve import CropServer, Request
    H, W = 256, 320
    path = os.path.join(tmp_path, "gaussian_model.pth.tar")
    model, lo, hi = _reference_checkpoint(path, device, H, W)
    rep = TwoLevel.from_checkpoint(path, H, W, device=device)
    assert rep.l1.kind == "cov2d" and (rep.l1.H, rep.l1.W) == (H // 2, W // 2)
    with torch.no_grad():
        r1 = model.level_models[0]()["render"].float()
        r2 = model.level_models[1]()["render"].float()
        want = torch.clamp(F.interpolate(r1, size=(H, W), mode="bilinear")
                           + (r2 * (hi - lo) - 0.5 + lo), 0, 1)[0].permute(1, 2, 0)
        full = rep.render_full()
    err = float((full - want).abs().max())
    print(f"served reference checkpoint vs train.py composition: max err {err:.2e}", flush=True)
    assert err <= 2e-6
    assert float(want.std()) > 0.05
    with torch.no_grad():
        half = rep.render_crop(0.0, 0.0, W // 2, H // 2, 0.5)
        for (x0, y0, w, h) in ((0, 0, 64, 48), (100, 36, 128, 96), (W - 32, H - 16, 32, 16)):
            crop = rep.render_crop(float(x0), float(y0), w, h)
            assert torch.equal(crop, full[y0:y0 + h, x0:x0 + w]), (x0, y0)
            c2 = rep.render_crop(float(x0), float(y0), w // 2, h // 2, 0.5)
            i0, j0 = y0 // 2, x0 // 2
            assert torch.equal(c2, half[i0:i0 + h // 2, j0:j0 + w // 2]), (x0, y0)
    reqs = [Request(0, float(x), float(y), w, h, s) for x, y, w, h, s in
            ((0, 0, 64, 48, 1.0), (100, 36, 128, 96, 1.0), (40, 20, 64, 64, 0.5),
             (200, 100, 96, 80, 1.0), (0, 0, 160, 128, 0.5))]
    alone = [rep.render_crop(r.x0, r.y0, r.w, r.h, r.scale) for r in reqs]
    torch.cuda.synchronize()
    srv = CropServer([rep], slots=4, keep_results=True, device=device)
    srv.run_poisson(reqs, qps=2000.0)
    for r, a in zip(reqs, alone):
        assert torch.equal(r.result, a), (r.x0, r.y0, r.w, r.h, r.scale)


def test_two_level_fit_train_flow(device):
    """two_level_fit(flow='train'): level 1 fits the area downsample, level 2
    the normalised residual of the bilinearly resized clamped level-1 render,
    and the reported PSNR is that of train.py's composition of the two
    renders (train.py:187-196, recomputed here with torch's interpolate)."""
    import math
    import torch.nn.functional as F
    from lig_amd import data, fit
    H = W = 128
    gt = data.make_image(H, W, 2, device=device)
    l1, l2, p = fit.two_level_fit(gt, 800, 2400, steps=30, lr=1e-2, device=device, flow="train")
    with torch.no_grad():
        r1 = torch.clamp(l1.render(), 0, 1).permute(2, 0, 1)[None]
        up = F.interpolate(r1, size=(H, W), mode="bilinear")[0].permute(1, 2, 0)
        t1 = fit.downsample(gt, 4, "area")
        t2, lo, hi = fit.residual_target(gt, up)
        out = torch.clamp(up + (torch.clamp(l2.render(), 0, 1) * (hi - lo) - 0.5 + lo), 0, 1)
    mse = float(((out - gt) ** 2).mean())
    want = 10 * math.log10(1.0 / mse)
    print(f"train flow PSNR {p:.3f} dB, recomputed {want:.3f} dB", flush=True)
    assert p == pytest.approx(want, abs=1e-3)
    assert p > 8.0 and float(t1.mean()) == pytest.approx(float(gt.mean()), abs=1e-5)
