"""CPU tests of the caller layer's host-side semantics (no GPU calls):
train.py's residual normalisation and composition (train.py:113-135, 150-154,
187-196), its area downsample (train.py:95-98), and the checkpoint format
(train.py:51-67, 172-176) with the strict loader and the serving loader."""
import os

import numpy as np
import pytest
import torch
import torch.nn.functional as F

from lig_amd import fit


def _images(seed, H=24, W=20):
    g = torch.Generator().manual_seed(seed)
    gt = torch.rand(H, W, 3, generator=g)
    prev = torch.clamp(gt + 0.3 * torch.randn(H, W, 3, generator=g), 0, 1)
    return gt, prev


def test_residual_target_closed_form():
    """target = (gt - prev + 0.5 - min) / (max - min), with min / max over the
    whole residual + 0.5 (train.py:113-135): spans exactly [0, 1]."""
    gt, prev = _images(0)
    t, lo, hi = fit.residual_target(gt, prev)
    r = gt.double() - prev.double() + 0.5
    assert float(lo) == pytest.approx(float(r.min()), abs=1e-6)
    assert float(hi) == pytest.approx(float(r.max()), abs=1e-6)
    assert float(t.min()) == 0.0 and float(t.max()) == pytest.approx(1.0, abs=1e-6)
    np.testing.assert_allclose(t.numpy(), ((r - r.min()) / (r.max() - r.min())).numpy(),
                               atol=1e-6)


def test_compose_inverts_the_residual_normalisation():
    """A perfect level-2 fit (its clamped render equal to the normalised target)
    composes back to the ground truth: compose(residual_target(gt, prev)) = gt;
    and the test-time order of train.py:193-196 (prev + (r (max - min) - 0.5 +
    min), then clamp) agrees with the training-time order compose uses
    (train.py:150-154) to FP32 rounding."""
    for seed in range(3):
        gt, prev = _images(seed)
        t, lo, hi = fit.residual_target(gt, prev)
        out = fit.compose(torch.clamp(t, 0, 1), prev, lo, hi)
        np.testing.assert_allclose(out.numpy(), gt.numpy(), atol=2e-6)
        r2 = torch.rand_like(t)
        test_order = torch.clamp(prev + (r2 * (hi - lo) - 0.5 + lo), 0, 1)
        np.testing.assert_allclose(fit.compose(r2, prev, lo, hi).numpy(), test_order.numpy(),
                                   atol=1e-6)


def test_area_downsample_is_block_mean():
    """train.py's level target, interpolate(scale_factor=1/f, mode='area'), is
    the mean over each f x f block."""
    gt, _ = _images(4, 32, 48)
    for f in (2, 4):
        got = fit.downsample(gt, f, "area").numpy()
        want = gt.numpy().reshape(32 // f, f, 48 // f, f, 3).mean(axis=(1, 3))
        np.testing.assert_allclose(got, want, atol=1e-6)


def _mirror_levels():
    torch.manual_seed(0)
    return [fit.Gaussian2D(num_points=50, H=16, W=20, device="cpu", lr=1e-2),
            fit.Gaussian2D(num_points=80, H=32, W=40, device="cpu", lr=1e-2)]


def test_checkpoint_round_trip_and_strict_keys(tmp_path):
    """Gaussian2D levels saved in train.py's layout load back into Gaussian2D
    levels; a level whose keys do not match (a Cholesky Level given a cov2d
    checkpoint) raises instead of silently keeping its own values, and so does
    a level count that differs."""
    src = _mirror_levels()
    with torch.no_grad():
        for m in src:
            m.rgbs.uniform_(0, 1)
    path = os.path.join(tmp_path, "ck.pth.tar")
    fit.save_checkpoint(path, src, [torch.tensor(-0.25)], [torch.tensor(1.5)])
    ck = torch.load(path, weights_only=True)
    assert sorted(ck) == ["state_dict", "store_max", "store_min"]
    assert sorted(ck["state_dict"][0]) == ["cov2d", "means", "rgbs"]
    torch.manual_seed(1)
    dst = [fit.Gaussian2D(num_points=50, H=16, W=20, device="cpu", lr=1e-2),
           fit.Gaussian2D(num_points=80, H=32, W=40, device="cpu", lr=1e-2)]
    smin, smax = fit.load_checkpoint(path, dst)
    for a, b in zip(src, dst):
        for k in ("means", "cov2d", "rgbs"):
            assert torch.equal(getattr(a, k), getattr(b, k)), k
    assert float(smin[0]) == -0.25 and float(smax[0]) == 1.5
    chol_levels = [fit.Level(50, 16, 20, seed=1, device="cpu"),
                   fit.Level(80, 32, 40, seed=2, device="cpu")]
    with pytest.raises(KeyError, match="missing .*cholesky.*unexpected .*cov2d"):
        fit.load_checkpoint(path, chol_levels)
    with pytest.raises(ValueError, match="levels"):
        fit.load_checkpoint(path, dst[:1])


def test_two_level_from_checkpoint_layout(tmp_path):
    """render.TwoLevel.from_checkpoint reads train.py's layout: level s of n is
    int(H 2^(s+1-n)) on a side (gaussianlig.py:20-22), cov2d levels keep the
    covariance parameterisation, Cholesky levels theirs, and the level-2
    bounds are store_min[0] / store_max[0]; anything else raises."""
    from lig_amd.render import TwoLevel
    src = _mirror_levels()
    path = os.path.join(tmp_path, "ck.pth.tar")
    fit.save_checkpoint(path, src, [torch.tensor(-0.25)], [torch.tensor(1.5)])
    rep = TwoLevel.from_checkpoint(path, 32, 40, device="cpu")
    assert (rep.l1.H, rep.l1.W, rep.l2.H, rep.l2.W) == (16, 20, 32, 40)
    assert rep.l1.kind == rep.l2.kind == "cov2d" and rep.upsample == "bilinear"
    assert torch.equal(rep.l2.cov2d, src[1].cov2d.detach())
    assert (rep.lo, rep.hi) == (-0.25, 1.5)
    with pytest.raises(AttributeError):
        rep.l1.cholesky
    chol = [fit.Level(50, 16, 20, seed=1, device="cpu"), fit.Level(80, 32, 40, seed=2,
                                                                    device="cpu")]
    fit.save_checkpoint(path, chol, [torch.tensor(0.0)], [torch.tensor(1.0)])
    rep = TwoLevel.from_checkpoint(path, 32, 40, device="cpu", upsample="hermite")
    assert rep.l2.kind == "cholesky" and torch.equal(rep.l2.cholesky, chol[1].cholesky.detach())
    torch.save({"state_dict": [{"means": torch.zeros(2, 2), "rgbs": torch.zeros(2, 3)}] * 2,
                "store_min": [torch.tensor(0.0)], "store_max": [torch.tensor(1.0)]}, path)
    with pytest.raises(KeyError):
        TwoLevel.from_checkpoint(path, 32, 40, device="cpu")
    fit.save_checkpoint(path, src[:1], [], [])
    with pytest.raises(ValueError):
        TwoLevel.from_checkpoint(path, 32, 40, device="cpu")
    with pytest.raises(ValueError):
        TwoLevel(rep.l1, rep.l2, 0.0, 1.0, upsample="bicubic")


def test_crop_views_are_tile_aligned():
    """Crops render on the full image's 16-pixel tile grid (an indefinite
    covariance is drawn only inside its tile rectangle): the widened view
    starts on a tile edge, covers the window, ends on a tile edge or the
    image's edge, and the window's offset inside it is exact."""
    from lig_amd.render import LevelParams, _aligned_view, _tile_span
    rng = np.random.default_rng(0)
    for _ in range(500):
        full = int(rng.integers(1, 300))
        o = int(rng.integers(0, full))
        n = int(rng.integers(1, full - o + 1))
        a, b = _tile_span(o, n, full)
        assert a % 16 == 0 and a <= o < a + 16 and b >= o + n
        assert b == full or (b % 16 == 0 and b - (o + n) < 16)
    p = LevelParams(torch.zeros(1, 2), torch.ones(1, 3), torch.zeros(1, 3), 256, 320, "cov2d")
    (q, x0, y0, w, h, s), off = _aligned_view(p, 100.0, 36.0, 128, 96, 1.0)
    assert (x0, y0, w, h, off) == (96.0, 32.0, 144, 112, (4, 4))
    (q, x0, y0, w, h, s), off = _aligned_view(p, 100.0, 36.0, 64, 48, 0.5)
    assert (x0, y0, w, h, off) == (96.0, 32.0, 80, 64, (2, 2))
    (q, x0, y0, w, h, s), off = _aligned_view(p, 288.0, 240.0, 32, 16, 1.0)
    assert (x0, y0, w, h, off) == (288.0, 240.0, 32, 16, (0, 0))
    # off the output pixel grid: rendered as asked
    assert _aligned_view(p, 10.5, 3.0, 16, 16, 1.0) == ((p, 10.5, 3.0, 16, 16, 1.0), (0, 0))
