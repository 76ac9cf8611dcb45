"""GPU parity of the rasteriser (plain and gradient-aware, forward and backward)
and the upscaler against the CPU oracle.

Tolerances (stated per output):
  * images / wsum / dx / dy / dxy: max |gpu - oracle_fp64| <= 1e-5 over pixels
    with no ambiguous pair (a pair within relative 1e-5 of the 1/255 cut);
  * final_idx: exact on the same pixels;
  * gradients: norm-wise relative error <= 1e-4 against the reference-gradient
    restatement (FP32 pair math, oracle mode "ref32") and against the FP64
    oracle, over Gaussians with no ambiguous pair.
"""
import numpy as np
import pytest
import torch

from tests.scenes import random_scene

pytestmark = pytest.mark.gpu


def _t(a, dev, dtype=torch.float32):
    return None if a is None else torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(dev)


def _np(t):
    return t.detach().cpu().numpy()


def _prepare(oracle_lib, s, bw):
    H, W = s["H"], s["W"]
    xys, ext, con, nt = oracle_lib.project_gaussians_forward(
        s["par"], s["means"], H, W, bw, s["opac"], cholesky=s["cholesky"])
    m, cum, isect, gids, ks, gs, bins = oracle_lib.bin_and_sort(xys, ext, nt, H, W, bw)
    return dict(xys=xys, ext=ext, con=con, nt=nt, m=m, cum=cum, gs=gs, bins=bins)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


RASTER_CASES = [
    # n, H, W, bw, opacities, seed, scale
    (200, 64, 64, 16, True, 0, (0.5, 3.0)),
    (600, 96, 80, 16, False, 1, (0.5, 3.0)),
    (400, 64, 64, 16, True, 2, (2.0, 9.0)),
    (300, 45, 70, 8, True, 3, (0.5, 4.0)),
    (250, 40, 40, 5, False, 4, (0.5, 3.0)),
    (120, 33, 50, 12, True, 5, (1.0, 5.0)),
]


@pytest.mark.parametrize("ga", [False, True])
@pytest.mark.parametrize("n,H,W,bw,op,seed,scale", RASTER_CASES)
def test_raster_forward_backward(device, oracle_lib, n, H, W, bw, op, seed, scale, ga):
    from lig_amd.gsplat2d import cuda as C
    print(f"raster case n={n} H={H} W={W} bw={bw} op={op} seed={seed} ga={ga}", flush=True)
    s = random_scene(n, H, W, seed, opacities=op, scale=scale, margin=10.0)
    if op:
        s["opac"][1::13] = 0.002   # zero extent: phantom (tile 0, gid 0) slots
    p = _prepare(oracle_lib, s, bw)
    dev = device
    gs, bins = _t(p["gs"], dev, torch.int32), _t(p["bins"], dev, torch.int32)
    xys, con, col = _t(p["xys"], dev), _t(p["con"], dev), _t(s["colors"], dev)
    opac = _t(s["opac"], dev)
    img, wsum, dx, dy, dxy, fin, packed = C.rasterize_forward_packed(
        H, W, bw, gs, bins, xys, con, col, opac, ga)
    ref = oracle_lib.rasterize_forward(p["gs"], p["bins"], p["xys"], p["con"], s["colors"],
                                       s["opac"], H, W, bw, gradient_aware=ga)
    clean = ref["ambig"] == 0
    assert clean.mean() > 0.9
    np.testing.assert_array_equal(_np(fin)[clean], ref["final_idx"][clean])
    assert np.abs(_np(img) - ref["img"])[clean].max() <= 1e-5
    assert np.abs(_np(wsum) - ref["wsum"])[clean].max() <= 1e-5
    if ga:
        for a, b in ((dx, "dx"), (dy, "dy"), (dxy, "dxy")):
            assert np.abs(_np(a) - ref[b])[clean].max() <= 1e-5, b

    rng = np.random.default_rng(seed + 100)
    v_out = rng.normal(size=(H, W, 3)).astype(np.float32)
    v_w = rng.normal(size=(H, W)).astype(np.float32)
    v_d = [rng.normal(size=(H, W, 3)).astype(np.float32) for _ in range(3)] if ga else [None] * 3
    fin_np = _np(fin)
    cum = _t(p["cum"], dev, torch.int32)
    gs2, bins2, eids = C.bin_and_sort_fast(n, p["m"], xys, _t(p["ext"], dev), cum, H, W, bw,
                                           with_emission_ids=True)
    assert torch.equal(gs2, gs) and torch.equal(bins2, bins)
    args = (H, W, bw, gs, bins, packed, n, fin, _t(v_out, dev), _t(v_w, dev), op,
            *[_t(v, dev) for v in v_d])
    got_atomic = C.rasterize_backward_packed(*args)
    got = C.rasterize_backward_packed(*args, emission_ids_sorted=eids, cum_tiles_hit=cum)
    again = C.rasterize_backward_packed(*args, emission_ids_sorted=eids, cum_tiles_hit=cum)
    for a, b in zip(got, again):   # deterministic reduction: bit-identical run to run
        assert (a is None and b is None) or torch.equal(a, b)
    names = ("v_xy", "v_xy_abs", "v_conic", "v_rgb", "v_opacity")
    r64 = oracle_lib.rasterize_backward(p["gs"], p["bins"], p["xys"], p["con"], s["colors"],
                                        s["opac"], fin_np, v_out, v_w, H, W, bw, *v_d,
                                        mode="fp64")
    r32 = oracle_lib.rasterize_backward(p["gs"], p["bins"], p["xys"], p["con"], s["colors"],
                                        s["opac"], fin_np, v_out, v_w, H, W, bw, *v_d,
                                        mode="ref32")
    keep = r64["ambig"] == 0
    assert keep.mean() > 0.8
    for name, g, ga_ in zip(names, got, got_atomic):
        if g is None:
            assert name == "v_opacity" and not op and ga_ is None
            continue
        np.testing.assert_allclose(_np(ga_), _np(g), rtol=1e-5,
                                   atol=1e-6 * float(np.abs(_np(g)).max() + 1e-30))
        a = _np(g)[keep].astype(np.float64)
        e64 = _rel(a, r64[name][keep])
        e32 = _rel(a, r32[name][keep])
        print(f"  {name}: rel err vs fp64 {e64:.2e}, vs ref32 {e32:.2e}")
        assert e32 <= 1e-4, (name, e32)
        assert e64 <= 1e-4, (name, e64)


@pytest.mark.parametrize("bw", [16, 8, 12, 5])
def test_ga_absgrad_per_warp_group(device, oracle_lib, bw):
    """Gradient-aware xys.absgrad is the sum over 32-thread groups (thread
    rank ty*bw + tx) of |group sum of v_xy| (gradient_aware_rasterizer.cu:320,
    339-341), on a scene where that differs from the per-pixel sum of |v_xy|
    by more than 10%; the plain kernel keeps the per-pixel form."""
    from lig_amd.gsplat2d import cuda as C
    n, H, W = 150, 50, 61
    s = random_scene(n, H, W, 60 + bw, opacities=True, scale=(1.5, 5.0), margin=5.0)
    s["opac"] = np.clip(s["opac"], 0.1, 0.9).astype(np.float32)
    print(f"ga absgrad bw={bw}", flush=True)
    p = _prepare(oracle_lib, s, bw)
    dev = device
    gs, bins = _t(p["gs"], dev, torch.int32), _t(p["bins"], dev, torch.int32)
    cum = _t(p["cum"], dev, torch.int32)
    xys, con, col, opac = (_t(p["xys"], dev), _t(p["con"], dev), _t(s["colors"], dev),
                           _t(s["opac"], dev))
    rng = np.random.default_rng(bw)
    v_out = rng.normal(size=(H, W, 3)).astype(np.float32)
    v_w = rng.normal(size=(H, W)).astype(np.float32)
    v_d = [rng.normal(size=(H, W, 3)).astype(np.float32) for _ in range(3)]
    _, _, eids = C.bin_and_sort_fast(n, p["m"], xys, _t(p["ext"], dev), cum, H, W, bw,
                                     with_emission_ids=True)
    out = {}
    for ga in (False, True):
        f = C.rasterize_forward_packed(H, W, bw, gs, bins, xys, con, col, opac, ga)
        fin_np = _np(f[5])
        vd = v_d if ga else [None] * 3
        got = C.rasterize_backward_packed(H, W, bw, gs, bins, f[6], n, f[5], _t(v_out, dev),
                                          _t(v_w, dev), True, *[_t(v, dev) for v in vd],
                                          emission_ids_sorted=eids, cum_tiles_hit=cum)
        r64 = oracle_lib.rasterize_backward(p["gs"], p["bins"], p["xys"], p["con"],
                                            s["colors"], s["opac"], fin_np, v_out, v_w, H, W,
                                            bw, *vd, mode="fp64")
        keep = r64["ambig"] == 0
        a = _np(got[1])[keep].astype(np.float64)
        e = _rel(a, r64["v_xy_abs"][keep])
        print(f"  ga={ga}: v_xy_abs rel err vs fp64 {e:.2e}", flush=True)
        assert e <= 1e-4, (ga, e)
        out[ga] = _np(got[1]).astype(np.float64)
    # the per-pixel form of the GA gradient: with 1 x 1 blocks every pixel is
    # its own group, so the oracle's absgrad is then the sum of per-pixel |v_xy|
    p1 = _prepare(oracle_lib, s, 1)
    f1 = oracle_lib.rasterize_forward(p1["gs"], p1["bins"], p1["xys"], p1["con"], s["colors"],
                                      s["opac"], H, W, 1)
    r1 = oracle_lib.rasterize_backward(p1["gs"], p1["bins"], p1["xys"], p1["con"], s["colors"],
                                       s["opac"], f1["final_idx"], v_out, v_w, H, W, 1, *v_d,
                                       mode="fp64")
    per_pixel = r1["v_xy_abs"].sum()
    print(f"  GA absgrad total {out[True].sum():.4g} vs per-pixel form {per_pixel:.4g}",
          flush=True)
    assert out[True].sum() < 0.9 * per_pixel


@pytest.mark.parametrize("offdiag,op", [(0.9, False), (0.999, False), (0.95, True)])
def test_forward_region_cull_keeps_accepted_pairs(device, oracle_lib, offdiag, op):
    """The 16 x 16 forward drops the quadrants of a tile that a Gaussian's
    bounding box reaches but its accept region does not; on rotated, strongly
    elongated, large and tiny Gaussians (opacities up to 1.5 and below 1/255,
    means off the canvas) every accepted pair is still drawn: image and final
    index as the FP64 oracle's on pixels without an ambiguous pair."""
    from lig_amd.gsplat2d import cuda as C
    n, H, W, bw = 500, 80, 96, 16
    s = random_scene(n, H, W, 77 + int(op), opacities=op, scale=(0.3, 12.0), offdiag=offdiag,
                     margin=30.0)
    if op:
        s["opac"][::7] = 1.5
        s["opac"][3::11] = 0.002
    p = _prepare(oracle_lib, s, bw)
    img, wsum, _, _, _, fin, _ = C.rasterize_forward_packed(
        H, W, bw, _t(p["gs"], device, torch.int32), _t(p["bins"], device, torch.int32),
        _t(p["xys"], device), _t(p["con"], device), _t(s["colors"], device),
        _t(s["opac"], device), False)
    ref = oracle_lib.rasterize_forward(p["gs"], p["bins"], p["xys"], p["con"], s["colors"],
                                       s["opac"], H, W, bw)
    clean = ref["ambig"] == 0
    print(f"offdiag {offdiag} op {op}: {int(p['m'])} intersections, clean {clean.mean():.3f}",
          flush=True)
    assert clean.mean() > 0.9
    np.testing.assert_array_equal(_np(fin)[clean], ref["final_idx"][clean])
    assert np.abs(_np(img) - ref["img"])[clean].max() <= 1e-5
    assert np.abs(_np(wsum) - ref["wsum"])[clean].max() <= 1e-5


@pytest.mark.parametrize("H,skew,op", [(4096, True, False), (4096, True, True), (96, False, True)])
def test_api_launch_order_changes_no_bits(device, H, skew, op):
    """The API raster calls order their 16 x 16 tiles' launch in workspace
    (heaviest first when the launch is short or the per-XCD shares' loads
    differ); ordered and fixed-order launches give the same bits: image,
    weight sum, final index, and the deterministic backward's gradients."""
    import lig_amd.gsplat2d as g
    from lig_amd import fit
    from lig_amd.gsplat2d import cuda as C
    n = 60000 if H > 1000 else 3000
    print(f"api launch order H={H} skew={skew} op={op}", flush=True)
    lv = fit.Level(n, H, H, 8, 3.0, device=device)
    with torch.no_grad():
        if skew:
            lv.means[: n // 2, 1] *= 0.25
        xys, ext, con, nt = lv.project()
    opac = torch.rand(n, device=device) * 0.9 + 0.05 if op else None
    num, cum = g.compute_cumulative_intersects(nt)
    gids, bins, eids = C.bin_and_sort_fast(n, num, xys, ext, cum, H, H, 16,
                                           with_emission_ids=True)
    runs = []
    v_out = torch.randn(H, H, 3, device=device)
    v_w = torch.randn(H, H, device=device)
    for ordered in (True, False):
        img, wsum, _, _, _, fin, packed = C.rasterize_forward_packed(
            H, H, 16, gids, bins, xys, con, lv.rgbs.detach(), opac, False, ordered=ordered)
        grads = C.rasterize_backward_packed(H, H, 16, gids, bins, packed, n, fin, v_out, v_w, op,
                                            emission_ids_sorted=eids, cum_tiles_hit=cum,
                                            ordered=ordered)
        runs.append((img, wsum, fin) + tuple(x for x in grads if x is not None))
    assert torch.count_nonzero(runs[0][3]) > 0
    for k, (a, b) in enumerate(zip(*runs)):
        assert torch.equal(a, b), k


def test_raster_api_end_to_end(device, oracle_lib):
    """project -> rasterize -> loss -> backward through the autograd wrappers."""
    import lig_amd.gsplat2d as g
    n, H, W = 300, 64, 80
    s = random_scene(n, H, W, 9, opacities=True, margin=5.0)
    chol = _t(s["par"], device).requires_grad_(True)
    means = _t(s["means"], device).requires_grad_(True)
    col = _t(s["colors"], device).requires_grad_(True)
    opac = _t(s["opac"], device).requires_grad_(True)
    xys, ext, con, nt = g.project_gaussians_cholesky(chol, means, H, W, 16, opac)
    img, wsum, dx, dy, dxy = g.rasterize_gaussians(xys, ext, con, nt, col, opac, H, W, 16)
    assert dx is None and dy is None and dxy is None
    loss = (img * torch.linspace(0, 1, 3, device=device)).sum() + 0.3 * wsum.sum()
    loss.backward()
    for t in (chol, means, col, opac):
        assert t.grad is not None and torch.isfinite(t.grad).all()
    assert xys.absgrad.shape == (n, 2)

    # nothing on screen: the ones image (rasterize.py:80-90) and zero gradients
    far = _t(s["means"] + 1e5, device).requires_grad_(True)
    xys, ext, con, nt = g.project_gaussians_cholesky(chol, far, H, W, 16, opac)
    img, wsum, dx, dy, dxy = g.rasterize_gaussians(xys, ext, con, nt, col, opac, H, W, 16,
                                                   compute_upscale_gradients=True)
    assert torch.all(img == 1) and torch.all(wsum == 0) and torch.all(dx == 0)
    img.sum().backward()
    assert torch.all(far.grad == 0)


@pytest.mark.parametrize("ga", [False, True])
def test_record_addressing_modes_agree(device, oracle_lib, ga):
    """16 x 16 tiles without opacities: the deterministic backward addressing
    its records from the pack (bins sorted with gids) gives the same bits as
    the one addressing them through sorted emission ids, and the API path
    (which takes the pack mode) agrees with both."""
    import lig_amd.gsplat2d as g
    from lig_amd.gsplat2d import cuda as C
    n, H, W = 900, 100, 130
    s = random_scene(n, H, W, 21, opacities=False, scale=(0.5, 6.0), margin=10.0)
    xys, ext, con, nt = (_t(a, device) if i < 3 else _t(a, device, torch.int32) for i, a in
                         enumerate(oracle_lib.project_gaussians_forward(
                             s["par"], s["means"], H, W, 16, None, cholesky=True)))
    col = _t(s["colors"], device)
    num, cum = g.compute_cumulative_intersects(nt)
    g1, b1, eids = C.bin_and_sort_fast(n, num, xys, ext, cum, H, W, 16, with_emission_ids=True)
    g2, b2 = C.bin_and_sort_fast(n, num, xys, ext, cum, H, W, 16)
    assert torch.equal(g1, g2) and torch.equal(b1, b2)
    f1 = C.rasterize_forward_packed(H, W, 16, g1, b1, xys, con, col, None, ga)
    f2 = C.rasterize_forward_packed(H, W, 16, g2, b2, xys, con, col, None, ga,
                                    extents=ext, cum_tiles_hit=cum)
    for a, b in zip(f1[:6], f2[:6]):
        assert (a is None and b is None) or torch.equal(a, b)
    rng = np.random.default_rng(3)
    v_out = _t(rng.normal(size=(H, W, 3)), device)
    v_w = _t(rng.normal(size=(H, W)), device)
    extra = [_t(rng.normal(size=(H, W, 3)), device) for _ in range(3)] if ga else [None] * 3
    r1 = C.rasterize_backward_packed(H, W, 16, g1, b1, f1[6], n, f1[5], v_out, v_w, False,
                                     *extra, emission_ids_sorted=eids, cum_tiles_hit=cum)
    r2 = C.rasterize_backward_packed(H, W, 16, g2, b2, f2[6], n, f2[5], v_out, v_w, False,
                                     *extra, cum_tiles_hit=cum, slots_in_pack=True)
    for name, a, b in zip(("v_xy", "v_xy_abs", "v_conic", "v_colors"), r1[:4], r2[:4]):
        assert torch.equal(a, b), name
    assert torch.count_nonzero(r1[3]) > 0


UPSCALE_CASES = [
    (16, 16, 64, 64, None),
    (20, 30, 80, 120, None),
    (17, 13, 50, 41, None),
    (32, 32, 16, 16, None),
    (24, 24, 48, 48, (4.0, 2.0, 20.0, 18.0)),
    (24, 24, 37, 29, (2.5, 3.5, 21.0, 22.7)),
]


@pytest.mark.parametrize("sh,sw,dh,dw,roi", UPSCALE_CASES)
def test_upscale_forward_backward(device, oracle_lib, sh, sw, dh, dw, roi):
    import lig_amd.gsplat2d as g
    rng = np.random.default_rng(sh * 1000 + dh)
    f, fx, fy, fxy = (rng.normal(size=(sh, sw, 3)).astype(np.float32) for _ in range(4))
    ts = [_t(a, device).requires_grad_(True) for a in (f, fx, fy, fxy)]
    out = g.gradient_aware_upscale(*ts, dh, dw, roi)
    ref = oracle_lib.upscale_forward(f, fx, fy, fxy, dh, dw, roi)
    assert np.abs(_np(out) - ref).max() <= 1e-5 * max(1.0, np.abs(ref).max())
    gout = rng.normal(size=(dh, dw, 3)).astype(np.float32)
    out.backward(_t(gout, device))
    rg = oracle_lib.upscale_backward(gout, sh, sw, roi)
    for t, r in zip(ts, rg):
        assert np.abs(_np(t.grad) - r).max() <= 1e-5 * max(1.0, np.abs(r).max())


def test_upscale_backward_deterministic(device):
    from lig_amd.gsplat2d import cuda as C
    rng = np.random.default_rng(1)
    g = _t(rng.normal(size=(256, 256, 3)).astype(np.float32), device)
    src = torch.zeros(64, 64, 3, device=device)
    a = C.gradient_aware_upscale_backward(g, src, src, src, src, 256, 256, (0., 0., 64., 64.))
    b = C.gradient_aware_upscale_backward(g, src, src, src, src, 256, 256, (0., 0., 64., 64.))
    for x, y in zip(a, b):
        assert torch.equal(x, y)


@pytest.mark.parametrize("shape", [(64, 80, 3), (1, 7), (513, 257, 3)])
def test_fused_l2_loss_matches_torch(device, shape):
    """Fused clamp + MSE equals F.mse_loss(clamp(img, 0, 1), target) and its
    gradient (torch on the CPU in float64 as the reference)."""
    import torch.nn.functional as F
    from lig_amd.loss import l2_loss_clamped
    gen = torch.Generator().manual_seed(sum(shape))
    img = (torch.rand(shape, generator=gen) * 1.4 - 0.2)
    img.view(-1)[:5] = torch.tensor([0.0, 1.0, -0.0, 1.0000001, -1e-8])  # clamp bounds
    tgt = torch.rand(shape, generator=gen)
    x = img.to(device).requires_grad_(True)
    loss = l2_loss_clamped(x, tgt.to(device))
    (loss * 3.0).backward()
    xr = img.double().requires_grad_(True)
    lr = F.mse_loss(torch.clamp(xr, 0, 1), tgt.double())
    (lr * 3.0).backward()
    assert abs(loss.item() - lr.item()) <= 1e-6 * max(1.0, lr.item())
    np.testing.assert_allclose(x.grad.cpu().numpy(), xr.grad.numpy(), rtol=1e-5, atol=1e-12)


@pytest.mark.parametrize("case", ["pack", "eids_opac", "dense", "skew_opac"])
def test_api_backward_without_final_index(device, case):
    """16 x 16 plain API path without a final index: the forward writes the
    same image and weight sum without it, and the backward -- no pixel gated,
    each tile's Gaussians grouped by the rows they reach -- gives the
    gradients of the backward gated by the forward's final index bit for bit,
    with records addressed from the pack or through emission ids, with and
    without opacities, on tiles deeper than one 512-Gaussian sort chunk and on
    sizes that are not multiples of 16; without a wsum gradient the same bits
    as with a zero one; the atomic reduction agrees within its
    summation-order rounding."""
    import lig_amd.gsplat2d as g
    from lig_amd import fit
    from lig_amd.gsplat2d import cuda as C
    H, W, n, scale = {"pack": (100, 130, 900, 6.0), "eids_opac": (96, 80, 1200, 4.0),
                      "dense": (52, 70, 6000, 8.0), "skew_opac": (200, 150, 8000, 5.0)}[case]
    op = case.endswith("opac")
    lv = fit.Level(n, H, W, 31, scale, device=device)
    with torch.no_grad():
        if case == "skew_opac":
            lv.means[: n // 2, 1] *= 0.25
        xys, ext, con, nt = lv.project()
    opac = torch.rand(n, device=device, generator=torch.Generator(device=device).manual_seed(4)
                      ) * 0.9 + 0.05 if op else None
    num, cum = g.compute_cumulative_intersects(nt)
    col = lv.rgbs.detach()
    pack_mode = case in ("pack", "dense")
    if pack_mode:
        gids, bins = C.bin_and_sort_fast(n, num, xys, ext, cum, H, W, 16)
        kw = dict(extents=ext, cum_tiles_hit=cum)
        bk = dict(cum_tiles_hit=cum, slots_in_pack=True)
    else:
        gids, bins, eids = C.bin_and_sort_fast(n, num, xys, ext, cum, H, W, 16,
                                               with_emission_ids=True)
        kw = {}
        bk = dict(emission_ids_sorted=eids, cum_tiles_hit=cum)
    if case == "dense":
        depth = int((bins[:, 1] - bins[:, 0]).max())
        print(f"deepest tile: {depth} Gaussians", flush=True)
        assert depth > 512
    f_gate = C.rasterize_forward_packed(H, W, 16, gids, bins, xys, con, col, opac, False, **kw)
    f_free = C.rasterize_forward_packed(H, W, 16, gids, bins, xys, con, col, opac, False,
                                        final_index=False, **kw)
    assert f_free[5] is None
    assert torch.equal(f_gate[0], f_free[0]) and torch.equal(f_gate[1], f_free[1])
    gen = torch.Generator(device=device).manual_seed(7)
    v_out = torch.randn(H, W, 3, device=device, generator=gen)
    v_w = torch.randn(H, W, device=device, generator=gen)
    r_gate = C.rasterize_backward_packed(H, W, 16, gids, bins, f_gate[6], n, f_gate[5], v_out,
                                         v_w, op, **bk)
    r_free = C.rasterize_backward_packed(H, W, 16, gids, bins, f_free[6], n, None, v_out, v_w,
                                         op, **bk)
    names = ("v_xy", "v_xy_abs", "v_conic", "v_colors", "v_opacity")
    for name, a, b in zip(names, r_gate, r_free):
        assert (a is None) == (b is None), name
        if a is not None:
            assert torch.equal(a, b), (name, float((a - b).abs().max()))
    assert torch.count_nonzero(r_gate[3]) > 0
    # no wsum gradient (a loss of the image alone): the variant without the
    # wsum terms gives the bits of a zero wsum gradient
    r_none = C.rasterize_backward_packed(H, W, 16, gids, bins, f_free[6], n, None, v_out, None,
                                         op, **bk)
    r_zero = C.rasterize_backward_packed(H, W, 16, gids, bins, f_free[6], n, None, v_out,
                                         torch.zeros_like(v_w), op, **bk)
    for name, a, b in zip(names, r_none, r_zero):
        if a is not None:
            assert torch.equal(a, b), (name, float((a - b).abs().max()))
    # atomics (no records): the same sums in another order
    a_gate = C.rasterize_backward_packed(H, W, 16, gids, bins, f_gate[6], n, f_gate[5], v_out,
                                         v_w, op)
    a_free = C.rasterize_backward_packed(H, W, 16, gids, bins, f_free[6], n, None, v_out, v_w,
                                         op)
    for name, a, b in zip(names, a_gate, a_free):
        if a is not None:
            tol = 1e-4 * max(1.0, float(a.abs().max()))
            assert float((a - b).abs().max()) <= tol, name
