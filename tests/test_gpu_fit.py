"""GPU tests of the caller layer (gaussianlig.py / train.py semantics)."""
import os

import pytest
import torch

pytestmark = pytest.mark.gpu


def test_level_fit_reduces_loss(device):
    from lig_amd import data, fit
    gt = data.make_image(128, 160, 0, device=device)
    lvl = fit.Level(3000, 128, 160, seed=1, lr=1e-2, device=device)
    first = float(lvl.step(gt))
    for _ in range(40):
        last = float(lvl.step(gt))
    assert last < 0.7 * first, (first, last)


def test_reference_mirror_classes_train(device):
    from lig_amd import fit
    torch.manual_seed(0)
    model = fit.LIG(num_points=4000, H=128, W=128, n_scales=2, allo_ratio=0.5, device=device,
                    lr=1e-2)
    assert [m.H for m in model.level_models] == [64, 128]
    assert [m.init_num_points for m in model.level_models] == [500, 3500]
    gt = torch.rand(1, 3, 64, 64, device=device)
    psnrs = [model.level_models[0].train_iter(gt)[1] for _ in range(5)]
    assert all(p == p for p in psnrs)   # finite


def test_two_level_and_checkpoint(device, tmp_path):
    from lig_amd import data, fit
    gt = data.make_image(128, 128, 1, device=device)
    l1, l2, p = fit.two_level_fit(gt, 800, 2400, steps=30, lr=1e-2, device=device)
    assert p > 10.0
    path = os.path.join(tmp_path, "ck.pth")
    fit.save_checkpoint(path, [l1, l2], [torch.tensor(0.1)], [torch.tensor(0.9)])
    l1b = fit.Level(800, 32, 32, seed=9, device=device)
    l2b = fit.Level(2400, 128, 128, seed=9, device=device)
    smin, smax = fit.load_checkpoint(path, [l1b, l2b], map_location=device)
    assert torch.equal(l2b.means, l2.means) and float(smax[0]) == pytest.approx(0.9)


def _level_pair(n, H, W, seed, scale, device):
    from lig_amd import fit
    a = fit.Level(n, H, W, seed=seed, scale=scale, lr=1e-2, device=device, fused=True)
    b = fit.Level(n, H, W, seed=seed, scale=scale, lr=1e-2, device=device, fused=False)
    return a, b


@pytest.mark.parametrize("n,H,W,scale,seed", [
    (3000, 100, 130, 1.0, 3),     # edge tiles in x and y
    (6000, 128, 160, 2.5, 4),     # several 64-Gaussian batches per tile
    (800, 48, 40, 4.0, 5),        # Gaussians larger than the canvas
])
def test_fused_step_matches_api_chain(device, n, H, W, scale, seed):
    """The fused render + loss node gives the gradients of the API chain
    project_gaussians_cholesky -> rasterize_gaussians -> l2_loss_clamped
    bit for bit (same pair test, same record order) and the loss to FP32
    summation order; one Adam step later the parameters are still equal."""
    from lig_amd import data
    print(f"fused-vs-api n={n} H={H} W={W} scale={scale} seed={seed}", flush=True)
    tgt = data.make_image(H, W, seed, device=device)
    a, b = _level_pair(n, H, W, seed, scale, device)
    with torch.no_grad():
        b.rgbs.mul_(0.05)   # keep part of the image inside (0, 1) so the loss mask varies
        a.rgbs.copy_(b.rgbs)
    la, lb = a.loss(tgt), b.loss(tgt)
    assert float(la.detach()) == pytest.approx(float(lb.detach()), rel=1e-6)
    la.backward()
    lb.backward()
    for name in ("means", "cholesky", "rgbs"):
        ga, gb = getattr(a, name).grad, getattr(b, name).grad
        assert torch.count_nonzero(gb) > 0, name
        assert torch.equal(ga, gb), (name, float((ga - gb).abs().max()))
    a.optimizer.step()
    b.optimizer.step()
    for name in ("means", "cholesky", "rgbs"):
        assert torch.equal(getattr(a, name), getattr(b, name)), name


@pytest.mark.parametrize("n,H,W,seed,shares", [
    (12000, 96, 96, 6, 1),      # each tile's backward in one block
    (20000, 100, 130, 7, 0),    # partial edge rows and columns, shares chosen per launch
])
def test_fused_step_matches_api_chain_dense_tiles(device, n, H, W, seed, shares):
    """Thousands of Gaussians per tile: the fit backward groups each chunk of
    512 of a tile's Gaussians by the rows they reach (several chunks per tile
    here, Gaussians reaching no row among them; on the 100 x 130 canvas the
    edge tiles' row masks cut the groups, and the tile's batches may be split
    over several blocks) and still gives the API chain's gradients bit for bit."""
    import lig_amd._lib as L
    from lig_amd import data
    print(f"fused-vs-api dense n={n} H={H} W={W} seed={seed} shares={shares}", flush=True)
    lib = L.load()
    prev = lib.lig_set_backward_shares(shares)
    try:
        tgt = data.make_image(H, W, seed, device=device)
        a, b = _level_pair(n, H, W, seed, 3.0, device)
        with torch.no_grad():
            b.rgbs.mul_(0.01)
            a.rgbs.copy_(b.rgbs)
            nt = a.project()[3]
        ntiles = ((H + 15) // 16) * ((W + 15) // 16)
        assert int(nt.sum()) > ntiles * 2 * 512  # several 512-Gaussian chunks per tile
        la, lb = a.loss(tgt), b.loss(tgt)
        la.backward()
        lb.backward()
        for name in ("means", "cholesky", "rgbs"):
            ga, gb = getattr(a, name).grad, getattr(b, name).grad
            assert torch.count_nonzero(gb) > 0, name
            assert torch.equal(ga, gb), (name, float((ga - gb).abs().max()))
    finally:
        lib.lig_set_backward_shares(prev)


@pytest.mark.parametrize("skew", [False, True])
def test_launch_order_changes_no_bits(device, skew):
    """A 4096 x 4096 level (65536 tiles, many rounds of resident blocks):
    uniform, the fit kernels keep the per-XCD bands; with half the Gaussians
    pulled into the top quarter, the bands' loads differ and the tiles run
    heaviest first.  Either way the fused step's gradients equal the API
    chain's (which launches in band order) bit for bit."""
    from lig_amd import data
    H = W = 4096
    print(f"launch order skew={skew}", flush=True)
    tgt = data.make_image(H, W, 6, device=device)
    a, b = _level_pair(60000, H, W, 6, 3.0, device)
    with torch.no_grad():
        b.rgbs.mul_(0.05)
        if skew:
            b.means[:30000, 1] *= 0.25
        a.rgbs.copy_(b.rgbs)
        a.means.copy_(b.means)
    la, lb = a.loss(tgt), b.loss(tgt)
    assert float(la.detach()) == pytest.approx(float(lb.detach()), rel=1e-6)
    la.backward()
    lb.backward()
    for name in ("means", "cholesky", "rgbs"):
        ga, gb = getattr(a, name).grad, getattr(b, name).grad
        assert torch.count_nonzero(gb) > 0, name
        assert torch.equal(ga, gb), (name, float((ga - gb).abs().max()))


def test_fused_render_image_matches_api(device):
    from lig_amd import data
    from lig_amd.gsplat2d import cuda as C
    import lig_amd.gsplat2d as g
    H, W = 90, 150
    a, _ = _level_pair(2500, H, W, 7, 1.5, device)
    tgt = data.make_image(H, W, 7, device=device)
    with torch.no_grad():
        xys, ext, con, nt = a.project()
        num, cum = g.compute_cumulative_intersects(nt)
        gids, bins = C.bin_and_sort_fast(a.means.shape[0], num, xys, ext, cum, H, W, 16)
        img, loss, _ = C.fit_render_l2(H, W, gids, bins, xys, ext, cum, con, a.rgbs, tgt)
        ref = g.rasterize_gaussians(xys, ext, con, nt, a.rgbs, None, H, W, 16)[0]
    assert torch.equal(img, ref)
    want = ((ref.clamp(0, 1).double() - tgt.double()) ** 2).mean()
    assert float(loss) == pytest.approx(float(want), rel=1e-6)


def test_fused_no_intersections(device):
    """I < 1 (every Gaussian off screen): the reference renders ones; the loss
    is that of the ones image and every gradient is zero."""
    from lig_amd import fit
    H, W = 64, 64
    a, b = _level_pair(50, H, W, 2, 1.0, device)
    with torch.no_grad():
        a.means.add_(1e4)
        b.means.add_(1e4)
    tgt = torch.rand(H, W, 3, device=device)
    la, lb = a.loss(tgt), b.loss(tgt)
    assert float(la.detach()) == pytest.approx(float(lb.detach()), rel=1e-6)
    la.backward()
    lb.backward()
    for name in ("means", "cholesky", "rgbs"):
        assert torch.count_nonzero(getattr(a, name).grad) == 0
        assert torch.equal(getattr(a, name).grad, getattr(b, name).grad)


def test_batch_depth_changes_speed_not_bits(device, tmp_path, monkeypatch):
    """The forward's LDS batch depth (64 / 128 / 256) gives the same image,
    loss and gradients; the autotuner picks one of them."""
    from lig_amd import autotune, data, fit
    monkeypatch.setenv("LIG_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    H, W = 120, 150
    tgt = data.make_image(H, W, 9, device=device)
    outs = []
    for d in (64, 128, 256):
        lv = fit.Level(9000, H, W, 5, 3.0, device=device)   # deep tiles: several batches
        lv.batch_depth = d
        loss = lv.loss(tgt)
        loss.backward()
        outs.append((float(loss.detach()), lv.means.grad.clone(), lv.cholesky.grad.clone(),
                     lv.rgbs.grad.clone()))
    for o in outs[1:]:
        assert o[0] == outs[0][0]
        for a, b in zip(o[1:], outs[0][1:]):
            assert torch.equal(a, b)
    lv = fit.Level(9000, H, W, 5, 3.0, device=device)
    res = autotune.tune_level(lv, tgt, reps=2, use_cache=False)
    assert res["choice"] in (64, 128, 256) and lv.batch_depth == res["choice"]
    assert set(res["ms"]) == {64, 128, 256}


def test_level_graph_matches_eager_steps(device):
    """Level.graph(): the fused step replayed as a hipGraph (a batch of one)
    gives the parameters and optimiser state of the same number of eager
    Level.step calls, bit for bit, with eager steps interleaved."""
    from lig_amd import data, fit
    H, W = 96, 136
    tgt = data.make_image(H, W, 5, device=device)
    a = fit.Level(900, H, W, 7, 1.5, lr=1e-2, device=device)
    b = fit.Level(900, H, W, 7, 1.5, lr=1e-2, device=device)
    gb = b.graph(tgt)
    for _ in range(3):
        a.step(tgt)
    gb.run(3)
    a.step(tgt)
    b.step(tgt)
    for _ in range(2):
        a.step(tgt)
    gb.run(2)
    for name in ("means", "cholesky", "rgbs"):
        assert torch.equal(getattr(a, name), getattr(b, name)), name
        pa, pb = getattr(a, name), getattr(b, name)
        for key in ("exp_avg", "exp_avg_sq"):
            assert torch.equal(a.optimizer.state[pa][key], b.optimizer.state[pb][key])
        assert a.optimizer.state[pa]["step"] == b.optimizer.state[pb]["step"] == 6


def test_tile_size_tuning(device, tmp_path, monkeypatch):
    """tune_level_tiles times 16 x 16 (fused, each depth) and 8 x 8 (API path)
    steps and keeps the fastest; the image does not depend on the tile size
    and the gradients agree to FP32 summation order."""
    from lig_amd import autotune, data, fit
    monkeypatch.setenv("LIG_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    monkeypatch.setattr(autotune, "_cache", {})
    H, W = 96, 112
    tgt = data.make_image(H, W, 4, device=device)
    grads = []
    for fused, bw in ((True, 16), (False, 8)):
        lv = fit.Level(3000, H, W, 6, 1.5, device=device, block_width=bw, fused=fused)
        assert lv.fused == fused
        with torch.no_grad():
            img = lv.render()
        loss = lv.loss(tgt)
        loss.backward()
        grads.append((img, [p.grad.clone() for p in (lv.means, lv.cholesky, lv.rgbs)]))
    assert torch.equal(grads[0][0], grads[1][0])
    for a, b in zip(grads[0][1], grads[1][1]):
        torch.testing.assert_close(a, b, rtol=1e-4, atol=1e-6 * float(b.abs().max()))
    lv = fit.Level(3000, H, W, 6, 1.5, device=device)
    before = [p.detach().clone() for p in (lv.means, lv.cholesky, lv.rgbs)]
    res = autotune.tune_level_tiles(lv, tgt, reps=1)
    assert set(res["ms"]) == {"16/64", "16/128", "16/256", "8/api"}
    assert res["ms"][res["choice"]] == min(res["ms"].values())
    assert (lv.block_width, lv.fused) == ((8, False) if res["choice"] == "8/api" else (16, True))
    for a, b in zip(before, (lv.means, lv.cholesky, lv.rgbs)):
        assert torch.equal(a, b)   # timing runs take no optimiser step
    again = autotune.tune_level_tiles(lv, tgt, reps=1)
    assert again["cached"] and again["choice"] == res["choice"]
    lv.step(tgt)


def test_fit_level_captured_equals_eager(device):
    """fit.fit_level (captured steps, as two_level_fit runs them) leaves the
    parameters of the same number of eager steps, bit for bit."""
    from lig_amd import data, fit
    gt = data.make_image(64, 80, 2, device=device)
    a = fit.Level(500, 64, 80, 3, 1.0, lr=1e-2, device=device)
    b = fit.Level(500, 64, 80, 3, 1.0, lr=1e-2, device=device)
    for _ in range(7):
        a.step(gt)
    fit.fit_level(b, gt, 7)
    for name in ("means", "cholesky", "rgbs"):
        assert torch.equal(getattr(a, name), getattr(b, name)), name
