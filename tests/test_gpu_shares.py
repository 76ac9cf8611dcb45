"""The 16 x 16 backward split over shares (several blocks per tile, each
taking an equal share of the tile's 64-Gaussian batches) gives the same bits
for every share count: the fit step's records, the API records with and
without opacities, and the batched grid launch.  Each record has its own
emission slot, so only the launch shape changes; counts above a tile's batch
count leave some shares empty."""
import numpy as np
import pytest
import torch

from tests.scenes import random_scene

pytestmark = pytest.mark.gpu

SHARE_COUNTS = (1, 2, 3, 8, 0)   # 0: the library's own choice


@pytest.fixture
def shares():
    import lig_amd._lib as L
    lib = L.load()
    prev = lib.lig_set_backward_shares(0)

    def force(k):
        assert lib.lig_set_backward_shares(k) >= 0

    yield force
    lib.lig_set_backward_shares(prev)


def _t(a, dev, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(dev)


def test_fit_backward_same_bits_for_any_share_count(device, shares):
    from lig_amd import data, fit
    H, W = 120, 150
    tgt = data.make_image(H, W, 9, device=device)
    lv = fit.Level(9000, H, W, 5, 3.0, device=device)
    with torch.no_grad():
        lv.rgbs.mul_(0.02)   # keep the image inside (0, 1), where the clamp passes gradients
        nt = lv.project()[3]
    print(f"fit: {int(nt.sum())} intersections over 80 tiles", flush=True)
    ref = None
    for k in SHARE_COUNTS:
        shares(k)
        for p in (lv.means, lv.cholesky, lv.rgbs):
            p.grad = None
        lv.loss(tgt).backward()
        got = [p.grad.clone() for p in (lv.means, lv.cholesky, lv.rgbs)]
        if ref is None:
            ref = got
            assert all(torch.count_nonzero(g) > 0 for g in got)
        for name, a, b in zip(("means", "cholesky", "rgbs"), got, ref):
            assert torch.equal(a, b), (k, name)


@pytest.mark.parametrize("opac", [False, True])
def test_api_backward_same_bits_for_any_share_count(device, oracle_lib, shares, opac):
    import lig_amd.gsplat2d as g
    n, H, W = 3000, 70, 90
    s = random_scene(n, H, W, 31, opacities=opac, scale=(1.0, 8.0), margin=10.0)
    xys, ext, con, nt = oracle_lib.project_gaussians_forward(
        s["par"], s["means"], H, W, 16, s["opac"] if opac else None, cholesky=s["cholesky"])
    print(f"api opac={opac}: {int(nt.sum())} intersections over 30 tiles", flush=True)
    w = _t(np.random.default_rng(4).normal(size=(H, W, 3)), device)
    ref = None
    for k in SHARE_COUNTS:
        shares(k)
        xy = _t(xys, device).requires_grad_(True)
        cn = _t(con, device).requires_grad_(True)
        col = _t(s["colors"], device).requires_grad_(True)
        op = _t(s["opac"], device).requires_grad_(True) if opac else None
        img = g.rasterize_gaussians(xy, _t(ext, device), cn, _t(nt, device, torch.int32), col,
                                    op, H, W, 16)[0]
        (img * w).sum().backward()
        got = [xy.grad, cn.grad, col.grad] + ([op.grad] if opac else [])
        if ref is None:
            ref = got
            assert all(torch.count_nonzero(x) > 0 for x in got)
        for i, (a, b) in enumerate(zip(got, ref)):
            assert torch.equal(a, b), (k, i)


def test_batched_backward_same_bits_for_any_share_count(device, shares):
    from lig_amd.batch import BatchedFit, make_targets
    sizes = [(100, 130), (48, 40), (256, 200), (17, 300)]
    counts = [3000, 150, 9000, 300]
    seeds = [11, 12, 13, 14]
    targets = make_targets(sizes, seeds, device)
    ref = None
    for k in SHARE_COUNTS:
        shares(k)
        bf = BatchedFit(targets, counts, seeds, scale=2.5, lr=1e-2)
        with torch.no_grad():
            bf.rgbs.mul_(0.02)
        init = [t.detach().clone() for t in (bf.means, bf.cholesky, bf.rgbs)]
        for _ in range(2):
            bf.step()
        got = [t.detach().clone() for t in (bf.means, bf.cholesky, bf.rgbs)]
        if ref is None:
            ref = got
            assert all(not torch.equal(a, b) for a, b in zip(got, init))
        for i, (a, b) in enumerate(zip(got, ref)):
            assert torch.equal(a, b), (k, i)
