"""GPU tests of batched crop rendering (render.CropBatch: one projection,
binning, raster and composition launch for many crops of several
representations) and of the server's batching mode: every crop gets exactly
the bits TwoLevel.render_crop gives it alone -- Hermite and bilinear
composition, Cholesky and covariance levels, scale 1 and 0.5, crops at the
canvas edge, a view with no intersections (the render is ones,
rasterize.py:80-90) and a representation with no Gaussians."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def _reps(device, tmp_path):
    import os
    from lig_amd import data, fit
    from lig_amd.render import LevelParams, TwoLevel
    from lig_amd.serve import make_representations
    reps = make_representations(2, seed=7, device=device, sizes=(512, 768),
                                n_range=(20_000, 60_000), scale_px=2.0)
    # covariance levels composed bilinearly, from a checkpoint in train.py's layout
    torch.manual_seed(5)
    model = fit.LIG(num_points=6000, H=256, W=320, n_scales=2, allo_ratio=0.5, device=device,
                    lr=1e-2)
    with torch.no_grad():
        for m in model.level_models:
            m.cov2d.mul_(6.0)
            m.rgbs.uniform_(0.0, 0.15)
    path = os.path.join(tmp_path, "ck.pth.tar")
    fit.save_checkpoint(path, model.level_models, [torch.tensor(0.5)], [torch.tensor(1.5)])
    reps.append(TwoLevel.from_checkpoint(path, 256, 320, device=device))
    # every Gaussian in the top-left 64 x 64 corner of a 512^2 canvas
    m1, c1, r1 = data.init_gaussians(500, 32, 32, 3, 1.0, device)
    m2, c2, r2 = data.init_gaussians(3000, 64, 64, 4, 1.0, device)
    reps.append(TwoLevel(LevelParams(m1, c1, r1 * 0.1, 128, 128),
                         LevelParams(m2, c2, r2 * 0.1, 512, 512), lo=-0.3, hi=1.2))
    # no Gaussians at all
    z2, z3 = torch.zeros(0, 2, device=device), torch.zeros(0, 3, device=device)
    reps.append(TwoLevel(LevelParams(z2, z3, z3, 128, 128), LevelParams(z2, z3, z3, 512, 512),
                         lo=0.0, hi=1.0))
    return reps


CROPS = [  # (rep, x0, y0, w, h, scale)
    (0, 0.0, 0.0, 128, 96, 1.0), (0, 100.0, 36.0, 64, 48, 0.5), (1, 200.0, 120.0, 256, 128, 1.0),
    (1, 0.0, 0.0, 384, 384, 0.5), (2, 0.0, 0.0, 64, 48, 1.0), (2, 100.0, 36.0, 128, 96, 1.0),
    (2, 40.0, 20.0, 64, 64, 0.5), (2, 288.0, 240.0, 32, 16, 1.0), (3, 0.0, 0.0, 96, 80, 1.0),
    (3, 384.0, 384.0, 128, 128, 1.0), (4, 16.0, 16.0, 64, 64, 1.0), (0, 496.0, 0.0, 16, 512, 1.0),
    (1, 752.0, 752.0, 16, 16, 1.0), (1, 4.0, 8.0, 200, 120, 1.0),
]


def test_crop_batch_equals_render_crop(device, tmp_path):
    from lig_amd.render import CropBatch, render_crops
    reps = _reps(device, tmp_path)
    with torch.no_grad():
        alone = [reps[r].render_crop(x, y, w, h, s) for r, x, y, w, h, s in CROPS]
        cb = CropBatch([(reps[r], x, y, w, h, s) for r, x, y, w, h, s in CROPS])
        assert cb.counts[CROPS.index((3, 384.0, 384.0, 128, 128, 1.0))] == 0
        assert cb.counts[CROPS.index((4, 16.0, 16.0, 64, 64, 1.0))] == 0
        assert sum(cb.counts) == cb.num_intersects > 0
        got = cb.finish()
        # single crops and a batch of repeats
        one = render_crops([(reps[2], 100.0, 36.0, 128, 96, 1.0)])
        rep3 = render_crops([(reps[0], 0.0, 0.0, 128, 96, 1.0)] * 3)
    torch.cuda.synchronize()
    for c, a, b in zip(CROPS, alone, got):
        assert b.shape == a.shape, c
        assert torch.equal(a, b), (c, float((a - b).abs().max()))
    assert torch.equal(one[0], alone[5])
    assert all(torch.equal(x, alone[0]) for x in rep3)
    # no Gaussians: both levels render ones, so the composition clamps
    # 1 * (hi - lo) + up(1) - 0.5 + lo = 1.5 (up to rounding) to exactly 1
    assert torch.equal(got[CROPS.index((4, 16.0, 16.0, 64, 64, 1.0))],
                       torch.ones(64, 64, 3, device=device))
    print("crop batch: intersections", cb.num_intersects, "tiles", cb.num_tiles,
          "HBM need", cb.need)


def test_crop_batch_rejects_bad_crops(device, tmp_path):
    from lig_amd.render import CropBatch
    reps = _reps(device, tmp_path)
    with pytest.raises(ValueError):
        CropBatch([])
    with pytest.raises(ValueError):
        CropBatch([(reps[0], 0.0, 0.0, 0, 16, 1.0)])
    with pytest.raises(ValueError):   # bilinear composition off the output pixel grid
        CropBatch([(reps[2], 0.5, 0.0, 16, 16, 1.0)])


@pytest.mark.parametrize("slots", [1, 2])
def test_server_batches_give_the_same_bits(device, tmp_path, slots):
    from lig_amd.serve import CropServer, Request
    reps = _reps(device, tmp_path)
    crops = CROPS * 3
    with torch.no_grad():
        alone = [reps[r].render_crop(x, y, w, h, s) for r, x, y, w, h, s in crops]
    torch.cuda.synchronize()
    srv = CropServer(reps, slots=slots, keep_results=True, device=device, max_batch=8)
    reqs = [Request(*c) for c in crops]
    st = srv.run_poisson(reqs, qps=1e5)
    assert st.requests == len(crops)
    for r, a in zip(reqs, alone):
        assert r.error is None and torch.equal(r.result, a), (r.rep, r.x0, r.y0, r.w, r.h)
    s = st.summary()
    print("batched serve summary", s)
    assert sum(st.batch_sizes) == len(crops) and max(st.batch_sizes) > 1
    assert s["max_batch_seen"] <= 8


def test_server_batched_admission(device, tmp_path):
    from lig_amd.serve import CropServer, Request
    reps = _reps(device, tmp_path)
    srv = CropServer(reps, slots=2, hbm_budget=1, device=device, max_batch=4)
    reqs = [Request(*c) for c in CROPS]
    st = srv.run_poisson(reqs, qps=1e5)
    assert st.peak_reserved == max(r.reserved for r in reqs) > 0
    # a bad request fails alone; the rest of its batch is served
    srv = CropServer(reps, slots=1, keep_results=True, device=device, max_batch=4)
    reqs = [Request(*CROPS[4]), Request(2, 0.5, 0.0, 16, 16, 1.0), Request(*CROPS[5])]
    srv.start()
    for r in reqs:
        srv.submit(r)
    srv.stop()
    assert isinstance(reqs[1].error, ValueError)
    assert reqs[0].error is None and reqs[2].error is None
    assert reqs[0].result is not None and reqs[2].result is not None
