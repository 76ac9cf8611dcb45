"""GPU tests of batched fitting (lig_amd.batch): several images of uneven size
fitted in one launch per stage give every image the parameters, gradients and
loss its plain per-image fit (lig_amd.fit.Level, fused path) gives it."""
import pytest
import torch

pytestmark = pytest.mark.gpu

SIZES = [(100, 130), (48, 40), (256, 200), (17, 300), (64, 64), (33, 47)]
COUNTS = [600, 150, 2500, 300, 400, 90]
SEEDS = [11, 12, 13, 14, 15, 16]


def test_batched_matches_plain(device):
    from lig_amd.batch import BatchedFit, make_targets
    from lig_amd.fit import Level
    targets = make_targets(SIZES, SEEDS, device)
    bf = BatchedFit(targets, COUNTS, SEEDS, scale=1.5, lr=1e-2)
    levels = [Level(c, h, w, s, 1.5, lr=1e-2, device=device)
              for (h, w), c, s in zip(SIZES, COUNTS, SEEDS)]
    for step in range(3):
        bf.step()
        plain = [float(lv.step(t)) for lv, t in zip(levels, targets)]
        batched = bf.last_losses.tolist()
        for i, (a, b) in enumerate(zip(batched, plain)):
            assert a == pytest.approx(b, rel=1e-6), (step, i)
    o = 0
    for i, (lv, c) in enumerate(zip(levels, COUNTS)):
        for name in ("means", "cholesky", "rgbs"):
            assert torch.equal(getattr(bf, name)[o:o + c], getattr(lv, name)), (i, name)
        o += c
    psnr = bf.psnr()
    assert psnr.shape == (len(SIZES),) and (psnr > 0).all()
