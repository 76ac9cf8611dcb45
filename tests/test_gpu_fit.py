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
