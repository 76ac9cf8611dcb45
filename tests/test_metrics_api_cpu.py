"""Argument checks of lig_amd.metrics that run before any device call, and
the window taps against the oracle's."""
import numpy as np
import pytest
import torch

from oracle import ssim as O


def test_window_taps_match_oracle():
    from lig_amd.metrics import gauss_window
    for size, sigma in ((11, 1.5), (7, 1.0), (3, 0.7)):
        np.testing.assert_allclose(gauss_window(size, sigma).numpy(), O.gauss_window(size, sigma),
                                   rtol=1e-6, atol=0)


def test_argument_errors_before_device_use():
    from lig_amd import metrics as M
    X = torch.rand(1, 3, 200, 200)
    with pytest.raises(ValueError):
        M.ms_ssim(X, X[:, :2])
    with pytest.raises(ValueError):
        M.ssim(X[0], X[0])
    with pytest.raises(ValueError):
        M.ms_ssim(X, X.double())
    with pytest.raises(ValueError):
        M.ms_ssim(X, X, win_size=10)
    with pytest.raises(AssertionError):
        M.ms_ssim(X[..., :160], X[..., :160])
    with pytest.raises(NotImplementedError):
        M.ms_ssim(X.clone().requires_grad_(True), X)
    with pytest.raises(RuntimeError, match="CUDA"):
        M.ms_ssim(X, X)
