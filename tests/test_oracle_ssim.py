"""Pins of the FP64 SSIM / MS-SSIM oracle (oracle/ssim.py) by library
routines, brute force and closed forms."""
import numpy as np
import pytest
import torch
from scipy.signal import correlate2d

from oracle import ssim as O


def test_window_closed_form():
    g = O.gauss_window(11, 1.5)
    assert g.shape == (11,)
    assert abs(g.sum() - 1.0) < 1e-15
    np.testing.assert_allclose(g, g[::-1], rtol=0, atol=0)
    k = np.arange(11) - 5
    np.testing.assert_allclose(g / g[5], np.exp(-k ** 2 / 4.5), rtol=1e-14)
    with pytest.raises(ValueError):
        O.gauss_window(10, 1.5)


def test_filter_matches_scipy_correlate():
    rng = np.random.default_rng(0)
    img = rng.random((23, 31))
    for size, sigma in ((11, 1.5), (7, 1.0), (3, 0.8)):
        g = O.gauss_window(size, sigma)
        want = correlate2d(img, np.outer(g, g), mode="valid")
        np.testing.assert_allclose(O.filter_valid(img, g), want, rtol=1e-12, atol=1e-14)
    # an asymmetric window catches a flipped (convolution) orientation
    g = np.array([0.1, 0.2, 0.7])
    np.testing.assert_allclose(O.filter_valid(img, g), correlate2d(img, np.outer(g, g), "valid"),
                               rtol=1e-12)


def test_maps_match_brute_force_windows():
    rng = np.random.default_rng(1)
    x = rng.random((14, 13))
    y = np.clip(x + 0.2 * rng.standard_normal(x.shape), 0, 1)
    g = O.gauss_window()
    w2 = np.outer(g, g)
    R = 2.0
    C1, C2 = (0.01 * R) ** 2, (0.03 * R) ** 2
    ss, cs = O.ssim_maps(x * R, y * R, R, g)
    assert ss.shape == (4, 3)
    for i in range(4):
        for j in range(3):
            a = x[i:i + 11, j:j + 11] * R
            b = y[i:i + 11, j:j + 11] * R
            ma, mb = (w2 * a).sum(), (w2 * b).sum()
            va = (w2 * (a - ma) ** 2).sum()
            vb = (w2 * (b - mb) ** 2).sum()
            cab = (w2 * (a - ma) * (b - mb)).sum()
            c = (2 * cab + C2) / (va + vb + C2)
            s = (2 * ma * mb + C1) / (ma * ma + mb * mb + C1) * c
            assert abs(cs[i, j] - c) < 1e-12
            assert abs(ss[i, j] - s) < 1e-12


def test_constant_images_closed_form():
    a, b = 0.3, 0.55
    C1 = 0.01 ** 2
    want = (2 * a * b + C1) / (a * a + b * b + C1)
    x = np.full((1, 2, 20, 24), a)
    y = np.full((1, 2, 20, 24), b)
    ss, cs = O.ssim_cs(x, y)
    np.testing.assert_allclose(cs, 1.0, rtol=1e-12)
    np.testing.assert_allclose(ss, want, rtol=1e-12)
    # 176 x 192 halves without padding four times: every scale stays constant,
    # cs = 1 and only the last weight (on ssim) survives
    x = np.full((1, 1, 176, 192), a)
    y = np.full((1, 1, 176, 192), b)
    assert abs(O.ms_ssim(x, y) - want ** O.MS_WEIGHTS[-1]) < 1e-12
    # data range scales out
    assert abs(O.ms_ssim(x * 255, y * 255, data_range=255) - want ** 0.1333) < 1e-12


def test_identity_and_symmetry():
    rng = np.random.default_rng(2)
    x = rng.random((2, 3, 177, 181))
    y = rng.random((2, 3, 177, 181))
    assert abs(O.ms_ssim(x, x) - 1.0) < 1e-12
    assert abs(O.ssim(x, x) - 1.0) < 1e-12
    assert abs(O.ms_ssim(x, y) - O.ms_ssim(y, x)) < 1e-14
    assert O.ms_ssim(x, y, size_average=False).shape == (2,)


@pytest.mark.parametrize("shape", [(1, 3, 8, 8), (2, 1, 9, 7), (1, 2, 5, 12), (1, 1, 11, 11)])
def test_pool_matches_torch_avg_pool2d(shape):
    rng = np.random.default_rng(3)
    x = rng.random(shape)
    t = torch.from_numpy(x)
    want = torch.nn.functional.avg_pool2d(t, kernel_size=2, stride=2,
                                          padding=(shape[2] % 2, shape[3] % 2),
                                          count_include_pad=True).numpy()
    np.testing.assert_allclose(O.avg_pool_pad(x), want, rtol=1e-15, atol=0)


def test_one_hot_weights_single_out_each_scale():
    rng = np.random.default_rng(4)
    x = rng.random((1, 2, 170, 165))
    y = np.clip(x + 0.1 * rng.standard_normal(x.shape), 0, 1)
    xs, ys = x, y
    for lvl in range(5):
        ss, cs = O.ssim_cs(xs, ys)
        w = np.zeros(5)
        w[lvl] = 1.0
        got = O.ms_ssim_planes(x, y, weights=w)
        want = np.maximum(ss if lvl == 4 else cs, 0.0)
        np.testing.assert_allclose(got, want, rtol=1e-12)
        xs, ys = O.avg_pool_pad(xs), O.avg_pool_pad(ys)
