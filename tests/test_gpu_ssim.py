"""GPU parity of lig_ssim_multiscale (lig_amd.metrics.ms_ssim / ssim) with the
FP64 oracle (oracle/ssim.py).

Tolerances: per-scale plane means of ssim and cs within 2e-5 absolute (FP32
window sums with FP64 accumulation of the maps), MS-SSIM within 2e-5.  Exact
properties: MS-SSIM of an image with itself is 1.0, and a strided (NHWC) view
gives the same bits as the contiguous NCHW copy."""
import numpy as np
import pytest
import torch

from oracle import ssim as O

pytestmark = pytest.mark.gpu

TOL = 2e-5


def pair(shape, seed, noise=0.15):
    rng = np.random.default_rng(seed)
    x = rng.random(shape).astype(np.float32)
    # smooth structure plus noise so that the scales differ
    x = (0.5 * x + 0.5 * np.cumsum(x, axis=-1) / np.arange(1, shape[-1] + 1)).astype(np.float32)
    y = np.clip(x + noise * rng.standard_normal(shape), 0, 1).astype(np.float32)
    return x, y


@pytest.mark.parametrize("shape", [(1, 3, 176, 192), (2, 3, 181, 173), (1, 1, 300, 700),
                                   (1, 2, 165, 409)])
def test_ms_ssim_matches_oracle(device, shape):
    from lig_amd import metrics as M
    x, y = pair(shape, sum(shape))
    X, Y = torch.from_numpy(x).to(device), torch.from_numpy(y).to(device)
    got = M.ms_ssim(X, Y, data_range=1).item()
    want = O.ms_ssim(x, y, data_range=1)
    print(f"shape {shape}: hip {got:.8f} oracle {want:.8f}", flush=True)
    assert abs(got - want) < TOL
    per = M.ms_ssim(X, Y, data_range=1, size_average=False).cpu().numpy()
    np.testing.assert_allclose(per, O.ms_ssim(x, y, 1, size_average=False), atol=TOL, rtol=0)
    # every scale's plane means
    _, ss = M._run(X, Y, 1.0, M.gauss_window(), (0.01, 0.03), M.MS_WEIGHTS)
    xs, ys = x.astype(np.float64), y.astype(np.float64)
    for lvl in range(5):
        s_want, _ = O.ssim_cs(xs, ys, 1.0)
        np.testing.assert_allclose(ss[lvl].cpu().numpy(), s_want, atol=TOL, rtol=0)
        xs, ys = O.avg_pool_pad(xs), O.avg_pool_pad(ys)


@pytest.mark.parametrize("win_size,sigma", [(11, 1.5), (7, 1.0), (3, 0.7), (9, 2.0)])
def test_ssim_matches_oracle(device, win_size, sigma):
    from lig_amd import metrics as M
    x, y = pair((2, 3, 70, 133), win_size)
    X, Y = torch.from_numpy(x).to(device), torch.from_numpy(y).to(device)
    got = M.ssim(X, Y, data_range=1, win_size=win_size, win_sigma=sigma,
                 size_average=False).cpu().numpy()
    want = O.ssim(x, y, 1, size_average=False, win_size=win_size, win_sigma=sigma)
    np.testing.assert_allclose(got, want, atol=TOL, rtol=0)


def test_ssim_custom_window_and_nonnegative(device):
    from lig_amd import metrics as M
    x, _ = pair((1, 3, 64, 64), 9)
    y = (1.0 - x).astype(np.float32)          # anti-correlated: negative ssim
    X, Y = torch.from_numpy(x).to(device), torch.from_numpy(y).to(device)
    raw = M.ssim(X, Y, data_range=1).item()
    want = O.ssim(x, y, 1)
    assert want < 0 and abs(raw - want) < TOL
    nn_val = M.ssim(X, Y, data_range=1, nonnegative_ssim=True, size_average=False)
    assert torch.all(nn_val >= 0)
    g = torch.tensor([0.1, 0.2, 0.4, 0.2, 0.1])
    win = g.view(1, 1, 1, 5).repeat(3, 1, 1, 1)
    got = M.ssim(X, Y, data_range=1, win=win).item()
    ss, _ = O.ssim_cs(x, y, 1.0, g.double().numpy())
    assert abs(got - ss.mean()) < TOL


def test_ms_ssim_nhwc_view_bit_identical(device):
    from lig_amd import metrics as M
    x, y = pair((2, 3, 190, 200), 4)
    X = torch.from_numpy(x).to(device)
    Y = torch.from_numpy(y).to(device)
    Xh = X.permute(0, 2, 3, 1).contiguous().permute(0, 3, 1, 2)   # NHWC storage
    assert not Xh.is_contiguous()
    a = M.ms_ssim(X, Y, data_range=1, size_average=False)
    b = M.ms_ssim(Xh, Y, data_range=1, size_average=False)
    assert torch.equal(a, b)


def test_ms_ssim_closed_forms(device):
    from lig_amd import metrics as M
    x, y = pair((1, 3, 181, 207), 11)
    X, Y = torch.from_numpy(x).to(device), torch.from_numpy(y).to(device)
    assert M.ms_ssim(X, X, data_range=1).item() == 1.0
    assert abs(M.ms_ssim(X, Y, 1).item() - M.ms_ssim(Y, X, 1).item()) < 1e-6
    # constants: cs = 1 at every scale, only the last weight survives
    a, b = 0.3, 0.55
    A = torch.full((1, 3, 176, 192), a, device=device)
    B = torch.full((1, 3, 176, 192), b, device=device)
    C1 = 1e-4
    want = ((2 * a * b + C1) / (a * a + b * b + C1)) ** 0.1333
    assert abs(M.ms_ssim(A, B, data_range=1).item() - want) < 1e-6
    # the data range scales out
    s1 = M.ms_ssim(X, Y, data_range=1).item()
    s255 = M.ms_ssim(X * 255, Y * 255, data_range=255).item()
    assert abs(s1 - s255) < 1e-5


def test_ms_ssim_full_size_properties(device):
    """8192 x 8192 x 3 (the two-level fit's image size): identity is exactly 1
    and a constant pair gives the closed form."""
    from lig_amd import metrics as M
    g = torch.Generator(device=device).manual_seed(0)
    X = torch.rand(1, 3, 8192, 8192, device=device, generator=g)
    assert M.ms_ssim(X, X, data_range=1).item() == 1.0
    Y = torch.clamp(X + 0.1 * torch.randn(X.shape, device=device, generator=g), 0, 1)
    v = M.ms_ssim(X, Y, data_range=1).item()
    assert 0.0 < v < 1.0
    A = torch.full_like(X, 0.25)
    B = torch.full_like(X, 0.75)
    want = ((2 * 0.25 * 0.75 + 1e-4) / (0.25 ** 2 + 0.75 ** 2 + 1e-4)) ** 0.1333
    assert abs(M.ms_ssim(A, B, data_range=1).item() - want) < 1e-6


def test_ms_ssim_errors(device):
    from lig_amd import metrics as M
    X = torch.rand(1, 3, 200, 200, device=device)
    with pytest.raises(ValueError):
        M.ms_ssim(X, X[:, :2])
    with pytest.raises(ValueError):
        M.ms_ssim(X, X, win_size=10)
    with pytest.raises(AssertionError):
        M.ms_ssim(X[..., :160, :], X[..., :160, :])
    Xg = X.clone().requires_grad_(True)
    with pytest.raises(NotImplementedError):
        M.ms_ssim(Xg, X)


def test_evaluate_two_level_image(device):
    from lig_amd.fit import evaluate
    x, y = pair((1, 3, 256, 256), 21, noise=0.05)
    X = torch.from_numpy(x[0]).permute(1, 2, 0).contiguous().to(device)
    Y = torch.from_numpy(y[0]).permute(1, 2, 0).contiguous().to(device)
    p, m = evaluate(X, Y)
    mse = float(((x - y).astype(np.float64) ** 2).mean())
    assert abs(p - 10 * np.log10(1 / mse)) < 1e-3
    assert abs(m - O.ms_ssim(x, y, 1)) < TOL
