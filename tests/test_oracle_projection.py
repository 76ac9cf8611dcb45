"""Pins the oracle's projection (helpers.cuh:47-93, forward.cu:13-105,
backward.cu:185-247) against closed forms, special cases and FP64 finite
differences -- not against a retyped copy of its own formula."""
import numpy as np
import pytest

import oracle


def _cov_scene(n, seed):
    rng = np.random.default_rng(seed)
    l11 = rng.uniform(0.5, 6, n)
    l22 = rng.uniform(0.5, 6, n)
    l21 = rng.uniform(-2, 2, n)
    cov = np.stack([l11 ** 2, l11 * l21, l21 ** 2 + l22 ** 2], 1)
    means = rng.uniform(-10, 110, (n, 2))
    return cov.astype(np.float32), means.astype(np.float32), \
        np.stack([l11, l21, l22], 1).astype(np.float32)


def test_conic_is_inverse_covariance():
    cov, means, _ = _cov_scene(500, 0)
    conics, ext = oracle.compute_cov2d_bounds(cov)
    c64 = cov.astype(np.float64)
    S = np.stack([np.stack([c64[:, 0], c64[:, 1]], -1), np.stack([c64[:, 1], c64[:, 2]], -1)], -2)
    Si = np.linalg.inv(S)
    ref = np.stack([Si[:, 0, 0], Si[:, 0, 1], Si[:, 1, 1]], 1)
    np.testing.assert_allclose(conics, ref, rtol=2e-5, atol=1e-7)


def test_extent_closed_form_and_coverage():
    """extent = ceil(sqrt(2 ln(255 o) * Sigma_ii)); every pixel with alpha >= 1/255
    lies inside |p - mu| <= extent."""
    cov, means, _ = _cov_scene(400, 1)
    rng = np.random.default_rng(2)
    opac = rng.uniform(0.01, 1.0, 400).astype(np.float32)
    conics, ext = oracle.compute_cov2d_bounds(cov, opac)
    t = 2 * np.log(255.0 * opac.astype(np.float64))
    rx = np.sqrt(t * cov[:, 0].astype(np.float64))
    ry = np.sqrt(t * cov[:, 2].astype(np.float64))
    ok = (np.abs(rx - np.round(rx)) > 1e-4) & (np.abs(ry - np.round(ry)) > 1e-4)
    np.testing.assert_array_equal(ext[ok, 0], np.ceil(rx[ok]))
    np.testing.assert_array_equal(ext[ok, 1], np.ceil(ry[ok]))
    # coverage: max |dx| over the ellipse d^T S^-1 d = t is sqrt(t S_xx)
    for g in range(0, 400, 37):
        S = np.array([[cov[g, 0], cov[g, 1]], [cov[g, 1], cov[g, 2]]], np.float64)
        Si = np.linalg.inv(S)
        ys, xs = np.mgrid[-60:61, -60:61] * 0.5
        d = np.stack([xs.ravel(), ys.ravel()], 1)
        sig = 0.5 * np.einsum("ni,ij,nj->n", d, Si, d)
        inside = opac[g] * np.exp(-sig) >= 1 / 255
        assert np.all(np.abs(d[inside, 0]) <= ext[g, 0] + 1e-6)
        assert np.all(np.abs(d[inside, 1]) <= ext[g, 1] + 1e-6)


def test_special_cases():
    H = W = 64
    cov = np.array([[1, 1, 1],        # det == 0 -> nothing written
                    [1, 2, 1],        # det < 0 (indefinite) accepted
                    [4, 0, 4],        # normal
                    [4, 0, 4],        # off screen
                    [4, 0, 4]],       # opacity below 1/255 -> t <= 0
                   np.float32)
    means = np.array([[10, 10], [20, 20], [32, 32], [-500, 30], [40, 40]], np.float32)
    opac = np.array([1, 1, 1, 1, 0.003], np.float32)
    xys, ext, con, nt = oracle.project_gaussians_forward(cov, means, H, W, 16, opac)
    assert np.all(con[0] == 0) and nt[0] == 0 and np.all(xys[0] == 0)
    # det = 1 - 4 = -3: conic = (c, -b, a)/det; t*c/inv_det = t*Sigma_xx stays
    # positive, so the extent is ceil(sqrt(2 ln 255)) = 4 and the Gaussian is binned
    np.testing.assert_allclose(con[1], np.array([1, -2, 1]) / -3.0, rtol=1e-6)
    assert nt[1] == 1 and np.all(ext[1] == 4)
    assert nt[2] > 0 and np.all(xys[2] == means[2]) and np.all(ext[2] > 0)
    assert nt[3] == 0 and np.all(xys[3] == 0) and np.all(ext[3] == 0)
    np.testing.assert_allclose(con[3], [0.25, 0, 0.25])   # conics written before the bbox test
    # phantom: zero extent but the centre's tile counts once (helpers.cuh:68-71)
    assert nt[4] == 1 and np.all(ext[4] == 0) and np.all(xys[4] == means[4])


def test_tile_bbox_rule():
    """Centres on tile edges: bbox = [trunc((mu - e)/bw), trunc((mu + e)/bw + 1))."""
    H = W = 128
    cov = np.array([[1.0, 0, 1.0]] * 4, np.float32)   # extent ceil(sqrt(2 ln 255)) = 4
    means = np.array([[32, 32], [31.99, 47.5], [0, 0], [127.9, 5]], np.float32)
    xys, ext, con, nt = oracle.project_gaussians_forward(cov, means, H, W, 16)
    np.testing.assert_array_equal(ext, 4.0)
    # x: (32-4)/16=1.75 -> 1, (32+4)/16+1=3.25 -> 3 : tiles 1,2 ; same in y -> 4
    assert nt[0] == 4
    # x: 27.99/16 -> 1, 35.99/16+1 -> 3 => 2 ; y: 43.5/16 -> 2, 51.5/16+1 -> 4 => 2
    assert nt[1] == 4
    # at the origin: min clamps to 0, max = trunc(4/16 + 1) = 1
    assert nt[2] == 1
    # x near the right edge: (127.9+4)/16+1 = 9.2 clamps to 8, min 7; y: [0, 1)
    assert nt[3] == 1


def test_cholesky_equals_covariance_path():
    cov, means, chol = _cov_scene(300, 3)
    # covariance computed in FP32 exactly as the Cholesky kernel does
    l11, l21, l22 = chol[:, 0], chol[:, 1], chol[:, 2]
    cov32 = np.stack([l11 * l11, l11 * l21, l21 * l21 + l22 * l22], 1).astype(np.float32)
    a = oracle.project_gaussians_forward(chol, means, 100, 90, 16, cholesky=True)
    b = oracle.project_gaussians_forward(cov32, means, 100, 90, 16)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def _conic64(cov):
    det = cov[..., 0] * cov[..., 2] - cov[..., 1] ** 2
    return np.stack([cov[..., 2] / det, -cov[..., 1] / det, cov[..., 0] / det], -1)


def test_projection_backward_matches_fd():
    """v_cov = J^T v_conic with J = d conic / d cov (FP64 central differences)."""
    cov, means, chol = _cov_scene(200, 4)
    rng = np.random.default_rng(5)
    v_con = rng.normal(size=(200, 3)).astype(np.float32)
    v_xy = rng.normal(size=(200, 2)).astype(np.float32)
    xys, ext, con, nt = oracle.project_gaussians_forward(cov, means, 100, 100, 16)
    v_cov, v_mean = oracle.project_gaussians_backward(ext, con, v_xy, v_con)
    on = (ext[:, 0] > 0) & (ext[:, 1] > 0)
    c64 = cov.astype(np.float64)
    fd = np.zeros((200, 3))
    h = 1e-6
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        fd[:, k] = ((_conic64(c64 + e) - _conic64(c64 - e)) / (2 * h) * v_con).sum(1)
    np.testing.assert_allclose(v_cov[on], fd[on], rtol=1e-3, atol=1e-5 * np.abs(fd).max())
    np.testing.assert_array_equal(v_mean[on], v_xy[on])
    assert np.all(v_cov[~on] == 0) and np.all(v_mean[~on] == 0)


def test_cholesky_backward_matches_fd():
    cov, means, chol = _cov_scene(200, 6)
    rng = np.random.default_rng(7)
    v_con = rng.normal(size=(200, 3)).astype(np.float32)
    v_xy = rng.normal(size=(200, 2)).astype(np.float32)
    xys, ext, con, nt = oracle.project_gaussians_forward(chol, means, 100, 100, 16,
                                                         cholesky=True)
    v_chol, v_mean = oracle.project_gaussians_backward(ext, con, v_xy, v_con, cholesky=chol)
    on = (ext[:, 0] > 0) & (ext[:, 1] > 0)

    def f(L):
        c = np.stack([L[:, 0] ** 2, L[:, 0] * L[:, 1], L[:, 1] ** 2 + L[:, 2] ** 2], 1)
        return (_conic64(c) * v_con).sum(1)

    L64 = chol.astype(np.float64)
    fd = np.zeros((200, 3))
    h = 1e-6
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        fd[:, k] = (f(L64 + e) - f(L64 - e)) / (2 * h)
    np.testing.assert_allclose(v_chol[on], fd[on], rtol=1e-3, atol=1e-5 * np.abs(fd).max())


def test_cholesky_backward_nonfinite_zeroed():
    chol = np.array([[np.inf, 0, 1], [1, 0, 1]], np.float32)
    ext = np.ones((2, 2), np.float32)
    con = np.array([[1, 0, 1], [1, 0, 1]], np.float32)
    v, m = oracle.project_gaussians_backward(ext, con, np.ones((2, 2), np.float32),
                                             np.ones((2, 3), np.float32), cholesky=chol)
    assert np.all(v[0] == 0) and np.all(np.isfinite(v[1])) and np.any(v[1] != 0)
