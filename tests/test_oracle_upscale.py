"""Pins the oracle's gradient-aware upscaler (gradient_aware_upscaler.cu:137-371):
reproduction of bicubic polynomials from exact derivatives, corner values at
integer-aligned samples, and the adjoint identity <y, A x> = <A^T y, x>."""
import numpy as np
import pytest

import oracle


def _poly_field(coef, H, W):
    """Bicubic polynomial p(x, y) = sum a_ij x^i y^j sampled at source pixel
    k (the upscaler's corner k sits at coordinate k), with exact derivatives."""
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    f = np.zeros((H, W, 3))
    fx = np.zeros((H, W, 3))
    fy = np.zeros((H, W, 3))
    fxy = np.zeros((H, W, 3))
    for ch in range(3):
        for i in range(4):
            for j in range(4):
                a = coef[ch, i, j]
                f[..., ch] += a * xs ** i * ys ** j
                if i:
                    fx[..., ch] += a * i * xs ** (i - 1) * ys ** j
                if j:
                    fy[..., ch] += a * j * xs ** i * ys ** (j - 1)
                if i and j:
                    fxy[..., ch] += a * i * j * xs ** (i - 1) * ys ** (j - 1)
    return f, fx, fy, fxy


def _eval_poly(coef, x, y):
    out = np.zeros(x.shape + (3,))
    for ch in range(3):
        for i in range(4):
            for j in range(4):
                out[..., ch] += coef[ch, i, j] * x ** i * y ** j
    return out


@pytest.mark.parametrize("sh,sw,dh,dw", [(8, 8, 32, 32), (6, 9, 17, 40), (10, 10, 10, 10)])
def test_reproduces_bicubic_polynomials(sh, sw, dh, dw):
    rng = np.random.default_rng(sh * dw)
    coef = rng.normal(size=(3, 4, 4)) * 0.05 ** np.add.outer(np.arange(4), np.arange(4))
    f, fx, fy, fxy = _poly_field(coef, sh, sw)
    out = oracle.upscale_forward(f, fx, fy, fxy, dh, dw)
    # destination centre -> source coordinate (half-pixel grid)
    sx = (np.arange(dw) + 0.5) * sw / dw - 0.5
    sy = (np.arange(dh) + 0.5) * sh / dh - 0.5
    X, Y = np.meshgrid(sx, sy)
    interior = (X >= 0) & (X <= sw - 1) & (Y >= 0) & (Y <= sh - 1)
    ref = _eval_poly(coef, X, Y)
    # inputs are rounded to FP32 by the oracle's interface
    np.testing.assert_allclose(out[interior], ref[interior], rtol=1e-5, atol=1e-5)


def test_integer_aligned_samples_return_corner_values():
    rng = np.random.default_rng(0)
    f, fx, fy, fxy = (rng.normal(size=(7, 5, 3)).astype(np.float32) for _ in range(4))
    # dst == src with a shifted ROI of the same size maps d -> d + roi0 exactly
    out = oracle.upscale_forward(f, fx, fy, fxy, 4, 3, roi=(1.0, 2.0, 4.0, 6.0))
    np.testing.assert_allclose(out, f[2:6, 1:4].astype(np.float64), atol=0)


@pytest.mark.parametrize("sh,sw,dh,dw,roi", [(9, 7, 30, 22, None), (12, 12, 5, 7, None),
                                              (16, 16, 24, 40, (2.0, 3.0, 12.5, 14.0))])
def test_backward_is_adjoint(sh, sw, dh, dw, roi):
    rng = np.random.default_rng(1)
    xs = [rng.normal(size=(sh, sw, 3)).astype(np.float32) for _ in range(4)]
    y = rng.normal(size=(dh, dw, 3)).astype(np.float32)
    Ax = oracle.upscale_forward(*xs, dh, dw, roi)
    ATy = oracle.upscale_backward(y, sh, sw, roi)
    lhs = float((y.astype(np.float64) * Ax).sum())
    rhs = sum(float((a * x.astype(np.float64)).sum()) for a, x in zip(ATy, xs))
    assert abs(lhs - rhs) <= 1e-9 * max(1.0, abs(lhs))


def test_hermite_partition_of_unity():
    """Constant fields (zero derivatives) upscale to the same constant."""
    f = np.full((5, 6, 3), 0.37, np.float32)
    z = np.zeros_like(f)
    out = oracle.upscale_forward(f, z, z, z, 13, 19)
    np.testing.assert_allclose(out, 0.37, atol=1e-7)
