"""Pins of the optimiser oracle (oracle/optim.py) on the CPU."""
import numpy as np
import pytest
import torch

from oracle import optim as O


def _torch_adam(p0, grads, **kw):
    p = torch.tensor(p0, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([p], **kw)
    out = []
    for g in grads:
        p.grad = torch.tensor(g, dtype=torch.float64)
        opt.step()
        out.append(p.detach().numpy().copy())
    return out


@pytest.mark.parametrize("wd", [0.0, 0.05])
def test_adam_oracle_matches_torch(wd):
    rng = np.random.default_rng(0)
    p = rng.normal(size=257)
    grads = [rng.normal(size=257) for _ in range(6)]
    ref = _torch_adam(p, grads, lr=1e-2, betas=(0.9, 0.99), eps=1e-8, weight_decay=wd)
    st = {}
    for g, r in zip(grads, ref):
        p = O.adam_step(p, g, st, lr=1e-2, betas=(0.9, 0.99), eps=1e-8, weight_decay=wd)
        np.testing.assert_allclose(p, r, rtol=1e-12, atol=1e-14)


def test_adan_with_beta2_zero_is_adam():
    """beta2 = 0: diff_t = d_t contributes lr * 0 / 1, u_t = g_t, so Adan's
    update is Adam's with betas (beta1, beta3)."""
    rng = np.random.default_rng(1)
    p = rng.normal(size=300)
    grads = [rng.normal(size=300) for _ in range(7)]
    ref = _torch_adam(p, grads, lr=3e-3, betas=(0.98, 0.99), eps=1e-8)
    st = {}
    for g, r in zip(grads, ref):
        p = O.adan_step(p, g, st, lr=3e-3, betas=(0.98, 0.0, 0.99), eps=1e-8)
        np.testing.assert_allclose(p, r, rtol=1e-12, atol=1e-14)


def test_adan_first_step_and_constant_gradient():
    """d_1 = 0 and, for a constant gradient, d_t = 0, m_t = g (1 - b1^t),
    n_t = g^2 (1 - b3^t): every step moves p by lr * g / (|g| + eps)."""
    rng = np.random.default_rng(2)
    p0 = rng.normal(size=100)
    g = rng.normal(size=100)
    lr, eps = 1e-2, 1e-6
    st = {}
    p = p0
    for t in range(1, 6):
        p = O.adan_step(p, g, st, lr=lr, betas=(0.98, 0.92, 0.99), eps=eps)
        np.testing.assert_allclose(p, p0 - t * lr * g / (np.abs(g) + eps), rtol=1e-12,
                                   atol=1e-14)


@pytest.mark.parametrize("no_prox", [False, True])
def test_adan_weight_decay_first_step(no_prox):
    rng = np.random.default_rng(3)
    p0 = rng.normal(size=64)
    g = rng.normal(size=64)
    lr, wd, eps = 1e-2, 0.1, 1e-8
    p = O.adan_step(p0, g, {}, lr=lr, weight_decay=wd, eps=eps, no_prox=no_prox)
    step = lr * g / (np.abs(g) + eps)
    want = p0 * (1 - lr * wd) - step if no_prox else (p0 - step) / (1 + lr * wd)
    np.testing.assert_allclose(p, want, rtol=1e-12, atol=1e-14)


def test_adan_clip_scales_gradient():
    """The clipping factor multiplies the gradient before everything else: a
    clipped run equals an unclipped run on scaled gradients."""
    rng = np.random.default_rng(4)
    p0 = rng.normal(size=50)
    grads = [rng.normal(size=50) for _ in range(4)]
    a, b = {}, {}
    pa = pb = p0
    for g in grads:
        pa = O.adan_step(pa, g, a, clip=0.3)
        pb = O.adan_step(pb, 0.3 * g, b)
    np.testing.assert_allclose(pa, pb, rtol=1e-14)
