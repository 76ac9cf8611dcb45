"""Seeded random scenes shared by the tests."""
import numpy as np


def random_scene(n, H, W, seed, opacities=False, cholesky=True, scale=(0.5, 3.0),
                 offdiag=0.5, margin=0.0):
    """Means uniform over the image (plus margin), Cholesky diag ~U(scale),
    off-diagonal ~U(-offdiag, offdiag)*diag, colors ~U(0,1), opacities ~U(0.05, 0.95)."""
    rng = np.random.default_rng(seed)
    means = np.stack([rng.uniform(-margin, W + margin, n),
                      rng.uniform(-margin, H + margin, n)], 1).astype(np.float32)
    l11 = rng.uniform(*scale, n)
    l22 = rng.uniform(*scale, n)
    l21 = rng.uniform(-offdiag, offdiag, n) * np.sqrt(l11 * l22)
    chol = np.stack([l11, l21, l22], 1).astype(np.float32)
    colors = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    opac = rng.uniform(0.05, 0.95, n).astype(np.float32) if opacities else None
    if cholesky:
        par = chol
    else:
        a = l11 * l11
        b = l11 * l21
        c = l21 * l21 + l22 * l22
        par = np.stack([a, b, c], 1).astype(np.float32)
    return dict(par=par, means=means, colors=colors, opac=opac, H=H, W=W,
                cholesky=cholesky)
