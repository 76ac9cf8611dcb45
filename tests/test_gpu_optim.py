"""GPU parity of the fused optimiser kernels (lig_adam_step, lig_adan_step)
through lig_amd.optim: Adam against torch.optim.Adam on the same device
(FP32, relative 1e-6), Adan against the FP64 oracle (relative 2e-6 of the
parameter scale), over several steps."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("wd", [0.0, 0.01])
@pytest.mark.parametrize("n", [1, 7, 1000, 4099])
def test_fused_adam_matches_torch(device, n, wd):
    from lig_amd.optim import FusedAdam
    torch.manual_seed(n)
    p0 = torch.randn(n, device=device)
    a = p0.clone().requires_grad_(True)
    b = p0.clone().requires_grad_(True)
    oa = FusedAdam([a], lr=1e-2, betas=(0.9, 0.99), eps=1e-8, weight_decay=wd)
    ob = torch.optim.Adam([b], lr=1e-2, betas=(0.9, 0.99), eps=1e-8, weight_decay=wd)
    for _ in range(5):
        g = torch.randn(n, device=device)
        a.grad, b.grad = g.clone(), g.clone()
        oa.step()
        ob.step()
    torch.testing.assert_close(a.detach(), b.detach(), rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("cfg", [
    dict(betas=(0.98, 0.92, 0.99), weight_decay=0.0, no_prox=False, max_grad_norm=0.0),
    dict(betas=(0.9, 0.8, 0.95), weight_decay=0.02, no_prox=False, max_grad_norm=0.0),
    dict(betas=(0.98, 0.92, 0.99), weight_decay=0.02, no_prox=True, max_grad_norm=0.0),
    dict(betas=(0.98, 0.92, 0.99), weight_decay=0.0, no_prox=False, max_grad_norm=0.5),
])
def test_adan_matches_oracle(device, cfg):
    from lig_amd.optim import Adan
    from oracle import optim as O
    rng = np.random.default_rng(5)
    shapes = [(1000, 2), (1000, 3), (33,)]
    ps = [torch.tensor(rng.normal(size=s), dtype=torch.float32, device=device)
          .requires_grad_(True) for s in shapes]
    ref = [p.detach().cpu().numpy().astype(np.float64) for p in ps]
    states = [{} for _ in ps]
    opt = Adan(ps, lr=2e-3, eps=1e-8, **cfg)
    for _ in range(5):
        gs = [rng.normal(size=s).astype(np.float32) for s in shapes]
        for p, g in zip(ps, gs):
            p.grad = torch.tensor(g, device=device)
        clip = 1.0
        if cfg["max_grad_norm"] > 0:
            norm = np.sqrt(sum(float((g.astype(np.float64) ** 2).sum()) for g in gs))
            clip = min(cfg["max_grad_norm"] / (norm + 1e-8), 1.0)
        opt.step()
        ref = [O.adan_step(r, g, st, lr=2e-3, betas=cfg["betas"], eps=1e-8,
                           weight_decay=cfg["weight_decay"], no_prox=cfg["no_prox"], clip=clip)
               for r, g, st in zip(ref, gs, states)]
    for p, r in zip(ps, ref):
        np.testing.assert_allclose(p.detach().cpu().numpy(), r, rtol=0, atol=2e-6)
