"""Optimisers whose per-tensor update is one fused HIP kernel.

* FusedAdam: torch.optim.Adam's update (no amsgrad), weight decay added to the
  gradient; state 'step', 'exp_avg', 'exp_avg_sq' as torch's.
* Adan: the reference's Adan (optimizer.py:39-235: constructor arguments,
  parameter-group state 'step', per-parameter state 'exp_avg', 'exp_avg_sq',
  'exp_avg_diff', 'neg_pre_grad', global-norm clipping when max_grad_norm > 0,
  no_prox) with the single-tensor algorithm of optimizer.py:238-293 run by
  lig_adan_step.  Unlike the reference, p.grad is not scaled in place by the
  clipping factor.
"""
from __future__ import annotations

import torch
from torch.optim.optimizer import Optimizer

from . import _lib as L


def _stream(t):
    return L.stream_of(t)


class FusedAdam(Optimizer):
    def __init__(self, params, lr=1e-3, betas=(0.9, 0.999), eps=1e-8, weight_decay=0.0):
        if not 0.0 <= lr:
            raise ValueError(f"Invalid learning rate: {lr}")
        if not 0.0 <= eps:
            raise ValueError(f"Invalid epsilon value: {eps}")
        for i, b in enumerate(betas):
            if not 0.0 <= b < 1.0:
                raise ValueError(f"Invalid beta parameter at index {i}: {b}")
        super().__init__(params, dict(lr=lr, betas=betas, eps=eps, weight_decay=weight_decay))

    @torch.no_grad()
    def step(self, closure=None):
        loss = None
        if closure is not None:
            with torch.enable_grad():
                loss = closure()
        for group in self.param_groups:
            b1, b2 = group["betas"]
            for p in group["params"]:
                if p.grad is None:
                    continue
                L.check_device("param", p, torch.float32)
                st = self.state[p]
                if not st:
                    st["step"] = 0
                    st["exp_avg"] = torch.zeros_like(p, memory_format=torch.contiguous_format)
                    st["exp_avg_sq"] = torch.zeros_like(p, memory_format=torch.contiguous_format)
                st["step"] += 1
                g = p.grad.contiguous()
                with torch.cuda.device(p.device):
                    L.call("lig_adam_step", p.numel(), L.ptr(p), L.ptr(g), L.ptr(st["exp_avg"]),
                           L.ptr(st["exp_avg_sq"]), group["lr"], b1, b2, group["eps"],
                           group["weight_decay"], int(st["step"]), _stream(p))
        return loss


class Adan(Optimizer):
    def __init__(self, params, lr=1e-3, betas=(0.98, 0.92, 0.99), eps=1e-8, weight_decay=0.0,
                 max_grad_norm=0.0, no_prox=False, foreach: bool = True, fused: bool = True):
        if not 0.0 <= max_grad_norm:
            raise ValueError(f"Invalid Max grad norm: {max_grad_norm}")
        if not 0.0 <= lr:
            raise ValueError(f"Invalid learning rate: {lr}")
        if not 0.0 <= eps:
            raise ValueError(f"Invalid epsilon value: {eps}")
        for i, b in enumerate(betas):
            if not 0.0 <= b < 1.0:
                raise ValueError(f"Invalid beta parameter at index {i}: {b}")
        super().__init__(params, dict(lr=lr, betas=betas, eps=eps, weight_decay=weight_decay,
                                      max_grad_norm=max_grad_norm, no_prox=no_prox,
                                      foreach=foreach, fused=fused))

    def __setstate__(self, state):
        super().__setstate__(state)
        for group in self.param_groups:
            group.setdefault("no_prox", False)

    @torch.no_grad()
    def restart_opt(self):
        for group in self.param_groups:
            group["step"] = 0
            for p in group["params"]:
                if p.requires_grad:
                    st = self.state[p]
                    st["exp_avg"] = torch.zeros_like(p)
                    st["exp_avg_sq"] = torch.zeros_like(p)
                    st["exp_avg_diff"] = torch.zeros_like(p)

    @torch.no_grad()
    def step(self, closure=None):
        loss = None
        if closure is not None:
            with torch.enable_grad():
                loss = closure()
        clip = 1.0
        if self.defaults["max_grad_norm"] > 0:
            grads = [p.grad for g in self.param_groups for p in g["params"] if p.grad is not None]
            if grads:
                norm = torch.sqrt(sum(gr.pow(2).sum() for gr in grads))
                eps = self.param_groups[-1]["eps"]
                clip = float(torch.clamp(self.defaults["max_grad_norm"] / (norm + eps), max=1.0))
        for group in self.param_groups:
            b1, b2, b3 = group["betas"]
            group["step"] = group.get("step", 0) + 1
            step = int(group["step"])
            for p in group["params"]:
                if p.grad is None:
                    continue
                L.check_device("param", p, torch.float32)
                st = self.state[p]
                if "exp_avg" not in st:
                    st["exp_avg"] = torch.zeros_like(p)
                    st["exp_avg_sq"] = torch.zeros_like(p)
                    st["exp_avg_diff"] = torch.zeros_like(p)
                g = p.grad.contiguous()
                if "neg_pre_grad" not in st or step == 1:
                    # optimizer.py:171-173; on step 1 the kernel sets it itself
                    st["neg_pre_grad"] = (g * -clip if step > 1 else torch.empty_like(p))
                with torch.cuda.device(p.device):
                    L.call("lig_adan_step", p.numel(), L.ptr(p), L.ptr(g), L.ptr(st["exp_avg"]),
                           L.ptr(st["exp_avg_sq"]), L.ptr(st["exp_avg_diff"]),
                           L.ptr(st["neg_pre_grad"]), group["lr"], b1, b2, b3, group["eps"],
                           group["weight_decay"], int(bool(group["no_prox"])), clip, step,
                           _stream(p))
        return loss
