"""Fused clamp + L2 loss of the fit step on the HIP path.

`l2_loss_clamped(img, target)` equals `F.mse_loss(torch.clamp(img, 0, 1), target)`
(the caller's loss, gaussianlig.py:95-103 with utils.py:21-22) up to FP32
summation order, with the same gradient w.r.t. `img` (clamp mask inclusive at
both bounds).  The target is data: no gradient flows to it.
"""
from __future__ import annotations

import torch

from . import _lib as L


class _L2Clamped(torch.autograd.Function):
    @staticmethod
    def forward(ctx, img, target):
        img = img.contiguous()
        target = target.contiguous()
        for name, t in (("img", img), ("target", target)):
            L.check_device(name, t, torch.float32)
        if img.shape != target.shape:
            raise ValueError(f"img {tuple(img.shape)} and target {tuple(target.shape)} differ")
        n = img.numel()
        loss = torch.empty((), dtype=torch.float32, device=img.device)
        ws = torch.empty(L.size_query("lig_l2_loss_workspace_size", n), dtype=torch.uint8,
                         device=img.device)
        with torch.cuda.device(img.device):
            L.call("lig_l2_clamped_loss_forward", n, L.ptr(img), L.ptr(target), L.ptr(loss),
                   L.ptr(ws), ws.numel(), L.stream_of(img))
        ctx.save_for_backward(img, target)
        return loss

    @staticmethod
    def backward(ctx, grad_loss):
        img, target = ctx.saved_tensors
        g = grad_loss.reshape(()).to(torch.float32).contiguous()
        grad = torch.empty_like(img)
        with torch.cuda.device(img.device):
            L.call("lig_l2_clamped_loss_backward", img.numel(), L.ptr(img), L.ptr(target),
                   L.ptr(g), L.ptr(grad), L.stream_of(img))
        return grad, None


def l2_loss_clamped(img: torch.Tensor, target: torch.Tensor) -> torch.Tensor:
    return _L2Clamped.apply(img, target)
