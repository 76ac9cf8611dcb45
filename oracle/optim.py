"""Test infrastructure only: FP64 numpy restatements of the optimiser steps,
the checker for the HIP kernels in lig_amd/csrc/optim.hip.

* adam_step: torch.optim.Adam's documented update (no amsgrad, weight decay
  added to the gradient).  Pinned against torch.optim.Adam itself (FP64) in
  tests/test_oracle_optim.py.
* adan_step: optimizer.py:238-293 (_single_tensor_adan) with the state set-up
  of optimizer.py:147-186.  Pinned by closed forms (first step and constant
  gradients: every step moves by lr * g / (|g| + eps)), by its reduction to
  Adam when beta2 = 0 (then diff_t, t_t and the second update vanish, checked
  against torch.optim.Adam), and by the decoupled weight-decay forms of
  no_prox / prox on the first step.
"""
from __future__ import annotations

import numpy as np


def adam_step(p, g, state, lr=1e-3, betas=(0.9, 0.999), eps=1e-8, weight_decay=0.0):
    p = np.asarray(p, np.float64).copy()
    g = np.asarray(g, np.float64)
    b1, b2 = betas
    st = state
    if not st:
        st.update(step=0, m=np.zeros_like(p), v=np.zeros_like(p))
    st["step"] += 1
    t = st["step"]
    if weight_decay:
        g = g + weight_decay * p
    st["m"] = b1 * st["m"] + (1 - b1) * g
    st["v"] = b2 * st["v"] + (1 - b2) * g * g
    denom = np.sqrt(st["v"]) / np.sqrt(1 - b2 ** t) + eps
    return p - (lr / (1 - b1 ** t)) * st["m"] / denom


def adan_step(p, g, state, lr=1e-3, betas=(0.98, 0.92, 0.99), eps=1e-8, weight_decay=0.0,
              no_prox=False, clip=1.0):
    """One step of optimizer.py:238-293 for one tensor; `state` is a dict
    carrying step / m / v / diff / neg_pre between calls."""
    p = np.asarray(p, np.float64).copy()
    g = np.asarray(g, np.float64) * clip
    b1, b2, b3 = betas
    st = state
    if not st:
        st.update(step=0, m=np.zeros_like(p), v=np.zeros_like(p), diff=np.zeros_like(p))
    st["step"] += 1
    t = st["step"]
    if "neg_pre" not in st or t == 1:
        st["neg_pre"] = -g
    d = st["neg_pre"] + g
    st["m"] = b1 * st["m"] + (1 - b1) * g
    st["diff"] = b2 * st["diff"] + (1 - b2) * d
    u = b2 * d + g
    st["v"] = b3 * st["v"] + (1 - b3) * u * u
    denom = np.sqrt(st["v"]) / np.sqrt(1 - b3 ** t) + eps
    step_size = lr / (1 - b1 ** t)
    step_size_diff = lr * b2 / (1 - b2 ** t)
    if no_prox:
        p = p * (1 - lr * weight_decay)
        p = p - step_size * st["m"] / denom - step_size_diff * st["diff"] / denom
    else:
        p = p - step_size * st["m"] / denom - step_size_diff * st["diff"] / denom
        p = p / (1 + lr * weight_decay)
    st["neg_pre"] = -g
    return p
