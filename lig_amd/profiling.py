"""HIP-event timing of individual C-ABI entry points.

    with KernelTimer("lig_rasterize_backward") as kt:
        ...  # work that calls the entry point
    kt.mean_ms()

Events are recorded on the stream the call is issued on (the current torch
stream of the device, which is the stream passed to the C-ABI), immediately
before and after the call, so the span covers the kernels that call launches.
"""
from __future__ import annotations

import torch

from . import _lib


class KernelTimer:
    def __init__(self, name: str, device=None):
        self.name = name
        self.device = device
        self.pairs = []
        self._open = None

    def _hook(self, name, phase):
        ev = torch.cuda.Event(enable_timing=True)
        ev.record(torch.cuda.current_stream(self.device))
        if phase == "begin":
            self._open = ev
        else:
            self.pairs.append((self._open, ev))
            self._open = None

    def __enter__(self):
        _lib.hooks[self.name] = self._hook
        return self

    def __exit__(self, *exc):
        _lib.hooks.pop(self.name, None)
        return False

    def times_ms(self):
        torch.cuda.synchronize(self.device)
        return [a.elapsed_time(b) for a, b in self.pairs]

    def mean_ms(self) -> float:
        t = self.times_ms()
        return sum(t) / len(t) if t else float("nan")
