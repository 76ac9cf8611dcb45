"""Autotuning of the fit step's launch configuration.

The tuned knob is the forward's LDS batch depth: how many Gaussians a tile's
wave stages in LDS per batch (64, 128 or 256; results do not depend on it).
Each candidate is timed on the caller's own workload with HIP events around
the fused render call (lig_amd.profiling.KernelTimer), a few repetitions
after one warm-up, and the fastest is kept.  Choices are cached per workload
signature in memory and, when writable, in a JSON file
($LIG_AUTOTUNE_CACHE, default ~/.cache/lig_amd/autotune.json).

For batched fitting the launch form is tuned too: one workgroup per tile
(the hardware deals workgroups to the XCDs round-robin) or a persistent grid
fed by per-XCD tile queues with stealing (lig_fit_render_l2_batched_persistent),
which wins when some images' tiles cost much more than others'.

Tile size: the fused fit kernels are written for 16 x 16 tiles, and the
reference API caps block_width at 16 (rasterize.py:26), so 32 is not a
candidate.  `tune_level_tiles` also times 8 x 8 tiles through the gsplat2d API
path (Level(fused=False)) against the fused 16 x 16 step at every depth, on
the whole forward + backward.  Images do not depend on the tile size (a
pixel's sum runs over the same Gaussians in gid order); gradients are summed
over a different number of per-tile partials, so their last bits do.  The
batched kernels are 16 x 16 only.
"""
from __future__ import annotations

import json
import os
from typing import Callable, Dict, Optional, Sequence

import torch

from .profiling import KernelTimer

DEPTHS = (64, 128, 256)
_cache: Dict[str, int] = {}


def _cache_path() -> str:
    return os.environ.get("LIG_AUTOTUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "lig_amd",
                                       "autotune.json"))


def _load():
    try:
        with open(_cache_path()) as f:
            _cache.update({k: v for k, v in json.load(f).items()})
    except (OSError, ValueError):
        pass


def _store():
    try:
        path = _cache_path()
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as f:
            json.dump(_cache, f, indent=1, sort_keys=True)
    except OSError:
        pass


def tune(key: str, run: Callable[[int], None], entry: str, device,
         candidates: Sequence[int] = DEPTHS, reps: int = 3, use_cache: bool = True) -> dict:
    """Time `run(depth)` for each depth (the C-ABI call `entry` timed by HIP
    events); returns {"choice": depth, "ms": {depth: mean ms}}."""
    if use_cache:
        if not _cache:
            _load()
        if key in _cache:
            return {"choice": _cache[key], "ms": None, "cached": True}
    ms = {}
    for d in candidates:
        run(d)   # warm-up
        t = KernelTimer(entry, device)
        with t:
            for _ in range(reps):
                run(d)
        torch.cuda.synchronize(device)
        ms[d] = t.mean_ms()
    choice = min(ms, key=ms.get)
    _cache[key] = choice
    _store()
    return {"choice": choice, "ms": {k: round(v, 4) for k, v in ms.items()}, "cached": False}


def tune_level(level, target: torch.Tensor, reps: int = 3, use_cache: bool = True) -> dict:
    """Pick lig_amd.fit.Level.batch_depth for `level` fitting `target`."""
    key = f"level:{level.H}x{level.W}:n{level.means.shape[0]}"

    def run(d):
        with torch.no_grad():
            level.batch_depth = d
            level.loss(target)

    res = tune(key, run, "lig_fit_render_l2", level.means.device, reps=reps,
               use_cache=use_cache)
    level.batch_depth = res["choice"]
    return res


def tune_batched(bf, reps: int = 3, use_cache: bool = True) -> dict:
    """Pick lig_amd.batch.BatchedFit.batch_depth and .persistent: each depth
    is timed in both launch forms (HIP events around the render call) and the
    fastest pair is kept.  Returns {"choice": depth, "persistent": bool, "ms":
    {form: {depth: ms}}}."""
    key = f"batch:{len(bf.sizes)}:{bf.num_tiles}:n{bf.means.shape[0]}:{sum(bf.counts)}"
    if use_cache:
        if not _cache:
            _load()
        hit = _cache.get(key)
        if isinstance(hit, list) and len(hit) == 2:
            bf.batch_depth, bf.persistent = int(hit[0]), bool(hit[1])
            return {"choice": bf.batch_depth, "persistent": bf.persistent, "ms": None,
                    "cached": True}
    ms = {}
    for persistent in (False, True):
        bf.persistent = persistent

        def run(d):
            with torch.no_grad():
                bf.batch_depth = d
                bf.loss()

        entry = "lig_fit_render_l2_batched" + ("_persistent" if persistent else "")
        r = tune(key + f":{int(persistent)}", run, entry, bf.means.device, reps=reps,
                 use_cache=False)
        ms["persistent" if persistent else "grid"] = r["ms"]
    best = min(((f, d, t) for f, row in ms.items() for d, t in row.items()),
               key=lambda x: x[2])
    bf.persistent = best[0] == "persistent"
    bf.batch_depth = best[1]
    _cache[key] = [bf.batch_depth, bf.persistent]
    _store()
    return {"choice": bf.batch_depth, "persistent": bf.persistent, "ms": ms, "cached": False}


def tune_level_tiles(level, target: torch.Tensor, reps: int = 2,
                     use_cache: bool = True) -> dict:
    """Pick the tile size and LDS depth of `level`'s step: the fused 16 x 16
    step at each depth and 8 x 8 tiles through the API path, each timed with
    HIP events over loss + backward (gradients cleared after, no optimiser
    step, so the parameters do not change).  Sets level.fused, .block_width
    and .batch_depth; returns {"choice": {...}, "ms": {candidate: ms}}."""
    key = f"tiles:{level.H}x{level.W}:n{level.means.shape[0]}"
    cands = {f"16/{d}": (True, 16, d) for d in DEPTHS}
    cands["8/api"] = (False, 8, 0)

    def apply(c):
        level.fused, level.block_width, level.batch_depth = c

    if use_cache:
        if not _cache:
            _load()
        hit = _cache.get(key)
        if isinstance(hit, str) and hit in cands:
            apply(cands[hit])
            return {"choice": hit, "ms": None, "cached": True}
    dev = level.means.device
    ms = {}
    for name, c in cands.items():
        apply(c)

        def run():
            loss = level.loss(target)
            loss.backward()
            for p in (level.means, level.cholesky, level.rgbs):
                p.grad = None

        run()   # warm-up
        torch.cuda.synchronize(dev)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            run()
        e1.record()
        e1.synchronize()
        ms[name] = round(e0.elapsed_time(e1) / reps, 4)
    choice = min(ms, key=ms.get)
    apply(cands[choice])
    _cache[key] = choice
    _store()
    return {"choice": choice, "ms": ms, "cached": False}
