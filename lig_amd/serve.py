"""Crop-render serving with continuous batching and HBM admission.

Resident two-level representations (lig_amd.render.TwoLevel) stay in HBM.
Requests (representation, window origin, size, scale) arrive over time; a
fixed number of worker slots, each with its own HIP stream, take requests as
they arrive and run them concurrently on the device, so work of requests
admitted while others are in flight overlaps.  With max_batch > 1 a slot
also takes the requests already queued behind the one it takes (up to
max_batch) and renders them as one batch: one launch per stage for all of
them, the crops being images of one batch table (render.CropBatch).
Admission is by HBM budget: a request (or batch) first projects its windows
(small, and gives the exact intersection counts), then reserves the bytes
the rest of its render needs and waits while the reservations in flight
would exceed the budget.  A render uses only its own buffers and
deterministic kernels, so each result is the same whether it runs alone, in
a batch or in the mix.

Reported: per-request latency (arrival to completion) p50 / p99, output
Gpix/s, achieved request rate, peak reserved bytes.
"""
from __future__ import annotations

import queue
import threading
import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np
import torch

from .render import CropBatch, TwoLevel


@dataclass
class Request:
    rep: int
    x0: float
    y0: float
    w: int
    h: int
    scale: float = 1.0
    arrival: float = 0.0
    done: float = 0.0
    result: Optional[torch.Tensor] = None
    reserved: int = 0
    error: Optional[BaseException] = None


class HbmBudget:
    """Byte budget shared by the slots; `reserve` blocks until the request
    fits (a request larger than the whole budget is admitted alone)."""

    def __init__(self, budget: int):
        self.budget = int(budget)
        self.used = 0
        self.peak = 0
        self._cv = threading.Condition()

    def reserve(self, nbytes: int):
        with self._cv:
            while self.used > 0 and self.used + nbytes > self.budget:
                self._cv.wait()
            self.used += nbytes
            self.peak = max(self.peak, self.used)

    def release(self, nbytes: int):
        with self._cv:
            self.used -= nbytes
            self._cv.notify_all()


@dataclass
class ServeStats:
    latencies_ms: List[float] = field(default_factory=list)
    pixels: int = 0
    wall_s: float = 0.0
    requests: int = 0
    peak_reserved: int = 0
    batch_sizes: List[int] = field(default_factory=list)

    def summary(self) -> dict:
        lat = np.array(self.latencies_ms) if self.latencies_ms else np.zeros(1)
        extra = {}
        if self.batch_sizes:
            extra = {"mean_batch": round(float(np.mean(self.batch_sizes)), 2),
                     "max_batch_seen": int(max(self.batch_sizes))}
        return {**extra, "requests": self.requests,
                "p50_ms": round(float(np.percentile(lat, 50)), 3),
                "p99_ms": round(float(np.percentile(lat, 99)), 3),
                "mean_ms": round(float(lat.mean()), 3),
                "gpix_s": round(self.pixels / max(self.wall_s, 1e-9) / 1e9, 4),
                "achieved_qps": round(self.requests / max(self.wall_s, 1e-9), 2),
                "peak_reserved_gb": round(self.peak_reserved / 2**30, 3)}


class CropServer:
    """Continuous-batching crop renderer over resident representations.

    max_batch = 1: each slot renders one request at a time (its own
    projection, binning, raster and composition launches).  max_batch > 1:
    a slot that takes a request also takes up to max_batch - 1 more that are
    already queued and renders them together as one render.CropBatch (one
    launch per stage for all of them); every result has the same bits either
    way."""

    def __init__(self, reps: List[TwoLevel], slots: int = 4, hbm_budget: int = 16 << 30,
                 keep_results: bool = False, device="cuda:0", max_batch: int = 1):
        if max_batch < 1:
            raise ValueError("max_batch must be at least 1")
        self.reps = reps
        self.slots = slots
        self.max_batch = int(max_batch)
        self.budget = HbmBudget(hbm_budget)
        self.keep_results = keep_results
        self.device = torch.device(device)
        self._q: "queue.Queue[Optional[Request]]" = queue.Queue()
        self._done: List[Request] = []
        self._lock = threading.Lock()
        self._threads: List[threading.Thread] = []
        self.batch_sizes: List[int] = []

    # -- worker -----------------------------------------------------------
    def _take(self):
        """The next request and up to max_batch - 1 more already queued;
        (batch, stop) where stop says a stop marker was taken."""
        req = self._q.get()
        if req is None:
            return [], True
        batch = [req]
        while len(batch) < self.max_batch:
            try:
                nxt = self._q.get_nowait()
            except queue.Empty:
                break
            if nxt is None:
                return batch, True
            batch.append(nxt)
        return batch, False

    def _finish(self, reqs: List[Request], outs=None, error: Optional[BaseException] = None):
        now = time.perf_counter()
        for k, req in enumerate(reqs):
            req.done = now
            req.error = error
            req.result = outs[k] if (outs is not None and self.keep_results) else None
        with self._lock:
            self._done.extend(reqs)

    def _run_batch(self, batch: List[Request], stream):
        ok = []
        for req in batch:   # geometry errors are the request's own
            try:
                self.reps[req.rep].crop_geometry(req.x0, req.y0, req.w, req.h, req.scale)
                ok.append(req)
            except BaseException as e:
                self._finish([req], error=e)
        if not ok:
            return
        try:
            cb = CropBatch([(self.reps[r.rep], r.x0, r.y0, r.w, r.h, r.scale) for r in ok])
            need = cb.need
            for r in ok:
                r.reserved = need
            self.budget.reserve(need)
            try:
                outs = cb.finish()
                stream.synchronize()
            finally:
                del cb
                self.budget.release(need)
            with self._lock:
                self.batch_sizes.append(len(ok))
            self._finish(ok, outs)
        except BaseException as e:
            self._finish(ok, error=e)

    def _worker(self):
        torch.cuda.set_device(self.device)
        stream = torch.cuda.Stream(device=self.device)   # default priority
        with torch.cuda.stream(stream):
            while self.max_batch > 1:
                batch, stop = self._take()
                if batch:
                    self._run_batch(batch, stream)
                if stop:
                    return
            while True:
                req = self._q.get()
                if req is None:
                    break
                try:
                    rep = self.reps[req.rep]
                    plan = rep.crop_plan(req.x0, req.y0, req.w, req.h, req.scale)
                    req.reserved = plan[3]
                    self.budget.reserve(req.reserved)
                    try:
                        out = rep.finish_crop(plan, req.w, req.h)
                        stream.synchronize()
                    finally:
                        del plan
                        self.budget.release(req.reserved)
                    req.done = time.perf_counter()
                    req.result = out if self.keep_results else None
                except BaseException as e:   # reported per request
                    req.error = e
                    req.done = time.perf_counter()
                with self._lock:
                    self._done.append(req)

    def start(self):
        for _ in range(self.slots):
            t = threading.Thread(target=self._worker, daemon=True)
            t.start()
            self._threads.append(t)

    def submit(self, req: Request):
        req.arrival = time.perf_counter()
        self._q.put(req)

    def stop(self) -> List[Request]:
        for _ in self._threads:
            self._q.put(None)
        for t in self._threads:
            t.join()
        self._threads = []
        done, self._done = self._done, []
        return done

    # -- load generator ---------------------------------------------------
    def run_poisson(self, requests: List[Request], qps: float, seed: int = 3000) -> ServeStats:
        """Submit `requests` with exponential inter-arrival times of mean
        1/qps; returns latency / throughput statistics."""
        rng = np.random.default_rng(seed)
        gaps = rng.exponential(1.0 / qps, len(requests))
        self.start()
        t0 = time.perf_counter()
        nxt = t0
        for req, gap in zip(requests, gaps):
            nxt += gap
            delay = nxt - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            self.submit(req)
        done = self.stop()
        st = ServeStats()
        st.wall_s = time.perf_counter() - t0
        errs = [r.error for r in done if r.error is not None]
        if errs:
            raise errs[0]
        st.latencies_ms = [(r.done - r.arrival) * 1e3 for r in done]
        st.pixels = sum(r.w * r.h for r in done)
        st.requests = len(done)
        st.peak_reserved = self.budget.peak
        st.batch_sizes = list(self.batch_sizes)
        return st


# -- synthetic workload (C5) ----------------------------------------------

def make_representations(count: int, seed: int = 2000, device="cuda:0",
                         sizes=(4096, 6144, 8192), n_range=(500_000, 3_000_000),
                         scale_px: float = 7.2) -> List[TwoLevel]:
    """`count` random two-level representations: canvas side from `sizes`,
    level-2 Gaussian count uniform in n_range, level 1 with half as many on
    the 4x smaller canvas; parameters drawn as the fit's initialisation
    (seeded per representation)."""
    from .data import init_gaussians
    from .render import LevelParams
    rng = np.random.default_rng(seed)
    reps = []
    for i in range(count):
        S = int(rng.choice(sizes))
        n2 = int(rng.integers(n_range[0], n_range[1] + 1))
        n1 = n2 // 2
        m1, c1, r1 = init_gaussians(n1, S // 4, S // 4, seed + 17 * i + 1, scale_px, device)
        m2, c2, r2 = init_gaussians(n2, S, S, seed + 17 * i + 2, scale_px, device)
        reps.append(TwoLevel(LevelParams(m1, c1, r1 * 0.05, S // 4, S // 4),
                             LevelParams(m2, c2, r2 * 0.05, S, S), lo=-0.5, hi=1.5))
    return reps


def make_requests(reps: List[TwoLevel], count: int, seed: int = 3000,
                  crop_range=(256, 4096)) -> List[Request]:
    """Crops of log-uniform side in crop_range (multiples of 16, square or
    not), random offsets inside the canvas, scale 1 or 0.5."""
    rng = np.random.default_rng(seed)
    out = []
    lo, hi = np.log(crop_range[0]), np.log(crop_range[1])
    for _ in range(count):
        i = int(rng.integers(len(reps)))
        S = reps[i].l2.W
        scale = float(rng.choice([1.0, 0.5]))
        w = int(np.exp(rng.uniform(lo, hi))) // 16 * 16
        h = int(np.exp(rng.uniform(lo, hi))) // 16 * 16
        w = max(16, min(w, int(S * scale)))
        h = max(16, min(h, int(S * scale)))
        span_w, span_h = w / scale, h / scale
        x0 = float(rng.integers(0, int(S - span_w) + 1) // 4 * 4)
        y0 = float(rng.integers(0, int(S - span_h) + 1) // 4 * 4)
        out.append(Request(i, x0, y0, w, h, scale))
    return out
