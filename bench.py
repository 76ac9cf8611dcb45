#!/usr/bin/env python
"""Fit throughput of the two-level LIG workload on one MI355X.

Default workload (the configuration the headline metric is quoted on):
8192 x 8192 RGB target (seeded blobs + 1/f noise + edges + tiled cell
texture); level 1 fits the 4x bicubic-downsampled 2048^2 image with 1M
Gaussians, level 2 fits the normalised residual of the upscaled level-1
render at 8192^2 with 2M Gaussians.  Cholesky diagonal ~U(0.5, 3) * 7.2 px
(about 40 tiles per Gaussian), 16x16 tiles, sum rasterisation, clamp + L2
loss, Adam lr 1e-3.  One step = one optimiser step of level 1 plus one of
level 2; value = steps per second.

Also reported: render Gpix/s of the level-2 representation (image-only
renderer: project + bin + rasterise; the reference API's forward beside it),
the dominant kernel's roofline (HIP events around its C-ABI call on the
stream it runs on; algorithmic FLOPs from exact pair counts), and the CPU
oracle timed on a bounded crop of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0

# The two-level configurations start their colours at zero, as gaussianlig.py:58
# initialises rgbs: with ~40 tiles per Gaussian, colours ~U(0, 1) would render
# every pixel far above 1 (mean 14.8 at level 2), so the clamped loss would pass
# no gradient at all, and the GPU runs such all-zero gradients 8-12% faster than
# a fit that moves.
CONFIGS = {
    "c3": dict(H=8192, W=8192, seed=2, cells=True, n1=1_000_000, n2=2_000_000, scale=7.2,
               factor=4,
               workload="two-level LIG fit, 8192x8192 RGB: L1 2048^2 x 1M + L2 8192^2 x 2M "
                        "Gaussians, chol diag U(0.5,3)*7.2 px, colours from 0, 16x16 tiles, "
                        "L2 loss, Adam"),
    "c2": dict(H=2048, W=2048, seed=1, cells=False, n1=100_000, n2=300_000, scale=1.0, factor=4,
               workload="two-level LIG fit, 2048x2048 RGB: L1 512^2 x 100k + L2 2048^2 x 300k "
                        "Gaussians, chol diag U(0.5,3) px, colours from 0, 16x16 tiles, L2 loss, "
                        "Adam"),
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def raster_flops(E, C, backward):
    """Algorithmic FLOPs of the fit step's sum rasteriser (FMA = 2; exp, min,
    compare = 1): 17 per pair inside the SnugBox box, plus 7 (forward) or 29
    (backward) per contributing pair.  29 is the reference backward's 37
    without the 8 of v_xy_abs (two products and two |.| accumulations), which
    the fit step does not form: no parameter gradient depends on it."""
    return 17 * E + (29 if backward else 7) * C


def raster_bytes(I, T, P, N, backward):
    return 36 * I + 8 * T + 16 * P + (40 * N if backward else 0)


def level_stages(N, I, T, P, E, C):
    """Algorithmic (FLOP, byte) tallies of one fit step of one level, stage by
    stage: Cholesky projection forward / backward, tile intersection (scan,
    emit, one-pass-equivalent sort, bins), raster forward, clamp + L2 loss,
    raster backward, Adam."""
    return {
        "project_fwd": (0, 52 * N),
        "intersect": (0, 28 * N + 28 * I + 8 * T),
        "raster_fwd": (raster_flops(E, C, False), raster_bytes(I, T, P, N, False)),
        "loss": (0, 36 * P),
        "raster_bwd": (raster_flops(E, C, True), raster_bytes(I, T, P, N, True)),
        "project_bwd": (0, 72 * N),
        "adam": (0, 224 * N),
    }


def stage_roof_ms(flops, nbytes):
    return max(flops / (PEAK_FP32_TFLOPS * 1e12), nbytes / (PEAK_HBM_GBS * 1e9)) * 1e3


def level_counts(level, C, torch):
    """Exact (N, I, T, P, E, C) of a level's current parameters."""
    with torch.no_grad():
        xys, ext, con, nt = level.project()
        E, Cc = C.count_pairs(xys, ext, con, None, nt, level.H, level.W)
        I = int(nt.sum().item())
        mean_tiles = float(nt.float().mean().item())
    T = ((level.H + 15) // 16) * ((level.W + 15) // 16)
    return dict(N=int(level.means.shape[0]), I=I, T=T, P=level.H * level.W, E=E, C=Cc,
                mean_tiles_per_gaussian=round(mean_tiles, 2))


TRAFFIC_FILES = {"c3": os.path.join(ROOT, "profiles", "raster_c3_8192_traffic.json")}


def measured_traffic(config, kernel):
    """HBM bytes per launch of `kernel` on the level-2 canvas, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE record of this workload
    (each counter scaled by its reading for a copy of known size), or None."""
    path = TRAFFIC_FILES.get(config)
    if not path or not os.path.exists(path):
        return None, None
    with open(path) as f:
        rec = json.load(f)
    k = rec.get("kernels", {}).get(kernel)
    if not k:
        return None, None
    return int(k["traffic_bytes"]), os.path.relpath(path, ROOT)


def cpu_baseline(level, target, oracle, torch):
    """Oracle (FP64 CPU port) fit step of a whole level: Cholesky projection,
    binning, rasteriser forward, clamp + L2 loss gradient, rasteriser
    backward, projection backward, Adam (torch's update on the CPU).
    Returns (seconds, Gaussians, intersections)."""
    import numpy as np
    H, W = level.H, level.W
    with torch.no_grad():
        means = level.means.detach().cpu().numpy()
        chol = level.cholesky.detach().cpu().numpy()
        rgbs = level.rgbs.detach().cpu().numpy()
        tgt = target.cpu().numpy().astype(np.float64)
    params = [torch.from_numpy(a.copy()).requires_grad_(True) for a in (means, chol, rgbs)]
    opt = torch.optim.Adam(params, lr=1e-3)
    t0 = time.perf_counter()
    pxy, pext, pcon, pnt = oracle.project_gaussians_forward(chol, means, H, W, 16, None,
                                                            cholesky=True)
    m, cum, isect, gids, ks, gs, bins = oracle.bin_and_sort(pxy, pext, pnt, H, W, 16)
    del isect, gids, ks
    fwd = oracle.rasterize_forward(gs, bins, pxy, pcon, rgbs, None, H, W, 16)
    img = fwd["img"]
    inside = (img > 0) & (img < 1)
    v_out = ((2.0 / img.size) * (np.clip(img, 0, 1) - tgt) * inside).astype(np.float32)
    del img, inside
    bwd = oracle.rasterize_backward(gs, bins, pxy, pcon, rgbs, None, fwd["final_idx"], v_out,
                                    np.zeros((H, W), np.float32), H, W, 16)
    v_chol, v_mean = oracle.project_gaussians_backward(pext, pcon, bwd["v_xy"], bwd["v_conic"],
                                                       cholesky=chol)
    for p_, g_ in zip(params, (v_mean, v_chol, bwd["v_rgb"])):
        p_.grad = torch.from_numpy(np.ascontiguousarray(g_, dtype=np.float32))
    opt.step()
    return time.perf_counter() - t0, int(means.shape[0]), m


def batched_run(args, dev, torch):
    """Batched fitting (C4): 64 seeded images of uneven size at 0.05 Gaussians
    per pixel; the plain form steps each image's Level in turn (one image per
    launch sequence), the batched form steps all images with one launch per
    stage.  Both start from the same draws, so after the same number of steps
    their per-image PSNRs are equal (the batched form is bit-identical per
    image)."""
    import numpy as np
    from lig_amd import batch, fit
    sizes, counts, seeds = batch.batch_workload(64, seed=1000, gauss_per_px=0.05)
    targets = batch.make_targets(sizes, seeds, dev)
    scale = 2.0
    plain = [fit.Level(c, h, w, s, scale, device=dev) for (h, w), c, s in
             zip(sizes, counts, seeds)]
    bf = batch.BatchedFit(targets, counts, seeds, scale=scale)
    from lig_amd import autotune
    tuned = autotune.tune_batched(bf, use_cache=False)
    for lv, t in zip(plain, targets):
        autotune.tune_level(lv, t, reps=1, use_cache=False)
    res = {"batch_depth": tuned["choice"], "persistent": tuned["persistent"],
           "tune_ms": tuned["ms"],
           "plain_batch_depths": sorted({lv.batch_depth for lv in plain}),
           "images": len(sizes), "pixels": int(sum(h * w for h, w in sizes)),
           "gaussians": int(sum(counts)), "steps": args.batch_steps,
           "workload": "64 seeded images, height log-uniform 256-2048, width log-uniform "
                       "384-1536, 0.05 Gaussians/px, chol diag U(0.5,3)*2 px, L2, Adam"}

    def timed(stepper, steps):
        for _ in range(2):
            stepper()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            stepper()
        torch.cuda.synchronize()
        return round(steps / (time.perf_counter() - t0), 3)

    res["plain_steps_s"] = timed(lambda: [lv.step(t) for lv, t in zip(plain, targets)],
                                 args.batch_steps)
    res["batched_steps_s"] = timed(bf.step, args.batch_steps)   # the tuned launch form
    res["speedup"] = round(res["batched_steps_s"] / res["plain_steps_s"], 3)
    with torch.no_grad():
        p_plain = np.array([fit.psnr(torch.clamp(lv.render(), 0, 1), t)
                            for lv, t in zip(plain, targets)])
        bf.loss()
        p_batch = bf.psnr()
    res["max_abs_dpsnr_db"] = float(np.abs(p_plain - p_batch).max())
    res["mean_psnr_db"] = round(float(p_batch.mean()), 3)
    # the plain form with each image's step captured (Level.graph): k steps
    # of every image per call, one image after another
    graphs = [lv.graph(t) for lv, t in zip(plain, targets)]
    for gph in graphs:
        gph.run(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for gph in graphs:
        gph.run(args.batch_steps)
    torch.cuda.synchronize()
    res["plain_graph_steps_s"] = round(args.batch_steps / (time.perf_counter() - t0), 3)
    res["speedup_vs_plain_graph"] = round(res["batched_steps_s"] / res["plain_graph_steps_s"], 3)
    del graphs
    # both launch forms of the raster stages (bit-identical): one workgroup per
    # tile, and the persistent grid fed by per-XCD tile queues
    chosen = bf.persistent
    for form, flag in (("grid", False), ("persistent", True)):
        bf.persistent = flag
        res[f"{form}_steps_s"] = timed(bf.step, args.batch_steps)
    bf.persistent = chosen
    # the same step recorded once as a hipGraph (no host read of the
    # intersection count; bit-identical to the eager step)
    gs = batch.BatchedStepGraph(bf)
    res["graph_steps_s"] = timed(lambda: gs.run(1), args.batch_steps)
    res["graph_speedup"] = round(res["graph_steps_s"] / res["plain_steps_s"], 3)
    res["graph_capacity"] = gs.capacity
    res["graph_captures"] = gs.captures
    del gs, bf
    torch.cuda.empty_cache()
    # the two forms at a fitted PSNR: fresh parameters, K steps of each
    # image's captured plain step (one image after another) against K replays
    # of the captured batched step, lr 1e-2; mean PSNR of each form after
    if args.batch_fit_steps > 0:
        K = args.batch_fit_steps
        lr = 1e-2
        plain_f = [fit.Level(c, h, w, s, scale, lr=lr, device=dev) for (h, w), c, s in
                   zip(sizes, counts, seeds)]
        for lv, lp in zip(plain_f, plain):
            lv.batch_depth = lp.batch_depth
        bff = batch.BatchedFit(targets, counts, seeds, scale=scale, lr=lr)
        bff.batch_depth, bff.persistent = tuned["choice"], tuned["persistent"]
        graphs = [lv.graph(t) for lv, t in zip(plain_f, targets)]
        gsf = batch.BatchedStepGraph(bff)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for gph in graphs:
            gph.run(K)
        torch.cuda.synchronize()
        plain_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        gsf.run(K)
        torch.cuda.synchronize()
        batch_s = time.perf_counter() - t0
        with torch.no_grad():
            pp = np.array([fit.psnr(torch.clamp(lv.render(), 0, 1), t)
                           for lv, t in zip(plain_f, targets)])
            bff.loss()
            pb = bff.psnr()
        res["fitted"] = {"steps": K, "lr": lr,
                         "plain_graph_steps_s": round(K / plain_s, 3),
                         "batched_graph_steps_s": round(K / batch_s, 3),
                         "speedup": round(plain_s / batch_s, 3),
                         "mean_psnr_db_plain": round(float(pp.mean()), 4),
                         "mean_psnr_db_batched": round(float(pb.mean()), 4),
                         "max_abs_dpsnr_db": float(np.abs(pp - pb).max()),
                         "plain_captures": int(sum(gph.captures for gph in graphs)),
                         "batched_captures": gsf.captures}
        log(f"batched fitted {res['fitted']}")
        del graphs, gsf, bff, plain_f
        torch.cuda.empty_cache()
    # skewed batch: the same images with 8 of them at 20x the Gaussian density,
    # where whole XCD shares of the tile list cost far more than others
    skew = list(counts)
    for i in range(0, 16, 2):
        skew[i] *= 20
    bs = batch.BatchedFit(targets, skew, seeds, scale=scale)
    tuned_s = autotune.tune_batched(bs, use_cache=False)
    sk = {"gaussians": int(sum(skew)), "tuned_persistent": tuned_s["persistent"],
          "tuned_depth": tuned_s["choice"]}
    for form, flag in (("grid", False), ("persistent", True)):
        bs.persistent = flag
        sk[f"{form}_steps_s"] = timed(bs.step, 3)
    sk["persistent_speedup"] = round(sk["persistent_steps_s"] / sk["grid_steps_s"], 3)
    res["skewed"] = sk
    del bs
    log(f"batched {res}")
    del plain, targets
    torch.cuda.empty_cache()
    return res


def serve_run(args, dev, torch):
    """Crop serving (C5): resident two-level representations, Poisson crop
    requests, continuous batching over 4 stream slots, HBM admission; the
    load curve per request and with a slot's queued requests batched."""
    from lig_amd import serve
    torch.cuda.empty_cache()
    t0 = time.perf_counter()
    reps = serve.make_representations(args.serve_reps, seed=2000, device=dev)
    reqs = serve.make_requests(reps, args.serve_requests, seed=3000)
    # one request per representation first (untimed): first-use allocations
    warm = serve.CropServer(reps, slots=4, device=dev)
    warm.run_poisson([serve.Request(i, 0.0, 0.0, 256, 256, 1.0) for i in range(len(reps))],
                     qps=1e4)
    serve.CropServer(reps, slots=args.serve_batch_slots, device=dev,
                     max_batch=args.serve_batch).run_poisson(
        [serve.Request(r.rep, r.x0, r.y0, r.w, r.h, r.scale) for r in reqs[:64]], qps=1e4)
    torch.cuda.synchronize()
    log(f"serve setup {time.perf_counter() - t0:.1f}s ({len(reps)} representations)")
    srv = serve.CropServer(reps, slots=4, hbm_budget=32 << 30, device=dev)
    st = srv.run_poisson(reqs, qps=args.serve_qps, seed=3000)
    res = st.summary()
    res.update({"offered_qps": args.serve_qps, "slots": 4, "hbm_budget_gb": 32,
                "representations": len(reps),
                "workload": "32 two-level reps, 4096^2-8192^2 canvases, 0.5M-3M level-2 "
                            "Gaussians (+ half at level 1); crops 256^2-4096^2 log-uniform, "
                            "scale 1 or 0.5, Poisson arrivals"})
    # load curve: the same request stream offered faster; and one slot (a
    # plain FIFO queue, no overlap of requests) at the headline rate
    curve = []
    for q in [float(x) for x in args.serve_load_qps.split(",") if x]:
        st = serve.CropServer(reps, slots=4, hbm_budget=32 << 30,
                              device=dev).run_poisson(reqs, qps=q, seed=3000)
        curve.append(dict(offered_qps=q, **{k: v for k, v in st.summary().items()
                                           if k in ("p50_ms", "p99_ms", "achieved_qps",
                                                    "gpix_s")}))
    res["load_curve"] = curve
    # the same streams with queued crops batched per slot (render.CropBatch:
    # one launch per stage for a slot's queued requests)
    bcurve = []
    for q in [float(x) for x in args.serve_load_qps.split(",") if x]:
        st = serve.CropServer(reps, slots=args.serve_batch_slots, hbm_budget=32 << 30,
                              device=dev, max_batch=args.serve_batch).run_poisson(
                                  reqs, qps=q, seed=3000)
        bcurve.append(dict(offered_qps=q, **{k: v for k, v in st.summary().items()
                                            if k in ("p50_ms", "p99_ms", "achieved_qps",
                                                     "gpix_s", "mean_batch")}))
    res["batched"] = {"slots": args.serve_batch_slots, "max_batch": args.serve_batch,
                      "load_curve": bcurve}
    single = []
    for q in sorted({args.serve_qps, max([args.serve_qps] + [c["offered_qps"] for c in curve
                                                             if c["offered_qps"] <= 1200])}):
        st = serve.CropServer(reps, slots=1, hbm_budget=32 << 30,
                              device=dev).run_poisson(reqs, qps=q, seed=3000)
        single.append(dict(offered_qps=q, **{
            k: v for k, v in st.summary().items() if k in ("p50_ms", "p99_ms", "achieved_qps",
                                                             "gpix_s")}))
    res["single_slot"] = single
    log(f"serve {res}")
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c3", choices=sorted(CONFIGS))
    ap.add_argument("--render-reps", type=int, default=5)
    ap.add_argument("--cpu-baseline", type=int, default=1, help="0 to skip the CPU oracle leg")
    ap.add_argument("--serve-requests", type=int, default=600,
                    help="crop requests in the serving run (0 skips it)")
    ap.add_argument("--serve-qps", type=float, default=150.0)
    ap.add_argument("--serve-reps", type=int, default=32)
    ap.add_argument("--serve-load-qps", default="300,600,1200,2000,2400",
                    help="further offered rates of the serving run (comma list; empty skips)")
    ap.add_argument("--serve-batch", type=int, default=4,
                    help="most queued requests a slot renders as one batch in the batched curve")
    ap.add_argument("--serve-batch-slots", type=int, default=4)
    ap.add_argument("--skew-steps", type=int, default=3,
                    help="timed loss+backward calls on the density-skewed level 2 (0 skips)")
    ap.add_argument("--batch-steps", type=int, default=5,
                    help="timed steps of the batched-fitting comparison (0 skips it)")
    ap.add_argument("--batch-fit-steps", type=int, default=300,
                    help="steps of the batched-vs-plain comparison at a fitted PSNR (0 skips)")
    ap.add_argument("--api-steps", type=int, default=5,
                    help="timed steps of the reference-API chain at the headline workload")
    ap.add_argument("--captured-steps", type=int, default=-1,
                    help="steps of each level's fused step replayed as a hipGraph (0 skips; "
                         "default: 50 for the launch-bound c1/c2 sizes, 0 for c3)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1"}))
        sys.exit(2)

    import torch

    import lig_amd._lib as L
    from lig_amd import data, fit
    from lig_amd.gsplat2d import cuda as C
    from lig_amd.profiling import KernelTimer

    if not os.path.exists(L.LIB_PATH):
        import __graft_entry__
        __graft_entry__.build()
    L.load()
    cfg = CONFIGS[args.config]
    dev = torch.device("cuda:0")
    torch.manual_seed(cfg["seed"])

    t0 = time.perf_counter()
    gt = data.make_image(cfg["H"], cfg["W"], cfg["seed"], cells=cfg["cells"], device=dev)
    f = cfg["factor"]
    l1 = fit.Level(cfg["n1"], cfg["H"] // f, cfg["W"] // f, cfg["seed"] * 10 + 1, cfg["scale"],
                   device=dev)
    l2 = fit.Level(cfg["n2"], cfg["H"], cfg["W"], cfg["seed"] * 10 + 2, cfg["scale"], device=dev)
    with torch.no_grad():
        l1.rgbs.zero_()
        l2.rgbs.zero_()
    t1 = fit.downsample(gt, f)
    with torch.no_grad():
        up = torch.clamp(fit.upscale_render(l1, cfg["H"], cfg["W"]), 0, 1)
        t2, lo, hi = fit.residual_target(gt, up)
    del up
    torch.cuda.synchronize()
    log(f"setup {time.perf_counter() - t0:.1f}s")
    from lig_amd import autotune
    # tile size (16 fused / 8 through the API path) and LDS depth per level,
    # timed on forward + backward
    tuned = {}
    for name, lv, tt in (("l1", l1, t1), ("l2", l2, t2)):
        res = autotune.tune_level_tiles(lv, tt, use_cache=False)
        if not lv.fused:
            # the kernel roofline below times the fused 16 x 16 entry points
            best16 = min((k for k in res["ms"] if k.startswith("16/")), key=res["ms"].get)
            lv.fused, lv.block_width, lv.batch_depth = True, 16, int(best16.split("/")[1])
            res["measured_with"] = best16
        tuned[name] = res
    log(f"autotune {tuned}")

    for i in range(args.warmup):
        l1.step(t1)
        l2.step(t2)
        torch.cuda.synchronize()
        log(f"warmup {i + 1}/{args.warmup}")

    # per-level events are recorded every step and read after the timed
    # region, so the host never waits inside it
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    fwd_t = KernelTimer("lig_fit_render_l2", dev)
    bwd_t = KernelTimer("lig_fit_backward_records_l2", dev)
    torch.cuda.synchronize()
    start = time.perf_counter()
    for i in range(args.steps):
        ev = evs[i]
        ev[0].record()
        l1.step(t1)
        ev[1].record()
        with fwd_t, bwd_t:
            l2.step(t2)
        ev[2].record()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - start
    lvl_ms = {"l1": sum(ev[0].elapsed_time(ev[1]) for ev in evs),
              "l2": sum(ev[1].elapsed_time(ev[2]) for ev in evs)}
    steps_per_s = args.steps / elapsed
    log(f"timed {args.steps} steps in {elapsed:.3f}s")

    # launch-bound sizes: each level's fused step recorded once as a hipGraph
    # (Level.graph) and replayed; the levels are independent, so k two-level
    # steps take k replays of each level's graph
    captured = None
    ncap = args.captured_steps if args.captured_steps >= 0 else (
        0 if args.config == "c3" else 50)
    if ncap > 0:
        cap_s, caps = 0.0, 0
        for lv, tt in ((l1, t1), (l2, t2)):
            gph = lv.graph(tt)
            gph.run(2)
            torch.cuda.synchronize()
            c0 = time.perf_counter()
            gph.run(ncap)
            torch.cuda.synchronize()
            cap_s += time.perf_counter() - c0
            caps += gph.captures
            del gph
        captured = {"steps_s": round(ncap / cap_s, 3), "ms_per_step": round(cap_s / ncap * 1e3, 4),
                    "steps": ncap, "captures": caps,
                    "path": "each level's fused step replayed as a hipGraph (Level.graph)"}
        log(f"captured {captured}")

    # render throughput of the level-2 representation: the image-only renderer
    # (lig_amd.render.render_view: view projection, binning, image-only raster)
    # and, beside it, the reference API's forward (rasterize_gaussians, which
    # also writes the weight sum and final index a backward would need)
    from lig_amd import render as R

    def render_time(fn):
        with torch.no_grad():
            for _ in range(2):
                fn()
            torch.cuda.synchronize()
            r0 = time.perf_counter()
            for _ in range(args.render_reps):
                fn()
            torch.cuda.synchronize()
        return (time.perf_counter() - r0) / args.render_reps

    lp2 = R.LevelParams.of(l2)
    render_s = render_time(lambda: R.render_view(lp2, 0.0, 0.0, l2.W, l2.H))
    render_api_s = render_time(l2.render)
    counts = {"l1": level_counts(l1, C, torch), "l2": level_counts(l2, C, torch)}
    k2 = counts["l2"]
    I, T, P, E, Cc = k2["I"], k2["T"], k2["P"], k2["E"], k2["C"]
    fwd_ms, bwd_ms = fwd_t.mean_ms(), bwd_t.mean_ms()
    kernels = {
        "raster_fwd": dict(ms=fwd_ms, flops=raster_flops(E, Cc, False),
                           bytes=raster_bytes(I, T, P, l2.means.shape[0], False)),
        "raster_bwd": dict(ms=bwd_ms, flops=raster_flops(E, Cc, True),
                           bytes=raster_bytes(I, T, P, l2.means.shape[0], True)),
    }
    dom_name = max(kernels, key=lambda k: kernels[k]["ms"])
    dom = kernels[dom_name]
    achieved = dom["flops"] / (dom["ms"] * 1e-3) / 1e12
    traffic, traffic_src = measured_traffic(args.config, dom_name)

    # step roofline: sum over both levels' stages of max(FLOP time, byte time)
    t_roof = 0.0
    stage_ms = {}
    for lv, k in counts.items():
        for st, (fl, by) in level_stages(k["N"], k["I"], k["T"], k["P"], k["E"],
                                         k["C"]).items():
            ms = stage_roof_ms(fl, by)
            stage_ms[f"{lv}.{st}"] = round(ms, 4)
            t_roof += ms
    step_ms = 1e3 / steps_per_s
    step_roofline = {"t_roof_ms": round(t_roof, 3), "t_step_ms": round(step_ms, 3),
                     "frac": round(t_roof / step_ms, 4), "stage_roof_ms": stage_ms}
    # "mfma" is the compute-bound class of the roofline (against MI355X's FP32
    # figure); the kernel itself runs on the FP32 vector ALUs, no MFMA
    roofline = {"bound": "mfma", "compute_unit": "FP32 VALU (no MFMA instructions)",
                "kernel": dom_name, "achieved": round(achieved, 3),
                "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": traffic,
                "traffic_source": traffic_src,
                "kernel_ms": round(dom["ms"], 3), "algorithmic_flops": dom["flops"],
                "algorithmic_bytes": dom["bytes"],
                "hbm_gbs_algorithmic": round(dom["bytes"] / (dom["ms"] * 1e-3) / 1e9, 1)}

    # the rate an unchanged train.py gets through the reference API (the
    # alias of lig_amd.gsplat2d): Level(fused=False) runs project ->
    # rasterize_gaussians -> clamp + L2 -> backward -> Adam per level
    api_chain = None
    if args.api_steps > 0:
        for lv in (l1, l2):
            lv.fused = False
        for _ in range(2):
            l1.step(t1)
            l2.step(t2)
        torch.cuda.synchronize()
        a0 = time.perf_counter()
        for _ in range(args.api_steps):
            l1.step(t1)
            l2.step(t2)
        torch.cuda.synchronize()
        api_s = time.perf_counter() - a0
        for lv in (l1, l2):
            lv.fused = True
        api_chain = {"steps_s": round(args.api_steps / api_s, 3),
                     "ms_per_step": round(api_s / args.api_steps * 1e3, 3),
                     "steps": args.api_steps,
                     "path": "gsplat2d API: project_gaussians_cholesky -> rasterize_gaussians "
                             "-> clamp + L2 -> backward -> Adam (Level(fused=False))"}
        log(f"api chain {api_chain}")

    cpu = None
    if args.cpu_baseline:
        import oracle
        oracle.build()
        threads = oracle.num_threads()
        total = 0.0
        parts = []
        for name, lev, tgt in (("l1", l1, t1), ("l2", l2, t2)):
            secs, ng, m = cpu_baseline(lev, tgt, oracle, torch)
            total += secs
            parts.append(f"{name}: whole {lev.W}x{lev.H} level ({ng} Gaussians, {m} "
                         f"intersections) {secs:.2f}s")
            log(f"cpu baseline {parts[-1]}")
        cpu = {"value": round(1.0 / total, 6), "unit": "fit steps/s", "cores": threads,
               "kind": "port",
               "sample": "one whole two-level fit step: the oracle (C, FP64 raster, OpenMP) "
                         "project+bin+raster fwd/loss grad/raster bwd/proj bwd and torch's "
                         "CPU Adam on each full level; " + "; ".join(parts)}

    # the level-2 kernels on a density-skewed copy of the level (half the
    # Gaussians pulled into the top quarter of the canvas: 2.5x the mean
    # density there, 0.5x elsewhere), where the kernels' launch order has to
    # balance the XCDs; loss + backward only, parameters untouched
    skew = None
    if args.skew_steps > 0:
        sk = fit.Level(cfg["n2"], cfg["H"], cfg["W"], cfg["seed"] * 10 + 2, cfg["scale"],
                       device=dev)
        sk.batch_depth = l2.batch_depth
        with torch.no_grad():
            sk.means[: cfg["n2"] // 2, 1] *= 0.25
            sk.rgbs.copy_(l2.rgbs)  # the fitted level's colours: gradients that move
        sf = KernelTimer("lig_fit_render_l2", dev)
        sb = KernelTimer("lig_fit_backward_records_l2", dev)
        for i in range(args.skew_steps + 1):
            for p_ in (sk.means, sk.cholesky, sk.rgbs):
                p_.grad = None
            if i == 0:
                sk.loss(t2).backward()
                continue
            with sf, sb:
                sk.loss(t2).backward()
        skew = {"workload": "the 8192^2 level with half its Gaussians in the top quarter",
                "raster_fwd_ms": round(sf.mean_ms(), 3), "raster_bwd_ms": round(sb.mean_ms(), 3),
                "uniform_fwd_ms": round(fwd_ms, 3), "uniform_bwd_ms": round(bwd_ms, 3)}
        log(f"skewed level 2 {skew}")
        del sk

    del l1, l2, gt, t1, t2
    batched = batched_run(args, dev, torch) if args.batch_steps > 0 else None
    serve = serve_run(args, dev, torch) if args.serve_requests > 0 else None

    out = {
        "metric": "fit steps/s + render Gpix/s @8192², 2M Gaussians; p50/p99 crop latency",
        "value": round(steps_per_s, 4),
        "unit": "fit steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 / steps_per_s, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator, random-init Gaussians)",
        "config": {"workload": cfg["workload"], "config": args.config,
                   "level1": f"{cfg['H'] // f}x{cfg['W'] // f} x {cfg['n1']}",
                   "level2": f"{cfg['H']}x{cfg['W']} x {cfg['n2']}"},
        "level_ms": {k: round(v / args.steps, 3) for k, v in lvl_ms.items()},
        "render_gpix_s": round(P / render_s / 1e9, 4),
        "render_ms": round(render_s * 1e3, 3),
        "render_api_gpix_s": round(P / render_api_s / 1e9, 4),
        "workload_counts": counts,
        "kernel_ms_l2": {k: round(v["ms"], 3) for k, v in kernels.items()},
        "skewed_l2": skew,
        "roofline": roofline,
        "step_roofline": step_roofline,
        "cpu_baseline": cpu,
        "api_chain": api_chain,
        "captured": captured,
        "autotune": {k: {"tile/depth": v.get("measured_with", v["choice"]), "ms": v["ms"]}
                     for k, v in tuned.items()},
        "batched": batched,
        "serve": serve,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
