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


def _short(name: str) -> str:
    """Kernel name without namespaces, argument list and template noise."""
    base = name.replace("(anonymous namespace)::", "").split("(")[0]
    for pre in ("void ", "lig::", "at::native::"):
        base = base.replace(pre, "")
    return base.strip()


def summarize_rocpd(db_path: str, top: int = 25, by_grid: bool = True) -> str:
    """Per-kernel statistics from a rocprofv3 --kernel-trace database, grouped
    by (kernel, grid size) so launches of one kernel at different image sizes
    are reported separately.  Durations in microseconds."""
    import sqlite3
    con = sqlite3.connect(db_path)
    rows = con.execute("select name, duration, grid_x, grid_y, grid_z from kernels").fetchall()
    groups = {}
    for name, dur, gx, gy, gz in rows:
        key = (_short(name), gx * gy * gz if by_grid else 0)
        groups.setdefault(key, []).append(dur / 1e3)
    total = sum(sum(v) for v in groups.values())
    lines = [f"{'kernel':<48} {'threads':>10} {'calls':>6} {'avg_us':>10} {'min_us':>10} "
             f"{'max_us':>10} {'total_ms':>9} {'pct':>6}"]
    for (k, grid), v in sorted(groups.items(), key=lambda kv: -sum(kv[1]))[:top]:
        lines.append(f"{k[:48]:<48} {grid:>10} {len(v):>6} {sum(v) / len(v):>10.1f} "
                     f"{min(v):>10.1f} {max(v):>10.1f} {sum(v) / 1e3:>9.3f} "
                     f"{100 * sum(v) / total:>5.1f}%")
    lines.append(f"total kernel time {total / 1e3:.3f} ms over {len(rows)} dispatches")
    return "\n".join(lines)


def copy_probe(nbytes: int = 1 << 30) -> None:
    """One device-to-device copy of `nbytes` (torch's copy kernel): the
    reference reading for scaling FETCH_SIZE / WRITE_SIZE to bytes."""
    n = nbytes // 4
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {nbytes} bytes")


def summarize_pmc(db_path: str, pattern: str = "") -> str:
    """Per (kernel, launch size) mean of every collected counter per dispatch."""
    import sqlite3
    con = sqlite3.connect(db_path)
    cols = [r[1] for r in con.execute("pragma table_info(counters_collection)")]
    name_col = "kernel_name" if "kernel_name" in cols else "name"
    grid_col = "grid_size" if "grid_size" in cols else ("grid_x" if "grid_x" in cols else None)
    q = (f"select {name_col}, {grid_col if grid_col else '0'}, counter_name, value "
         f"from counters_collection")
    groups = {}
    for name, grid, cname, val in con.execute(q):
        if pattern and pattern not in name:
            continue
        key = (_short(name), grid, cname)
        groups.setdefault(key, []).append(float(val))
    lines = [f"{'kernel':<48} {'threads':>10} {'counter':<12} {'dispatches':>10} {'mean':>14}"]
    for (k, grid, cname), v in sorted(groups.items(), key=lambda kv: -sum(kv[1])):
        lines.append(f"{k[:48]:<48} {grid:>10} {cname:<12} {len(v):>10} {sum(v) / len(v):>14.1f}")
    return "\n".join(lines)


def _pmc_means(path: str) -> dict:
    """{(kernel short name, launch threads, counter): (mean per dispatch,
    dispatches)} over every *_results.db under `path`."""
    import glob
    import sqlite3
    out = {}
    dbs = [path] if path.endswith(".db") else sorted(
        glob.glob(f"{path}/**/*_results.db", recursive=True))
    for db in dbs:
        con = sqlite3.connect(db)
        cols = [r[1] for r in con.execute("pragma table_info(counters_collection)")]
        name_col = "kernel_name" if "kernel_name" in cols else "name"
        grid_col = "grid_size" if "grid_size" in cols else "grid_x"
        groups = {}
        for name, grid, cname, val in con.execute(
                f"select {name_col}, {grid_col}, counter_name, value from counters_collection"):
            groups.setdefault((_short(name), int(grid), cname), []).append(float(val))
        for k, v in groups.items():
            out[k] = (sum(v) / len(v), len(v))
    return out


def valu_issue(path: str, top: int = 16, cus: int = 256, xcds: int = 8) -> str:
    """VALU issue rate per (kernel, launch size) from one rocprofv3 run with
    `--pmc SQ_INSTS_VALU SQ_INSTS_SALU SQ_INSTS_LDS SQ_WAVES SQ_BUSY_CYCLES
    GRBM_GUI_ACTIVE GRBM_COUNT --kernel-trace`.  GRBM_COUNT counts GPU clocks
    over the dispatch summed over the `xcds` dies, so a CU's cycles are
    GRBM_COUNT / xcds and the clock is that over the dispatch's duration;
    the rates are wave-level VALU instructions per second (G) and per CU per
    cycle."""
    import glob
    import sqlite3
    dbs = [path] if path.endswith(".db") else sorted(
        glob.glob(f"{path}/**/*_results.db", recursive=True))
    per = {}
    meta = {}
    for db in dbs:
        con = sqlite3.connect(db)
        for did, name, grid, cname, val, dur in con.execute(
                "select dispatch_id, kernel_name, grid_size, counter_name, value, duration "
                "from counters_collection"):
            per.setdefault((db, did), {})[cname] = float(val)
            meta[(db, did)] = (_short(name), int(grid), float(dur))
    groups = {}
    for key, cs in per.items():
        name, grid, dur = meta[key]
        groups.setdefault((name, grid), []).append((cs, dur))
    lines = [f"{'kernel':<44} {'threads':>9} {'disp':>5} {'dur_us':>9} {'VALU':>10} "
             f"{'SALU':>9} {'LDS':>9} {'VALU_G/s':>8} {'clk_GHz':>7} {'VALU/CU/clk':>11}"]
    for (name, grid), lst in sorted(groups.items(),
                                    key=lambda kv: -sum(d for _, d in kv[1]))[:top]:
        n = len(lst)

        def mean(c):
            return sum(cs.get(c, 0.0) for cs, _ in lst) / n

        dur = sum(d for _, d in lst) / n
        cyc = mean("GRBM_COUNT") / xcds
        rate = mean("SQ_INSTS_VALU") / (cus * cyc) if cyc > 0 else float("nan")
        clk = cyc / dur if dur > 0 else float("nan")
        lines.append(f"{name[:44]:<44} {grid:>9} {n:>5} {dur / 1e3:>9.1f} "
                     f"{mean('SQ_INSTS_VALU'):>10.4g} {mean('SQ_INSTS_SALU'):>9.3g} "
                     f"{mean('SQ_INSTS_LDS'):>9.3g} {mean('SQ_INSTS_VALU') / dur:>8.1f} "
                     f"{clk:>7.2f} {rate:>11.3f}")
    return "\n".join(lines)


# kernels whose HBM traffic bench.py reports: (name prefix, launch threads)
TRAFFIC_KERNELS = {
    "raster_bwd": ("raster_bwd16_kernel<false, true, true", 16777216,
                   "fit-step backward, 8192^2"),
    "raster_fwd": ("raster_fwd16_kernel<false, true", 16777216, "fit-step forward, 8192^2"),
}


def traffic(fetch: str, write: str, copy_fetch: str, copy_write: str, copy_bytes: int,
            workload: str, commands: list) -> dict:
    """HBM bytes per launch of TRAFFIC_KERNELS from FETCH_SIZE and WRITE_SIZE
    runs (KiB per dispatch), each counter scaled by what it reads for a
    device copy of `copy_bytes` (copy-probe)."""
    cf = max(v[0] for (k, g, c), v in _pmc_means(copy_fetch).items() if c == "FETCH_SIZE")
    cw = max(v[0] for (k, g, c), v in _pmc_means(copy_write).items() if c == "WRITE_SIZE")
    fs, ws = copy_bytes / (cf * 1024.0), copy_bytes / (cw * 1024.0)
    fm, wm = _pmc_means(fetch), _pmc_means(write)
    rec = {"workload": workload,
           "calibration": {"copy_bytes": copy_bytes, "FETCH_SIZE_KiB": round(cf, 1),
                           "WRITE_SIZE_KiB": round(cw, 1), "fetch_scale": round(fs, 5),
                           "write_scale": round(ws, 5)},
           "commands": commands, "kernels": {}}
    for key, (prefix, threads, what) in TRAFFIC_KERNELS.items():
        # the instance launched most often: the fit steps' (the autotuner's
        # trial depths run once or twice each)
        f = sorted(((v[1], k, v[0]) for (k, g, c), v in fm.items()
                    if k.startswith(prefix) and g == threads and c == "FETCH_SIZE"),
                   reverse=True)
        if not f:
            continue
        name = f[0][1]
        w = [v[0] for (k, g, c), v in wm.items()
             if k == name and g == threads and c == "WRITE_SIZE"]
        if not w:
            continue
        rec["kernels"][key] = {
            "kernel": f"{name} ({what})", "launch_threads": threads, "dispatches": f[0][0],
            "FETCH_SIZE_KiB": round(f[0][2], 1), "WRITE_SIZE_KiB": round(w[0], 1),
            "traffic_bytes": int(round(f[0][2] * 1024 * fs + w[0] * 1024 * ws))}
    return rec


if __name__ == "__main__":
    import argparse
    import glob
    ap = argparse.ArgumentParser(description="rocprofv3 helpers")
    ap.add_argument("cmd", choices=["summarize", "pmc", "copy-probe", "traffic", "valu"])
    ap.add_argument("path", nargs="?", help="rocprofv3 output directory or *_results.db file")
    ap.add_argument("--top", type=int, default=25)
    ap.add_argument("--match", default="")
    ap.add_argument("--fetch", help="traffic: FETCH_SIZE run directory")
    ap.add_argument("--write", help="traffic: WRITE_SIZE run directory")
    ap.add_argument("--copy-fetch", help="traffic: copy-probe FETCH_SIZE run directory")
    ap.add_argument("--copy-write", help="traffic: copy-probe WRITE_SIZE run directory")
    ap.add_argument("--copy-bytes", type=int, default=1 << 30)
    ap.add_argument("--workload", default="c3")
    ap.add_argument("--command", action="append", default=[],
                    help="traffic: a command that produced the runs (repeatable)")
    a = ap.parse_args()
    if a.cmd == "copy-probe":
        copy_probe()
    elif a.cmd == "valu":
        print(valu_issue(a.path, a.top))
    elif a.cmd == "traffic":
        import json
        print(json.dumps(traffic(a.fetch, a.write, a.copy_fetch, a.copy_write, a.copy_bytes,
                                 a.workload, a.command), indent=1))
    else:
        dbs = [a.path] if a.path.endswith(".db") else sorted(
            glob.glob(f"{a.path}/**/*_results.db", recursive=True))
        for db in dbs:
            print(summarize_rocpd(db, a.top) if a.cmd == "summarize" else summarize_pmc(db, a.match))
