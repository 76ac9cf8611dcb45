"""bench.py's contract on the CPU: the one-GPU guard, the algorithmic FLOP and
byte tallies behind the roofline fields, and the committed HBM-traffic record
they are read beside.  The tallies are pinned by hand-counted small cases and
by the figures of a recorded C3 run (level 2 of the 8192^2 fit: the pair
counts E, C and intersections I that run printed, and the algorithmic FLOPs,
bytes and stage roofline it reported for them)."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402  (imports no torch at module level)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_raster_tallies_small_cases():
    # one pair inside the box that contributes: 17 + 7 forward, 17 + 29 backward
    assert bench.raster_flops(1, 1, False) == 24
    assert bench.raster_flops(1, 1, True) == 46
    # a pair in the box that does not contribute costs only the 17
    assert bench.raster_flops(1, 0, True) == 17
    # bytes: 36 per intersection, 8 per tile, 16 per pixel, and 40 per Gaussian backward
    assert bench.raster_bytes(1, 0, 0, 0, False) == 36
    assert bench.raster_bytes(0, 1, 0, 0, False) == 8
    assert bench.raster_bytes(0, 0, 1, 0, False) == 16
    assert bench.raster_bytes(0, 0, 0, 1, False) == 0
    assert bench.raster_bytes(0, 0, 0, 1, True) == 40


def test_raster_tallies_match_a_recorded_c3_run():
    N, I, T, P = 2_000_000, 79_052_767, 262_144, 8192 * 8192
    E, C = 14_339_165_451, 11_008_487_655
    assert bench.raster_flops(E, C, True) == 563_011_954_662
    assert bench.raster_bytes(I, T, P, N, True) == 4_001_738_588
    st = bench.level_stages(N, I, T, P, E, C)
    assert round(bench.stage_roof_ms(*st["raster_bwd"]), 4) == 3.5792
    assert round(bench.stage_roof_ms(*st["raster_fwd"]), 4) == 2.0396
    # the backward is bound by the FP32 VALU, the tile intersection by HBM
    fl, nb = st["raster_bwd"]
    assert fl / (bench.PEAK_FP32_TFLOPS * 1e12) > nb / (bench.PEAK_HBM_GBS * 1e9)
    assert st["intersect"][0] == 0


def test_stage_roof_takes_the_slower_bound():
    one_ms_flops = bench.PEAK_FP32_TFLOPS * 1e9      # 1 ms at the FP32 peak
    one_ms_bytes = bench.PEAK_HBM_GBS * 1e6          # 1 ms at the HBM peak
    assert bench.stage_roof_ms(one_ms_flops, 0) == pytest.approx(1.0)
    assert bench.stage_roof_ms(0, one_ms_bytes) == pytest.approx(1.0)
    assert bench.stage_roof_ms(2 * one_ms_flops, one_ms_bytes) == pytest.approx(2.0)


def test_committed_traffic_record_is_consistent():
    with open(bench.TRAFFIC_FILES["c3"]) as f:
        rec = json.load(f)
    cal = rec["calibration"]
    # the copy probe moves copy_bytes each way; each counter's scale is its
    # reading for that copy against the bytes moved
    assert cal["fetch_scale"] == pytest.approx(cal["copy_bytes"] / (cal["FETCH_SIZE_KiB"] * 1024),
                                               rel=1e-4)
    assert cal["write_scale"] == pytest.approx(cal["copy_bytes"] / (cal["WRITE_SIZE_KiB"] * 1024),
                                               rel=1e-4)
    for name, k in rec["kernels"].items():
        scaled = (k["FETCH_SIZE_KiB"] * cal["fetch_scale"] +
                  k["WRITE_SIZE_KiB"] * cal["write_scale"]) * 1024
        assert k["traffic_bytes"] == pytest.approx(scaled, rel=1e-4), name
        got, path = bench.measured_traffic("c3", name)
        assert got == k["traffic_bytes"] and path == os.path.join("profiles",
                                                                   "raster_c3_8192_traffic.json")
    assert bench.measured_traffic("c3", "no_such_kernel") == (None, None)
    assert bench.measured_traffic("c2", "raster_bwd") == (None, None)
