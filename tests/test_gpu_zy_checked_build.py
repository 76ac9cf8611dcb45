"""GPU run of the bounds-checked build (liblig_hip_checked.so, built with
LIG_CHECKED): the binning, raster, fit, batched-fit, crop and serving tests run
again in a child process against it, and after each test the failure count
of its device-side bounds checks (LDS ranks and outputs of the sort and the
row expansion, the emission's owner lookups, the backward's row-group sort and
record slots, the launch order, the batched crop kernels) must be zero.  A
failed check is counted, not trapped, so the kernels run to completion either
way."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SUBSET = ["tests/test_gpu_project_bin.py", "tests/test_gpu_raster.py", "tests/test_gpu_fit.py",
          "tests/test_gpu_batch.py", "tests/test_gpu_crop_batch.py", "tests/test_gpu_shares.py"]


def test_suite_passes_with_bounds_checks():
    import lig_amd._lib as L
    assert os.path.exists(L.CHECKED_LIB_PATH), "build with `python -m lig_amd.build`"
    env = dict(os.environ, LIG_LIB_VARIANT="checked", LIG_CHECK_ASSERT="1")
    cmd = [sys.executable, "-u", "-m", "pytest", "-x", "-q", "-m", "gpu", "-p",
           "no:cacheprovider", *SUBSET]
    print("running", " ".join(cmd), flush=True)
    res = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    tail = (res.stdout + res.stderr)[-3000:]
    print(tail, flush=True)
    assert res.returncode == 0, tail
    assert " passed" in res.stdout and "failed" not in res.stdout.split("\n")[-2], tail
