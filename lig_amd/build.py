"""Build the gfx950 HIP library ``lig_amd/liblig_hip.so`` (and its
bounds-checked variant ``liblig_hip_checked.so``) with hipcc.

Run ``python -m lig_amd.build`` (or ``__graft_entry__.build()``).  The library is
built in-tree so that it travels with the repository snapshot.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG_DIR, "csrc")
LIB_PATH = os.path.join(PKG_DIR, "liblig_hip.so")
CHECKED_LIB_PATH = os.path.join(PKG_DIR, "liblig_hip_checked.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # keep the compiler from emitting packed FP32 (v_pk_*_f32) instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")))


def build(verbose: bool = False) -> str:
    """Build the library and its bounds-checked variant (-DLIG_CHECKED,
    ``liblig_hip_checked.so``) side by side; returns the library's path."""
    jobs = []
    for path, extra in ((LIB_PATH, []), (CHECKED_LIB_PATH, ["-DLIG_CHECKED"])):
        tmp = path + ".tmp"
        cmd = [HIPCC, *FLAGS, *extra, "-o", tmp, *sources()]
        if verbose:
            print(" ".join(cmd), flush=True)
        jobs.append((path, tmp, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                                 stderr=subprocess.PIPE, text=True)))
    failed = []
    for path, tmp, proc in jobs:
        out, err = proc.communicate()
        if proc.returncode != 0:
            sys.stderr.write(out + err)
            failed.append(os.path.basename(path))
        else:
            os.replace(tmp, path)
    if failed:
        raise RuntimeError(f"hipcc failed building {', '.join(failed)}")
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
