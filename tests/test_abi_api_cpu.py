"""CPU-side checks of the boundary: the C-ABI library loads and exports every
symbol include/lig_abi.h declares; the Python API keeps the reference's
asserts, error types and no-fallback behaviour (no compute calls here)."""
import hashlib
import os
import re

import numpy as np
import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "lig_abi.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lig_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import ctypes
    import lig_amd._lib as L
    lib = L.load()
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(L.EXPORTED) == set(names)
    assert lib.lig_abi_version() >= 100
    assert lib.lig_strerror(-1) == b"invalid argument"
    # size queries are pure host functions
    assert L.size_query("lig_packed_gaussians_bytes", 10) == 640
    assert L.size_query("lig_bin_workspace_size", 0, 16) > 0


def test_backward_shares_override_is_host_state():
    """lig_set_backward_shares only sets host state: 0 (automatic) or 1..8,
    returning the previous value; other values are rejected."""
    import lig_amd._lib as L
    lib = L.load()
    prev = lib.lig_set_backward_shares(3)
    try:
        assert lib.lig_set_backward_shares(8) == 3
        assert lib.lig_set_backward_shares(9) == -1
        assert lib.lig_set_backward_shares(-1) == -1
        assert lib.lig_set_backward_shares(0) == 8
    finally:
        lib.lig_set_backward_shares(prev)


def test_tile_order_workspace_size():
    """The launch-order workspace holds one int per tile plus the class
    counts, band loads and cursors; none for an empty grid."""
    import lig_amd._lib as L
    assert L.size_query("lig_tile_order_workspace_size", 0) == 0
    for t in (1, 7, 16384, 262144):
        n = L.size_query("lig_tile_order_workspace_size", t)
        assert n >= 4 * t + 8 * 8 + 2 * 128 * 4, (t, n)
        assert n % 8 == 0


def test_invalid_arguments_rejected_without_gpu():
    import ctypes
    import lig_amd._lib as L
    with pytest.raises(L.LigError, match="invalid argument"):
        L.call("lig_project_gaussians_forward", 10, None, None, None, 64, 64, 17, None, None,
               None, None, None)
    with pytest.raises(L.LigError, match="invalid argument"):
        L.call("lig_gradient_aware_upscale_forward", 4, 4, 8, 8, 3.0, 0.0, 1.0, 4.0, None,
               None, None, None, None, None)


def test_public_names():
    import lig_amd.gsplat2d as g
    for name in ("project_gaussians", "project_gaussians_cholesky", "rasterize_gaussians",
                 "gradient_aware_upscale", "map_gaussian_to_intersects",
                 "bin_and_sort_gaussians", "compute_cumulative_intersects",
                 "compute_cov2d_bounds", "get_tile_bin_edges", "project_gaussians_2d",
                 "rasterize_gaussians_sum", "ProjectGaussians", "RasterizeGaussians"):
        assert hasattr(g, name), name
    from lig_amd.gsplat2d import cuda as C
    for name in ("rasterize_forward", "rasterize_backward", "project_gaussians_forward",
                 "project_gaussians_backward", "project_gaussians_forward_cholesky",
                 "project_gaussians_backward_cholesky", "compute_cov2d_bounds",
                 "map_gaussian_to_intersects", "get_tile_bin_edges",
                 "gradient_aware_upscale_forward", "gradient_aware_upscale_backward"):
        assert callable(getattr(C, name)), name


def test_alias_install():
    import sys
    import lig_amd.gsplat2d as g
    g.install_alias("gsplat2d_alias_test")
    mod = __import__("gsplat2d_alias_test.rasterize", fromlist=["rasterize_gaussians"])
    assert mod.rasterize_gaussians is g.rasterize_gaussians
    for k in [k for k in sys.modules if k.startswith("gsplat2d_alias_test")]:
        del sys.modules[k]


def test_reference_asserts_and_errors():
    import lig_amd.gsplat2d as g
    x = torch.zeros(4, 2)
    with pytest.raises(AssertionError):
        g.project_gaussians(torch.zeros(4, 3), x, 16, 16, 17)
    with pytest.raises(AssertionError):
        g.rasterize_gaussians(x, x, torch.zeros(4, 3), torch.zeros(4, dtype=torch.int32),
                              torch.zeros(4, 3), None, 16, 16, 1)
    with pytest.raises(ValueError):
        g.rasterize_gaussians(torch.zeros(4, 3), x, torch.zeros(4, 3),
                              torch.zeros(4, dtype=torch.int32), torch.zeros(4, 3), None, 16,
                              16, 16)
    with pytest.raises(ValueError):
        g.rasterize_gaussians(x, x, torch.zeros(4, 3), torch.zeros(4, dtype=torch.int32),
                              torch.zeros(4), None, 16, 16, 16)
    with pytest.raises(ValueError):
        g.project_gaussians(torch.zeros(0, 3), torch.zeros(0, 2), 16, 16, 16)


def test_no_cpu_fallback():
    """CPU tensors are refused (the reference's CHECK_CUDA), never computed."""
    import lig_amd.gsplat2d as g
    with pytest.raises(RuntimeError, match="CUDA tensor"):
        g.project_gaussians_cholesky(torch.ones(4, 3), torch.zeros(4, 2), 16, 16, 16)
    with pytest.raises(RuntimeError, match="CUDA tensor"):
        g.gradient_aware_upscale(*(torch.zeros(4, 4, 3) for _ in range(4)), 8, 8)


def test_product_does_not_import_oracle():
    """Nothing under lig_amd/ imports, loads or names the oracle package."""
    pat = re.compile(r"^\s*(import oracle|from oracle)|liblig_oracle", re.M)
    for dirpath, _, files in os.walk(os.path.join(ROOT, "lig_amd")):
        for f in files:
            if f.endswith((".py", ".hip", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert not pat.search(text), os.path.join(dirpath, f)


def test_c1_image_is_deterministic():
    from lig_amd import data
    img = data.make_image(512, 512, 0)
    assert img.shape == (512, 512, 3) and float(img.min()) == 0.0 and float(img.max()) == 1.0
    digest = hashlib.sha256(img.numpy().tobytes()).hexdigest()
    assert digest == "5c88d446e10635477bdef2e44d34e5aa2050dc40a16ad78f1ac9e6bcfdb3bfe0"


def test_crop_batch_tables_match_the_c_structs(tmp_path):
    """The host tables of batched crop rendering (render._VIEW_DTYPE,
    render._CROP_DTYPE) have the field offsets and sizes of LigView and
    LigCrop as a C compiler lays them out from lig_abi.h."""
    import subprocess
    from lig_amd import render
    fields = {"LigView": render._VIEW_DTYPE, "LigCrop": render._CROP_DTYPE}
    src = ["#include <stdio.h>", "#include <stddef.h>", '#include "lig_abi.h"',
           "int main(void) {"]
    for st, dt in fields.items():
        src.append(f'printf("{st} sizeof %zu\\n", sizeof({st}));')
        for name in dt.names:
            if name == "reserved":
                continue
            src.append(f'printf("{st} {name} %zu\\n", offsetof({st}, {name}));')
    src += ["return 0;", "}"]
    c = os.path.join(tmp_path, "layout.c")
    exe = os.path.join(tmp_path, "layout")
    open(c, "w").write("\n".join(src))
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), "-o", exe, c],
                   check=True)
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {tuple(line.split()[:2]): int(line.split()[2]) for line in out if line}
    for st, dt in fields.items():
        assert got[(st, "sizeof")] == dt.itemsize
        for name in dt.names:
            if name != "reserved":
                assert got[(st, name)] == dt.fields[name][1], (st, name)


def test_checked_build_exports_the_same_symbols():
    """The bounds-checked variant exports every declared entry point too, and
    the normal build reports that it has no checks (LIG_ERR_UNSUPPORTED)."""
    import ctypes
    import lig_amd._lib as L
    if not os.path.exists(L.CHECKED_LIB_PATH):
        pytest.skip("bounds-checked build not present")
    lib = ctypes.CDLL(L.CHECKED_LIB_PATH)
    for n in _declared():
        assert hasattr(lib, n), n
    assert L.check_status() is None or os.environ.get("LIG_LIB_VARIANT") == "checked"
