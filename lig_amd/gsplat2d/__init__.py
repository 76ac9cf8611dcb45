"""gsplat2d API on MI355X (gfx950) HIP kernels.

Drop-in for the reference package (gsplat2d/gsplat2d/__init__.py:1-135): the
same public names, plus the aliases `project_gaussians_2d` (= the Cholesky
projection) and `rasterize_gaussians_sum`.  `install_alias()` registers this
package and its submodules under the name `gsplat2d` so that unmodified
callers (`from gsplat2d.rasterize import rasterize_gaussians`) resolve here.
"""
import importlib
import sys
import warnings
from typing import Any

import torch

from .project_gaussians import project_gaussians
from .project_gaussians_cholesky import project_gaussians_cholesky
from .rasterize import rasterize_gaussians
from .upscale import gradient_aware_upscale
from .utils import (
    bin_and_sort_gaussians,
    compute_cov2d_bounds,
    compute_cumulative_intersects,
    get_tile_bin_edges,
    map_gaussian_to_intersects,
)
from .version import __version__

project_gaussians_2d = project_gaussians_cholesky
rasterize_gaussians_sum = rasterize_gaussians

__all__ = [
    "__version__",
    "project_gaussians",
    "rasterize_gaussians",
    "project_gaussians_cholesky",
    "project_gaussians_2d",
    "rasterize_gaussians_sum",
    "bin_and_sort_gaussians",
    "compute_cumulative_intersects",
    "compute_cov2d_bounds",
    "get_tile_bin_edges",
    "map_gaussian_to_intersects",
    "gradient_aware_upscale",
    "ProjectGaussians",
    "RasterizeGaussians",
    "BinAndSortGaussians",
    "ComputeCumulativeIntersects",
    "ComputeCov2dBounds",
    "GetTileBinEdges",
    "MapGaussiansToIntersects",
]

_SUBMODULES = ("project_gaussians", "project_gaussians_cholesky", "rasterize", "upscale",
               "utils", "cuda", "version")


def install_alias(name: str = "gsplat2d") -> None:
    """Make `import gsplat2d[.submodule]` resolve to this package."""
    pkg = sys.modules[__name__]
    sys.modules[name] = pkg
    for sub in _SUBMODULES:
        sys.modules[f"{name}.{sub}"] = importlib.import_module(f"{__name__}.{sub}")


def _deprecated(cls_name, fn, new_name):
    """Backwards-compatible autograd.Function names (gsplat2d/__init__.py:41-135):
    forward warns and forwards to the functional API, backward is not provided."""

    class _Shim(torch.autograd.Function):
        @staticmethod
        def forward(ctx, *args, **kwargs):
            warnings.warn(f"{cls_name} is deprecated, use {new_name} instead",
                          DeprecationWarning)
            return fn(*args, **kwargs)

        @staticmethod
        def backward(ctx: Any, *grad_outputs: Any) -> Any:
            raise NotImplementedError

    _Shim.__name__ = _Shim.__qualname__ = cls_name
    return _Shim


MapGaussiansToIntersects = _deprecated("MapGaussiansToIntersects", map_gaussian_to_intersects,
                                       "map_gaussian_to_intersects")
ComputeCumulativeIntersects = _deprecated("ComputeCumulativeIntersects",
                                          compute_cumulative_intersects,
                                          "compute_cumulative_intersects")
ComputeCov2dBounds = _deprecated("ComputeCov2dBounds", compute_cov2d_bounds,
                                 "compute_cov2d_bounds")
GetTileBinEdges = _deprecated("GetTileBinEdges", get_tile_bin_edges, "get_tile_bin_edges")
BinAndSortGaussians = _deprecated("BinAndSortGaussians", bin_and_sort_gaussians,
                                  "bin_and_sort_gaussians")
ProjectGaussians = _deprecated("ProjectGaussians", project_gaussians, "project_gaussians")
RasterizeGaussians = _deprecated("RasterizeGaussians", rasterize_gaussians,
                                 "rasterize_gaussians")
