"""midas_amd -- MI355X-native (gfx950, HIP) far-field caking of MIDAS.

The public names follow the reference package ``midas_integrate``
(packages/midas_integrate/midas_integrate/__init__.py:20-150); the compute
behind them is libmidas_caking.so (include/midas_caking.h).
"""
from __future__ import annotations

__version__ = "0.1.0"

from .bin_io import (MAP_HEADER_MAGIC, MAP_HEADER_SIZE, PXLIST_DTYPE, MapHeader, PixelMap,
                     canonical_string, header_for, load_map, write_map)
from .params import (IntegrationParams, bin_edges, parse_integrator_params,
                     parse_mapper_params, parse_params)
from .caking import (AREA_THRESHOLD, CSRGeometry, Context, Event, build_csr, build_map,
                     IntegrationGraph, default_context, hybrid_message, integrate,
                     integrate_async, integrate_batch, integrate_hybrid,
                     integrate_with_variance,
                     load_map_gpu, profile_1d, profile_1d_with_variance, time_integrate)
from .outputs import ResultFiles
from .postprocess import gauss_smooth_eta, median_filter_eta
from .fused_csr import build_fused_geometry
from .zarr_out import enrich_zarr

__all__ = [
    "__version__", "MAP_HEADER_MAGIC", "MAP_HEADER_SIZE", "PXLIST_DTYPE", "MapHeader",
    "PixelMap", "canonical_string", "header_for", "load_map", "write_map",
    "IntegrationParams", "bin_edges", "parse_integrator_params", "parse_mapper_params",
    "parse_params", "AREA_THRESHOLD", "CSRGeometry", "Context", "build_csr", "build_map",
    "IntegrationGraph", "default_context", "integrate", "integrate_batch",
    "integrate_with_variance", "Event", "integrate_async", "integrate_hybrid",
    "hybrid_message",
    "load_map_gpu", "profile_1d", "profile_1d_with_variance", "time_integrate", "ResultFiles", "gauss_smooth_eta",
    "median_filter_eta", "build_fused_geometry", "enrich_zarr",
]
