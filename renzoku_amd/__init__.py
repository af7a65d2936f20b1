"""renzoku_amd — MI355X-native engine for Renzoku's data-parallel hot path.

The compute lives in librzk.so (gfx950 HIP kernels behind the C ABI of
include/rzk.h).  This package holds its ctypes binding (`_lib`), the seeded
synthetic scenes the workloads are defined on (`scenes`), the VPL set of a
resident scene through the engine's generator and the pixel-tile helpers
(`vplgen`), scene / frame files and the render-configuration importer.
"""
from . import _lib, scenes, vplgen  # noqa: F401
from ._lib import (Batch, Clusters, Ctx, Guide, KdSlices, LightSlice, RzkError, Scene, Server, Vpls,  # noqa: F401
                   alphas, cluster_assign, gather, generate_vpls, map_vec6, reduced_matrix, refine_slices,
                   render_vpl, transport_matrix, visibility)

__all__ = ["Ctx", "Scene", "Vpls", "LightSlice", "Guide", "Batch", "Clusters", "Server", "RzkError", "gather", "render_vpl", "visibility",
           "reduced_matrix", "transport_matrix", "alphas", "cluster_assign", "refine_slices", "generate_vpls", "KdSlices", "map_vec6",
           "scenes", "vplgen"]
