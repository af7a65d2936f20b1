"""The VPL set that feeds the gather, and the pixel-tile slice helpers.

device_vpls is rzk_generate_vpls (the engine's bounce-1 VPL generation,
vpl_generator.cc:127-181) on a resident scene; the oracle restates it
(OrcScene.generate_vpls) and tests/test_gpu_vplgen.py pins the two together.
"""
from __future__ import annotations

import numpy as np

from . import _lib


def device_vpls(scene, num_vpls: int, seed: int = 2013):
    """(vpl (V, 21) f32, vmat (V,) i32, light samples used) from rzk_generate_vpls on `scene`
    (renzoku_amd.Scene)"""
    return _lib.generate_vpls(scene.ctx, scene, num_vpls, seed=seed)


def tile_centre_pixels(width, height, tile):
    """slice representatives of pixel-tile slices (row-major over tiles)"""
    tx, ty = (width + tile - 1) // tile, (height + tile - 1) // tile
    px, py = [], []
    for y in range(ty):
        for x in range(tx):
            px.append(min(x * tile + tile // 2, width - 1))
            py.append(min(y * tile + tile // 2, height - 1))
    return np.asarray(px, np.int32), np.asarray(py, np.int32)


def tile_neighbours(width, height, tile, k):
    """k nearest tiles of every tile by tile-centre distance, ties by index"""
    tx, ty = (width + tile - 1) // tile, (height + tile - 1) // tile
    S = tx * ty
    sx, sy = np.arange(S) % tx, np.arange(S) // tx
    out = np.zeros((S, k), np.int32)
    for s in range(S):
        d = (sx - sx[s]) ** 2 + (sy - sy[s]) ** 2
        order = np.lexsort((np.arange(S), d))
        out[s] = order[:k]
    return out


__all__ = ["device_vpls", "tile_centre_pixels", "tile_neighbours"]
