"""Bounce-1 VPL set generation (VplGenerator::generate_vpl with max_bounce = 2,
vpl_generator.cc:127-181; SubpathGenerator::sample_direct_vertex for an area
light, path.cc:171-220, 305-334; AreaLight::sample, area_light.cc:26-44).

A light sample picks an emitting triangle by area, a uniform point on it and a
cosine-weighted direction about its normal; the ray's closest hit on a surface
with a material becomes a VPL with power Le |cos| / (pdf_area * pdf_dir).
Samples that miss or reach a light are emitted but store nothing.  Sampling
is seeded (numpy PCG64) and the tracing is whatever `tracer` the caller hands
in, so the same seed gives the same VPL set on every backend that traces the
same hits.
"""
from __future__ import annotations

import math

import numpy as np

from .scenes import F, SceneData, cross, unit

INV_PI = F(0.31830988618379)
TWO_PI = F(6.28318530717959)


def _onb_from_n(n):
    """Onb::init_from_n (onb.cc:42-54), vectorised"""
    nn = unit(n)
    uu = cross(nn, np.broadcast_to(np.array([1, 0, 0], F), nn.shape))
    small = np.sqrt((uu * uu).sum(-1)) < F(0.01)
    uu2 = cross(nn, np.broadcast_to(np.array([0, 1, 0], F), nn.shape))
    uu = np.where(small[..., None], uu2, uu)
    vv = cross(nn, uu)
    return unit(uu), unit(vv), nn


def light_samples(scene: SceneData, n, rng):
    """n emitted light rays: (rays (n, 6), power (n, 3))."""
    lt = scene.light_tris
    P = scene.positions[lt].reshape(-1, 3, 3)
    p01, p02 = P[:, 0] - P[:, 1], P[:, 0] - P[:, 2]
    cr = cross(p01, p02)
    area = (F(0.5) * np.sqrt((cr * cr).sum(-1))).astype(F)
    total = F(area.sum(dtype=np.float64))
    nrm = unit(cr)
    cdf = np.cumsum(area.astype(np.float64)) / float(total)
    u = rng.random((n, 5)).astype(F)
    k = np.minimum(np.searchsorted(cdf, u[:, 0], side="right"), len(lt) - 1)
    s = np.sqrt(u[:, 1])
    b0, b1 = F(1) - s, u[:, 2] * s
    q = (b0[:, None] * P[k, 0] + b1[:, None] * P[k, 1] + (F(1) - b0 - b1)[:, None] * P[k, 2]).astype(F)
    phi = TWO_PI * u[:, 3]
    sin_t = np.sqrt(u[:, 4])
    x, y = sin_t * np.cos(phi), sin_t * np.sin(phi)
    z = np.sqrt(np.maximum(F(1) - sin_t * sin_t, F(0)))
    uu, vv, nn = _onb_from_n(nrm[k])
    wo = unit(x[:, None] * uu + y[:, None] * vv + z[:, None] * nn)
    pdf_q = F(1) / total
    pdf_wo = z * INV_PI
    light_ix = scene.surf_light[scene.tri_surface[lt[k]]]
    Le = scene.light_radiance[light_ix]
    with np.errstate(divide="ignore", invalid="ignore"):
        power = (Le * z[:, None] / (pdf_q * pdf_wo)[:, None]).astype(F)
    rays = np.concatenate([q, unit(wo)], 1).astype(F)
    ok = pdf_wo > 0
    return rays, power, ok


def generate_vpls(scene: SceneData, tracer, filler, num_vpls: int, seed: int = 2013, batch: int = 65536,
                  max_samples: int = 1 << 26):
    """Returns (vpl (V, 21) f32, vmat (V,) i32, n_emitted)."""
    rng = np.random.default_rng(seed)
    vpls, mats = [], []
    have = 0
    emitted = 0
    while have < num_vpls and emitted < max_samples:
        rays, power, ok = light_samples(scene, batch, rng)
        t, beta, gamma, tri = tracer(rays)
        recv, m, lt = filler(rays, t, beta, gamma, tri)
        valid = ok & (tri >= 0) & (m >= 0) & (lt < 0)
        idx = np.nonzero(valid)[0]
        need = num_vpls - have
        if len(idx) >= need:
            last = idx[need - 1]
            idx = idx[:need]
            emitted += int(last) + 1
        else:
            emitted += batch
        v = np.zeros((len(idx), 21), F)
        v[:, 0:3] = recv[idx, 0:3]          # p
        v[:, 3:6] = recv[idx, 6:9]          # dg.n (shading normal)
        v[:, 6:15] = recv[idx, 12:21]       # dg.uvn
        v[:, 15:18] = unit(-rays[idx, 3:6])  # wi = unit_vector(-r.dir())
        v[:, 18:21] = power[idx]
        vpls.append(v)
        mats.append(m[idx])
        have += len(idx)
    if have < num_vpls:
        raise RuntimeError(f"only {have} VPLs after {emitted} light samples")
    return np.concatenate(vpls).astype(F), np.concatenate(mats).astype(np.int32), float(emitted)


def gpu_tracer(scene_handle):
    """tracer/filler pair on the device scene (renzoku_amd.Scene)."""
    def tracer(rays):
        return scene_handle.trace_closest(rays, 1e-3, 1e6)

    def filler(rays, t, b, g, tri):
        return scene_handle.fill_receivers(rays, t, b, g, tri)

    return tracer, filler


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


__all__ = ["generate_vpls", "light_samples", "gpu_tracer", "tile_centre_pixels", "tile_neighbours", "math"]
