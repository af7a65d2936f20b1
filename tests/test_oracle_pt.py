"""Pins the oracle's guided path tracer (vpl.cc:934-1041 with the NEE of
area_light.cc:48-142): the Philox stream against the published known-answer
vectors, the BSDF samplers against their own densities (importance-sampling
identities that any dropped factor breaks), the one-bounce estimator against
the closed-form irradiance of a polygonal area light, and the guided
continuation against the unguided one (both unbiased, so their means agree for
ANY guide map; a guide pdf off by a constant factor does not)."""
import math

import numpy as np

import oracle
from renzoku_amd import scenes


def test_philox4x32_10_known_answers():
    # Salmon et al., "Parallel random numbers: as easy as 1, 2, 3" (SC'11),
    # Random123 kat_vectors for philox4x32 with 10 rounds
    cases = [
        ([0, 0, 0, 0], 0, 0, [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, 0xFFFFFFFF, 0xFFFFFFFF, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], 0xA4093822, 0x299F31D0,
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, k0, k1, want in cases:
        assert list(oracle.philox(ctr, k0, k1)) == want


def _frame(n):
    n = np.asarray(n, np.float64)
    n /= np.linalg.norm(n)
    a = np.array([1.0, 0.0, 0.0]) if abs(n[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    u = np.cross(a, n)
    u /= np.linalg.norm(u)
    v = np.cross(n, u)
    return np.concatenate([u, v, n]).astype(np.float32)


def _mat(kind, kd, ks=(0, 0, 0), p0=0.0, p1=0.0):
    m = np.zeros(12, np.float32)
    m[0], m[1:4], m[4:7], m[7], m[8] = kind, kd, ks, p0, p1
    return m


def test_bsdf_sample_density_identity():
    """E[cos(wi) / pdf(wi)] over the sampler's own draws equals the projected
    solid angle of the hemisphere, pi, when pdf is the sampler's density.
    Lambert is exact per sample; the mixtures are checked within 4 standard
    errors (cos/pdf <= pi / r_d is bounded)."""
    uvn = _frame([0.2, 1.0, -0.3])
    n = uvn[6:9].astype(np.float64)
    rng = np.random.default_rng(7)
    mats = [_mat(0, (0.5, 0.4, 0.3)),
            _mat(1, (0.3, 0.3, 0.3), (0.4, 0.4, 0.4), 40.0),
            _mat(2, (0.2, 0.3, 0.2), (0.5, 0.4, 0.5), 0.15, 0.35)]
    for m in mats:
        for wo in ([0.0, 1.0, 0.0], [0.8, 0.5, 0.1]):
            wo = np.asarray(wo, np.float32) / np.linalg.norm(wo)
            vals = []
            for _ in range(6000):
                u = rng.integers(0, 2 ** 32, 3, dtype=np.uint64).astype(np.uint32)
                wi, pdf, f = oracle.bsdf_sample(m, uvn, wo, u)
                c = float(np.dot(wi, n))
                if pdf > 0.0 and c > 0.0:
                    # the returned pdf is the density evaluated at wi
                    assert abs(pdf - oracle.bsdf_pdf(m, uvn, wo, wi)) <= 1e-6 * max(pdf, 1.0)
                    vals.append(c / pdf)
                else:
                    vals.append(0.0)
            vals = np.asarray(vals)
            if m[0] == 0:
                nz = vals[vals > 0]
                assert np.allclose(nz, math.pi, rtol=2e-5)
            se = vals.std() / math.sqrt(len(vals))
            assert abs(vals.mean() - math.pi) < 4 * se + 1e-3, (m[0], wo, vals.mean(), se)


def _plane_light_scene(W, H):
    """a Lambertian floor (y = 0, kd 0.5) under a 100 x 60 quad light facing
    down at y = 100, radiance (2, 3, 4); camera below the light plane, looking
    down steeply enough that no pixel footprint grazes the floor"""
    b = scenes._Builder()
    m = b.material(scenes.LAMBERT, (0.5, 0.5, 0.5))
    b.quads(b.surface(m), [[(-1000, 0, -1000), (-1000, 0, 1000), (1000, 0, 1000), (1000, 0, -1000)]])
    b.lights.append((2.0, 3.0, 4.0))
    y = 100.0
    light = [(-50, y, -30), (50, y, -30), (50, y, 30), (-50, y, 30)]
    b.quads(b.surface(-1, 0), [light])
    cam = scenes.camera((0, 90, -60), (0, 0, 10), (0, 1, 0), 50.0, W, H)
    return b.build(cam, "plane-light"), np.asarray(light, np.float64)


def _polygon_irradiance(p, n, poly):
    """Lambert's point-to-polygon formula: E / Le = 1/2 sum_i theta_i n . unit(r_i x r_i+1)"""
    r = poly[None, :, :] - p[:, None, :]
    r /= np.linalg.norm(r, axis=2, keepdims=True)
    acc = np.zeros(len(p))
    for i in range(len(poly)):
        a, b = r[:, i], r[:, (i + 1) % len(poly)]
        c = np.cross(a, b)
        th = np.arccos(np.clip((a * b).sum(1), -1, 1))
        acc += th * (c @ n) / np.linalg.norm(c, axis=1)
    return 0.5 * np.abs(acc)


def test_one_bounce_matches_polygon_irradiance():
    """max_bounce 1 with direct lighting: the pixel value is the MIS-combined
    NEE estimate of (kd / pi) * E(p), E from Lambert's closed form."""
    W = H = 24
    sc, light = _plane_light_scene(W, H)
    osc = oracle.OrcScene(sc)
    img = oracle.guided_pt(osc, sc.camera, spp=256, max_bounce=1, direct=1, guide_mode=0, seed=11, tmin=1e-3,
                           tile=8, nthreads=8)
    # hit point of each pixel centre (the pixel average differs from it by
    # second-order terms of E over a pixel footprint, well below the tolerance)
    px, py = np.meshgrid(np.arange(W), np.arange(H))
    rays = oracle.camera_rays(sc.camera, px.ravel().astype(np.int32), py.ravel().astype(np.int32))
    t, bb, gg, tri = osc.trace_closest(rays)
    floor = (tri >= 0) & (sc.tri_surface[np.maximum(tri, 0)] == 0)
    r = rays.reshape(-1, 6).astype(np.float64)
    p = r[:, :3] + r[:, 3:] * t[:, None]
    E = _polygon_irradiance(p[floor], np.array([0.0, 1.0, 0.0]), light)
    want = (0.5 / math.pi) * E[:, None] * np.array([2.0, 3.0, 4.0])[None, :]
    got = img.reshape(-1, 3)[floor].astype(np.float64)
    assert floor.sum() == W * H
    ratio = got / want
    assert abs(ratio.mean() - 1.0) < 4e-3, ratio.mean()
    assert np.quantile(np.abs(ratio - 1.0), 0.99) < 0.12   # per-pixel noise: about 2.5% rms at 256 spp
    # the three channels scale exactly with the light's radiance
    assert np.allclose(got[:, 1] / got[:, 0], 1.5, rtol=1e-5)
    assert np.allclose(got[:, 2] / got[:, 0], 2.0, rtol=1e-5)


def _random_guide(rng, nslices, res):
    maps = rng.gamma(0.5, 1.0, (nslices, res, res)).astype(np.float32)
    maps[:, :, : res // 4] = 0.0              # unreachable by the guide: the BSDF strategy must cover it
    totals = maps.reshape(nslices, -1).sum(1).astype(np.float32)
    rows = maps.sum(2)
    row_cdf = np.cumsum(rows / totals[:, None], 1).astype(np.float32)
    row_cdf[:, -1] = 1.0
    col_cdf = np.cumsum(maps / np.maximum(rows[:, :, None], 1e-30), 2).astype(np.float32)
    col_cdf[:, :, -1] = 1.0
    return maps, row_cdf, col_cdf, totals


def test_guided_continuation_is_unbiased():
    """guide_mode 1 (random per-slice maps) and guide_mode 0 estimate the same
    integral; a guide pdf scaled by 1/2 (totals doubled) biases the estimate
    and the same statistic detects it."""
    cs = scenes.cornell_box(32, 32)
    osc = oracle.OrcScene(cs)
    tile, res = 8, 8
    ns = (32 // tile) ** 2
    maps, row_cdf, col_cdf, totals = _random_guide(np.random.default_rng(3), ns, res)
    kw = dict(spp=64, max_bounce=3, direct=1, seed=5, tmin=1e-3, tile=tile, nthreads=8)
    base = oracle.guided_pt(osc, cs.camera, guide_mode=0, **kw)
    guided = oracle.guided_pt(osc, cs.camera, guide_mode=1, res=res, maps=maps, row_cdf=row_cdf, col_cdf=col_cdf,
                              totals=totals, **kw)
    wrong = oracle.guided_pt(osc, cs.camera, guide_mode=1, res=res, maps=maps, row_cdf=row_cdf, col_cdf=col_cdf,
                             totals=totals * 2.0, **kw)

    def z(a, b):
        d = (a - b).mean(2).ravel()
        return d.mean() / (d.std() / math.sqrt(len(d)))

    assert not np.array_equal(base, guided)
    assert abs(z(guided, base)) < 4.0, z(guided, base)
    assert abs(z(wrong, base)) > 6.0, z(wrong, base)


def test_window_and_determinism():
    """a pixel's value depends only on (seed, pixel, sample): rendering a
    window gives the same pixels as the full frame, in any thread count"""
    cs = scenes.cornell_box(24, 24)
    osc = oracle.OrcScene(cs)
    kw = dict(spp=4, max_bounce=4, direct=1, guide_mode=0, seed=9, tmin=1e-3, tile=8)
    full = oracle.guided_pt(osc, cs.camera, nthreads=8, **kw)
    win = oracle.guided_pt(osc, cs.camera, window=(5, 3, 11, 13), nthreads=1, **kw)
    assert np.array_equal(full[3:16, 5:16], win[3:16, 5:16])
    assert not win[:3].any() and not win[:, :5].any()
    assert full.sum() > 0


def test_rays_form_is_the_image_form():
    """guided_pt_rays on the very rays guided_pt draws (camera_sample_rays), summed
    per pixel in sample order and divided by spp, is guided_pt's image bit for bit
    (guided, stratified 4 spp); another sample number changes the value"""
    cs = scenes.cornell_box(24, 20)
    osc = oracle.OrcScene(cs)
    tile, res, spp = 8, 8, 4
    ns = 3 * 3
    maps, row_cdf, col_cdf, totals = _random_guide(np.random.default_rng(4), ns, res)
    g = dict(res=res, maps=maps, row_cdf=row_cdf, col_cdf=col_cdf, totals=totals)
    img = oracle.guided_pt(osc, cs.camera, spp, 4, 1, 1, 11, 1e-3, tile, nthreads=8, **g)
    py, px = np.mgrid[0:20, 0:24]
    px, py = px.ravel().astype(np.int32), py.ravel().astype(np.int32)
    acc = np.zeros((len(px), 3), np.float32)
    for s in range(spp):
        smp = np.full(len(px), s, np.int32)
        rays = oracle.camera_sample_rays(cs.camera, spp, 11, px, py, smp)
        acc += oracle.guided_pt_rays(osc, 24, 20, 4, 1, 1, 11, 1e-3, tile, rays, np.stack([px, py], 1), smp, **g)
    acc /= np.float32(spp)
    assert np.array_equal(acc.reshape(20, 24, 3).view(np.uint32), img.view(np.uint32))
    rays = oracle.camera_sample_rays(cs.camera, spp, 11, px, py, np.zeros(len(px), np.int32))
    a = oracle.guided_pt_rays(osc, 24, 20, 4, 1, 1, 11, 1e-3, tile, rays, np.stack([px, py], 1),
                              np.zeros(len(px), np.int32), **g)
    b = oracle.guided_pt_rays(osc, 24, 20, 4, 1, 1, 11, 1e-3, tile, rays, np.stack([px, py], 1),
                              np.full(len(px), 7, np.int32), **g)
    assert not np.array_equal(a, b)
