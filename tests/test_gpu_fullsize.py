"""The benchmark configurations at their full sizes, against the oracle.

* The LightSlice step exactly as the headline bench runs it: seeded room of
  200k triangles, 512x512, 65,536 VPLs, 32x32-pixel slices, 300 clusters.
  Primary hits bit-equal, V exact, R <= 1e-4, clustering bit-equal to the
  oracle's replay on the device's R, the final gather <= 1e-4 (rel-L2).
* The batched mix: 256 uneven jobs on the 1M-triangle scene (windows 32..256 px,
  1-64 spp, guides from 4k-100k VPLs) in one launch; a seeded sample of jobs
  bit-equal to the plain drop-in, and one job against the oracle's guided path
  tracer (<= 1e-6).
* The served form on the 2M-triangle scene with 200k resident VPLs: requests
  alone = in a mix = plain, bit for bit, and one request against the oracle."""
import math

import numpy as np
import pytest

import oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def test_c2_headline_configuration(ctx):
    import renzoku_amd as rz
    from renzoku_amd import scenes, vplgen
    W = H = 512
    sc = scenes.room(W, H, seed=2013, target_tris=200_000)
    gs = rz.Scene(ctx, sc)
    vpl, vmat, ne = vplgen.device_vpls(gs, 65_536, seed=2013)
    vp = rz.Vpls(ctx, vpl, vmat)
    ls = rz.LightSlice(ctx, gs, vp, width=W, height=H, tile=32, num_clusters=300, num_neighbor_slices=16,
                       map_resolution=32, map_epsilon=1e-3, clamp_mode=1, clamp=0.01, incident_clamp=0.01,
                       tmin=1e-3, seed=2013, map_gate=0, n_emitted=ne)
    ls.initialize(sc.camera)
    img = ls.gather()
    inf = ls.info()
    S, V = inf.num_slices, len(vmat)
    assert (S, V) == (256, 65_536)
    print(f"C2 full size: {sc.ntri} triangles, {S} slices, {inf.num_clusters} clusters", flush=True)
    osc = oracle.OrcScene(sc)
    # primary hits of every pixel, bit for bit
    px, py = scenes.pixel_grid(W, H)
    prays = oracle.camera_rays(sc.camera, px, py)
    t, b, g, tri = osc.trace_closest(prays)
    hits = ls.get("hits", np.float32, W * H * 4).reshape(-1, 4)
    assert np.array_equal(hits[:, 3].view(np.int32), tri)
    # reduced matrix of the slice rows
    row_pix = ls.get("row_pix", np.int32, S)
    recv, m, _ = osc.fill_receivers(prays, t, b, g, tri)
    R = ls.get("R", np.float32, S * V * 3).reshape(V, S, 3)
    Vd = ls.get("V", np.float32, S * V).reshape(V, S)
    oR, oV = osc.reduced_matrix(recv[row_pix], m[row_pix], vpl, vmat, 1, 0.01, 1e-3, nthreads=16)
    assert np.array_equal(Vd, oV)
    assert rel_l2(R, oR) <= 1e-4
    # clustering of the device's R, replayed by the oracle on the same MT stream
    oc = oracle.cluster_by_sampling(R, 300, 2013, nthreads=16)
    assert np.array_equal(ls.get("centers", np.int32, inf.num_centers), oc["centers"])
    assert np.array_equal(ls.get("assign", np.int32, V), oc["assign"])
    reps = ls.get("reps", np.int32, inf.num_clusters)
    rep_w = ls.get("rep_w", np.float32, inf.num_clusters)
    assert np.array_equal(reps, oc["reps"])
    assert np.array_equal(rep_w.view(np.uint32), oc["rep_w"].view(np.uint32))
    # the final gather: every pixel over the 300 representatives
    ref = osc.gather_clusters(recv, m, vpl[reps], vmat[reps], rep_w, 1, 0.01, 1e-3, ne,
                              nthreads=16).reshape(H, W, 3)
    err = rel_l2(img, ref)
    print(f"C2 full size: gather rel-L2 {err:.3e}", flush=True)
    assert err <= 1e-4 and ref.sum() > 0


@pytest.fixture(scope="module")
def instanced_1m(ctx):
    import renzoku_amd as rz
    from renzoku_amd import scenes
    import bench
    sc = scenes.instanced_meshes(1280, 720, seed=2013, target_tris=1_000_000)
    gs = rz.Scene(ctx, sc)
    guides = bench.guides_for(ctx, sc, gs, bench.BATCH_LEVELS)
    return sc, gs, guides


def _job_mix(guides, n):
    """the batched benchmark's seeded job mix (windows 32..256 px, 1-64 spp)"""
    rng = np.random.default_rng(2013)
    specs = []
    for i in range(n):
        w, h = int(rng.integers(32, 257)), int(rng.integers(32, 257))
        nv, vp, ls, g = guides[int(rng.integers(0, len(guides)))]
        specs.append(dict(guide=g, ls=ls, window=(int(rng.integers(0, 1280 - w + 1)), int(rng.integers(0, 720 - h + 1)),
                                                  w, h),
                          spp=int(rng.integers(1, 65)), max_bounce=5, direct_lighting=0, guide_mode=1,
                          seed=int(rng.integers(0, 2 ** 31))))
    return specs


def _guide_arrays(ls):
    S, res = ls.info().num_slices, 32
    return dict(res=res, maps=ls.get("maps", np.float32, S * res * res), row_cdf=ls.get("row_cdf", np.float32, S * res),
                col_cdf=ls.get("col_cdf", np.float32, S * res * res), totals=ls.get("map_totals", np.float32, S))


def test_c4_batched_full_mix(ctx, instanced_1m):
    import renzoku_amd as rz
    sc, gs, guides = instanced_1m
    specs = _job_mix(guides, 256)
    b = rz.Batch(ctx, gs)
    imgs, rays, ms = b.render(specs, count_rays=True)
    print(f"C4: 256 jobs in {ms:.1f} ms, rays {rays}", flush=True)
    for k in range(0, 256, 37):
        s = specs[k]
        ref, _ = s["ls"].render_guided(sc.camera, spp=s["spp"], max_bounce=5, direct_lighting=0, guide_mode=1,
                                       seed=s["seed"], window=s["window"], count_rays=False)
        x0, y0, w, h = s["window"]
        assert np.array_equal(imgs[k].view(np.uint32), ref[y0:y0 + h, x0:x0 + w].view(np.uint32)), k
    # one job of moderate size against the oracle's guided path tracer
    k = next(i for i, s in enumerate(specs)
             if 4096 <= s["window"][2] * s["window"][3] and 8 <= s["spp"] and
             s["window"][2] * s["window"][3] * s["spp"] <= 400_000)
    s = specs[k]
    osc = oracle.OrcScene(sc)
    ref = oracle.guided_pt(osc, sc.camera, s["spp"], 5, 0, 1, s["seed"], 1e-3, 16, window=s["window"], nthreads=16,
                           **_guide_arrays(s["ls"]))
    x0, y0, w, h = s["window"]
    err = rel_l2(imgs[k], ref[y0:y0 + h, x0:x0 + w])
    print(f"C4 job {k} {s['window']} x {s['spp']} spp against the oracle: rel-L2 {err:.3e}", flush=True)
    assert err <= 1e-6 and ref.sum() > 0


def _room_camera(rng, w, h):
    from renzoku_amd import scenes
    eye = (rng.uniform(200, 1800), rng.uniform(120, 480), rng.uniform(100, 900))
    look = (rng.uniform(200, 1800), rng.uniform(0, 300), rng.uniform(1000, 1900))
    return scenes.camera(eye, look, (0, 1, 0), float(rng.uniform(45, 75)), w, h)


def test_c5_served_full_scene(ctx):
    import renzoku_amd as rz
    from renzoku_amd import scenes, vplgen
    sc = scenes.instanced_meshes(1280, 720, seed=2013, target_tris=2_000_000)
    gs = rz.Scene(ctx, sc)
    vpl, vmat, ne = vplgen.device_vpls(gs, 200_000, seed=2013)
    vp = rz.Vpls(ctx, vpl, vmat)
    ls = rz.LightSlice(ctx, gs, vp, width=1280, height=720, tile=32, num_clusters=1000, num_neighbor_slices=16,
                       n_emitted=ne, seed=2013)
    ls.initialize(sc.camera)
    cl = ls.clusters()
    gparams = dict(tile=16, map_resolution=32, map_epsilon=1e-3, clamp_mode=1, incident_clamp=0.01, tmin=1e-3,
                   map_gate=0)
    print(f"C5: {sc.ntri} triangles, {len(vmat)} VPLs, {ls.info().num_clusters} resident clusters", flush=True)
    rng = np.random.default_rng(17)
    reqs = []
    for i in range(5):
        side = int(rng.integers(128, 385))
        reqs.append((_room_camera(rng, side, side), dict(spp=int(rng.integers(1, 7)), max_bounce=5,
                                                         direct_lighting=0, guide_mode=1,
                                                         seed=int(rng.integers(0, 2 ** 31)))))
    srv = rz.Server(0, gs, cl, memory_budget=4 << 30, guide=gparams)
    alone, _ = srv.wait(srv.submit(reqs[0][0], **reqs[0][1]), timeout=110.0)
    tickets = [srv.submit(c, **kw) for c, kw in reqs]
    mixed = [srv.wait(t, timeout=110.0)[0] for t in tickets]
    srv.close()
    assert np.array_equal(alone.view(np.uint32), mixed[0].view(np.uint32))
    for (cam, kw), img in zip(reqs, mixed):
        g = rz.Guide.create(ctx, gs, cl, cam, gparams)
        ref, _ = g.render(**kw)
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))
    # one small request against the oracle, with the guide tables of its camera
    cam = _room_camera(rng, 128, 128)
    kw = dict(spp=4, max_bounce=5, direct_lighting=0, guide_mode=1, seed=99)
    g = rz.Guide.create(ctx, gs, cl, cam, gparams)
    img, _ = g.render(**kw)
    S = math.ceil(128 / 16) ** 2
    tabs = dict(res=32, maps=g.get("maps", S * 32 * 32), row_cdf=g.get("row_cdf", S * 32),
                col_cdf=g.get("col_cdf", S * 32 * 32), totals=g.get("totals", S))
    osc = oracle.OrcScene(sc)
    ref = oracle.guided_pt(osc, cam, 4, 5, 0, 1, 99, 1e-3, 16, nthreads=16, **tabs)
    err = rel_l2(img, ref)
    print(f"C5 request 128x128 x 4 spp against the oracle: rel-L2 {err:.3e}", flush=True)
    assert err <= 1e-6 and ref.sum() > 0
