"""GPU parity: the HIP kernels (through the C ABI) against the CPU oracle on the
same seeded inputs.

Bars (fp32): hit triangle ids and t bit-equal (ray/box/triangle arithmetic is
restated without FMA and with correctly rounded div/sqrt); visibility,
traversal counts, BVH topology, alphas and cluster assignments bit-exact;
Lambertian-only radiance bit-exact up to summation order (asserted <= 1e-6
rel-L2); Phong/Ward radiance rel-L2 <= 1e-4 (glibc vs device powf/expf ULPs).
"""
import numpy as np
import pytest

import oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def random_rays(sc, n, seed):
    rng = np.random.default_rng(seed)
    lo = sc.positions.reshape(-1, 3).min(0)
    hi = sc.positions.reshape(-1, 3).max(0)
    o = (lo + (hi - lo) * rng.random((n, 3))).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    return np.stack([oracle.ray_make(o[i], d[i]) for i in range(n)])


@pytest.fixture(scope="module")
def scenes_gpu(ctx, cornell, small_room):
    import renzoku_amd as rz
    return {"cornell": (cornell, rz.Scene(ctx, cornell), oracle.OrcScene(cornell)),
            "room": (small_room, rz.Scene(ctx, small_room), oracle.OrcScene(small_room))}


@pytest.mark.parametrize("name", ["cornell", "room"])
def test_bvh_topology_matches_oracle(scenes_gpu, name):
    sc, gs, os_ = scenes_gpu[name]
    g = gs.export_bvh()
    o = os_.export_bvh()
    for a, b, what in zip(g, o, ["sorted_tri", "morton", "child", "range", "box"]):
        assert a.shape == b.shape, what
        assert np.array_equal(a.view(np.uint32) if a.dtype == np.float32 else a,
                              b.view(np.uint32) if b.dtype == np.float32 else b), what
    assert gs.info().depth == os_.depth()


@pytest.mark.parametrize("name", ["cornell", "room"])
def test_closest_hit_parity(scenes_gpu, name):
    sc, gs, os_ = scenes_gpu[name]
    w, h = int(sc.camera[17]), int(sc.camera[18])
    from renzoku_amd import scenes
    px, py = scenes.pixel_grid(w, h)
    rays = np.concatenate([oracle.camera_rays(sc.camera, px, py), random_rays(sc, 4096, 7)])
    t, b, g, tri = gs.trace_closest(rays)
    ot, ob, og, otri = os_.trace_closest(rays)
    assert np.array_equal(tri, otri), f"{(tri != otri).sum()} triangle id mismatches"
    hit = otri >= 0
    assert np.array_equal(t[hit].view(np.uint32), ot[hit].view(np.uint32))
    assert np.array_equal(b[hit].view(np.uint32), ob[hit].view(np.uint32))
    assert np.array_equal(g[hit].view(np.uint32), og[hit].view(np.uint32))
    # traversal work identical too
    st = gs.trace_stats(rays, 0)
    _, _, _, _, ost = os_.trace_closest(rays, stats=True)
    assert np.array_equal(st, ost)


@pytest.mark.parametrize("name", ["cornell", "room"])
def test_any_hit_and_segment_parity(scenes_gpu, name):
    sc, gs, os_ = scenes_gpu[name]
    rays = random_rays(sc, 8192, 11)
    for tmax in (1e6, 50.0):
        a = gs.trace_any(rays, 1e-3, tmax)
        oa = os_.trace_any(rays, 1e-3, tmax)
        assert np.array_equal(a, oa)
    st = gs.trace_stats(rays, 1)
    _, ost = os_.trace_any(rays, stats=True)
    assert np.array_equal(st, ost)
    rng = np.random.default_rng(3)
    lo, hi = sc.positions.reshape(-1, 3).min(0), sc.positions.reshape(-1, 3).max(0)
    p = (lo + (hi - lo) * rng.random((8192, 3))).astype(np.float32)
    q = (lo + (hi - lo) * rng.random((8192, 3))).astype(np.float32)
    assert np.array_equal(gs.trace_segments(p, q), os_.trace_segments(p, q))


@pytest.mark.parametrize("name", ["cornell", "room"])
def test_receiver_parity(scenes_gpu, name):
    sc, gs, os_ = scenes_gpu[name]
    from renzoku_amd import scenes
    w, h = int(sc.camera[17]), int(sc.camera[18])
    px, py = scenes.pixel_grid(w, h)
    rays, hits, recv, m, l = gs.camera_primary(sc.camera, px, py)
    orays = oracle.camera_rays(sc.camera, px, py)
    assert np.array_equal(rays.view(np.uint32), orays.view(np.uint32))
    t, b, g, tri = os_.trace_closest(orays)
    orecv, om, ol = os_.fill_receivers(orays, t, b, g, tri)
    assert np.array_equal(m, om) and np.array_equal(l, ol)
    assert np.array_equal(recv.view(np.uint32), orecv.view(np.uint32))


@pytest.fixture(scope="module")
def cornell_vpls(ctx, scenes_gpu):
    import renzoku_amd as rz
    from renzoku_amd import vplgen
    sc, gs, os_ = scenes_gpu["cornell"]
    vpl, vmat, ne = vplgen.device_vpls(gs, 1024, seed=2013)
    # the same VPL set from the oracle's restatement of the generator
    ovpl, ovmat, one = os_.generate_vpls(1024, seed=2013)
    assert np.array_equal(vpl.view(np.uint32), ovpl.view(np.uint32)) and ne == one
    return vpl, vmat, ne, rz.Vpls(ctx, vpl, vmat)


def test_gather_parity_cornell(ctx, scenes_gpu, cornell_vpls):
    import renzoku_amd as rz
    from renzoku_amd import scenes
    sc, gs, os_ = scenes_gpu["cornell"]
    vpl, vmat, ne, vp = cornell_vpls
    px, py = scenes.pixel_grid(256, 256)
    sel = np.arange(0, 256 * 256, 7)
    rays = oracle.camera_rays(sc.camera, px[sel], py[sel])
    t, b, g, tri = os_.trace_closest(rays)
    recv, m, _ = os_.fill_receivers(rays, t, b, g, tri)
    bits = rz.visibility(ctx, gs, vp, recv, m, 1e-3)
    # visibility bit-exact against the oracle's segment test
    k = np.arange(0, len(vmat), 37)
    for kk in k:
        blocked = os_.trace_segments(recv[:, :3], np.repeat(vpl[kk:kk + 1, :3], len(recv), 0))
        blocked = blocked | (m < 0)
        words = np.zeros(bits.shape[1], np.uint64)
        for i in np.nonzero(blocked)[0]:
            words[i // 64] |= np.uint64(1) << np.uint64(i % 64)
        tail = len(recv) % 64
        if tail:
            words[-1] |= ~((np.uint64(1) << np.uint64(tail)) - np.uint64(1))
        assert np.array_equal(bits[kk], words), f"VPL {kk}"
    out = rz.gather(ctx, gs, vp, recv, m, 1, 0.01, 1e-3, ne)
    ref = os_.gather(recv, m, vpl, vmat, 1, 0.01, 1e-3, ne, nthreads=8)
    assert rel_l2(out, ref) <= 1e-6


def test_render_vpl_parity_cornell(ctx, scenes_gpu, cornell_vpls):
    import renzoku_amd as rz
    from renzoku_amd import scenes
    sc, gs, os_ = scenes_gpu["cornell"]
    vpl, vmat, ne, vp = cornell_vpls
    img = rz.render_vpl(ctx, gs, vp, sc.camera, 1, 0.01, 1e-3, ne)
    px, py = scenes.pixel_grid(256, 256)
    rays = oracle.camera_rays(sc.camera, px, py)
    t, b, g, tri = os_.trace_closest(rays)
    recv, m, _ = os_.fill_receivers(rays, t, b, g, tri)
    ref = os_.gather(recv, m, vpl, vmat, 1, 0.01, 1e-3, ne, nthreads=8).reshape(256, 256, 3)
    assert rel_l2(img, ref) <= 1e-6
    assert img.mean() > 0


def test_render_vpl_ragged_image(ctx, scenes_gpu, cornell_vpls):
    """an image whose sides are not multiples of the 8x8 wave tiles (partial
    tiles in the primary-hit walk and the gather)"""
    import renzoku_amd as rz
    from renzoku_amd import scenes
    sc, gs, os_ = scenes_gpu["cornell"]
    vpl, vmat, ne, vp = cornell_vpls
    cam = np.array(sc.camera, np.float32).copy()
    W, H = 61, 45
    cam[17], cam[18] = W, H
    img = rz.render_vpl(ctx, gs, vp, cam, 1, 0.01, 1e-3, ne)
    px, py = scenes.pixel_grid(W, H)
    rays = oracle.camera_rays(cam, px, py)
    t, b, g, tri = os_.trace_closest(rays)
    recv, m, _ = os_.fill_receivers(rays, t, b, g, tri)
    ref = os_.gather(recv, m, vpl, vmat, 1, 0.01, 1e-3, ne, nthreads=8).reshape(H, W, 3)
    assert img.shape == (H, W, 3)
    assert rel_l2(img, ref) <= 1e-6


@pytest.fixture(scope="module")
def room_setup(ctx, scenes_gpu):
    import renzoku_amd as rz
    from renzoku_amd import vplgen
    sc, gs, os_ = scenes_gpu["room"]
    vpl, vmat, ne = vplgen.device_vpls(gs, 2048, seed=5)
    return sc, gs, os_, vpl, vmat, ne, rz.Vpls(ctx, vpl, vmat)


def test_reduced_matrix_alphas_assign_parity(ctx, room_setup):
    import renzoku_amd as rz
    sc, gs, os_, vpl, vmat, ne, vp = room_setup
    px, py = vplgen_rows(128, 128, 32)
    rays = oracle.camera_rays(sc.camera, px, py)
    t, b, g, tri = os_.trace_closest(rays)
    recv, m, _ = os_.fill_receivers(rays, t, b, g, tri)
    R, V = rz.reduced_matrix(ctx, gs, vp, recv, m, 1, 0.01, 1e-3)
    oR, oV = os_.reduced_matrix(recv, m, vpl, vmat, 1, 0.01, 1e-3, nthreads=8)
    assert np.array_equal(V, oV)
    assert rel_l2(R, oR) <= 1e-4
    # alphas / magnitudes / assignment bit-exact given the same R
    a, mags = rz.alphas(ctx, oR)
    assert np.array_equal(a.view(np.uint32), oracle.alphas(oR).view(np.uint32))
    assert np.array_equal(mags.view(np.uint32), oracle.column_magnitudes(oR).view(np.uint32))
    rng = np.random.default_rng(1)
    centers = rng.choice(len(vmat), 40, replace=False).astype(np.int32)
    w = rng.uniform(0.5, 50, 40).astype(np.float32)
    asg = rz.cluster_assign(ctx, oR, centers, w)
    assert np.array_equal(asg, oracle.cluster_assign(oR, centers, w, nthreads=8))


@pytest.mark.parametrize("cols,rows", [(1, 1), (3, 2), (1023, 3), (1024, 1), (2051, 5), (5000, 17)])
def test_alphas_ragged_shapes(ctx, cols, rows):
    """alphas and magnitudes bit-exact for sums whose length is not a multiple
    of the 1,024-value chunks (or of 4) of the ordered-sum kernel"""
    import renzoku_amd as rz
    rng = np.random.default_rng(cols * 31 + rows)
    R = rng.uniform(0, 3, (cols, rows, 3)).astype(np.float32) * (rng.uniform(size=(cols, rows, 1)) < 0.8)
    a, mags = rz.alphas(ctx, R)
    assert np.array_equal(a.view(np.uint32), oracle.alphas(R).view(np.uint32))
    assert np.array_equal(mags.view(np.uint32), oracle.column_magnitudes(R).view(np.uint32))


def vplgen_rows(w, h, tile):
    from renzoku_amd import vplgen
    return vplgen.tile_centre_pixels(w, h, tile)


def test_lightslice_pipeline_parity(ctx, room_setup):
    import renzoku_amd as rz
    from renzoku_amd import scenes, vplgen
    sc, gs, os_, vpl, vmat, ne, vp = room_setup
    W = H = 128
    ls = rz.LightSlice(ctx, gs, vp, width=W, height=H, tile=16, num_clusters=64, num_neighbor_slices=4,
                       n_emitted=ne, seed=2013)
    ls.initialize(sc.camera)
    img = ls.gather()
    inf = ls.info()
    S, V = inf.num_slices, len(vmat)
    R = ls.get("R", np.float32, S * V * 3).reshape(V, S, 3)
    # clustering on the device's R, replayed by the oracle on the same MT stream
    oc = oracle.cluster_by_sampling(R, 64, 2013, nthreads=8)
    assert np.array_equal(ls.get("centers", np.int32, inf.num_centers), oc["centers"])
    assert np.array_equal(ls.get("assign", np.int32, V), oc["assign"])
    reps = ls.get("reps", np.int32, inf.num_clusters)
    rep_w = ls.get("rep_w", np.float32, inf.num_clusters)
    assert np.array_equal(reps, oc["reps"]) and np.array_equal(rep_w.view(np.uint32), oc["rep_w"].view(np.uint32))
    # rows and reduced matrix against the oracle
    px, py = vplgen.tile_centre_pixels(W, H, 16)
    rays = oracle.camera_rays(sc.camera, px, py)
    t, b, g, tri = os_.trace_closest(rays)
    rrecv, rm, _ = os_.fill_receivers(rays, t, b, g, tri)
    oR, oV = os_.reduced_matrix(rrecv, rm, vpl, vmat, 1, 0.01, 1e-3, nthreads=8)
    assert rel_l2(R, oR) <= 1e-4
    # final gather: Σ_c w_c gather_point(rep_c) / n_emitted per pixel
    pxa, pya = scenes.pixel_grid(W, H)
    prays = oracle.camera_rays(sc.camera, pxa, pya)
    t, b, g, tri = os_.trace_closest(prays)
    # the primary hits (one wave per 8x8 tile walking the reference tree) are the oracle's, bit for bit
    hits = ls.get("hits", np.float32, W * H * 4).reshape(-1, 4)
    assert np.array_equal(hits[:, 3].view(np.int32), tri)
    ok = tri >= 0
    for k, ref in enumerate((t, b, g)):
        assert np.array_equal(hits[ok, k].view(np.uint32), ref[ok].view(np.uint32))
    recv, m, _ = os_.fill_receivers(prays, t, b, g, tri)
    ref = os_.gather_clusters(recv, m, vpl[reps], vmat[reps], rep_w, 1, 0.01, 1e-3, ne, nthreads=8).reshape(H, W, 3)
    assert rel_l2(img, ref) <= 1e-4
    # per-slice guiding data against the oracle's restatement
    nb = ls.get("neighbors", np.int32, S * 4).reshape(S, 4)
    assert np.array_equal(nb, vplgen.tile_neighbours(W, H, 16, 4))
    upper = ls.get("slice_upper", np.int32, S * len(reps)).reshape(S, -1)
    Vd = ls.get("V", np.float32, S * V).reshape(V, S)
    nr = len(reps)
    gpdf = ls.get("slice_pdf", np.float32, S * nr).reshape(S, -1)
    gcdf = ls.get("slice_cdf", np.float32, S * nr).reshape(S, -1)
    swi = ls.get("slice_wi", np.float32, S * nr * 3).reshape(S, nr, 3)
    res = 32
    gmaps = ls.get("maps", np.float32, S * res * res).reshape(S, res, res)
    growc = ls.get("row_cdf", np.float32, S * res).reshape(S, res)
    gcolc = ls.get("col_cdf", np.float32, S * res * res).reshape(S, res, res)
    gtot = ls.get("map_totals", np.float32, S)
    rows_recv = ls.get("rows_recv", np.float32, S * 21).reshape(S, 21)
    for s in range(0, S, 5):
        wi, up = os_.slice_cluster_info(rrecv[s], reps, vpl, vmat, Vd[:, s].copy(), 1, 0.01)
        assert np.array_equal(up, upper[s])
        w = oracle.slice_cluster_pdf(R, nb[s], reps, rep_w, up)
        ok, pdf, cdf = oracle.pdf_build(w)
        assert np.array_equal(gpdf[s].view(np.uint32), pdf.view(np.uint32))
        assert np.array_equal(gcdf[s].view(np.uint32), cdf.view(np.uint32))
        # the slice's luminance map and its 2-D distribution (build_map, lightslice.cc:2058-2151)
        om = oracle.build_map(res, swi[s], upper[s], rep_w, R[reps, s, :], rows_recv[s, 12:21], 1e-3, 1)
        assert np.array_equal(gmaps[s].view(np.uint32), om["map"].view(np.uint32)), f"slice {s} map"
        assert np.array_equal(growc[s].view(np.uint32), om["row_cdf"].view(np.uint32))
        assert np.array_equal(gcolc[s].view(np.uint32), om["col_cdf"].view(np.uint32))
        assert gtot[s] == np.float32(om["total"])


def test_render_vpl_chunked_gather_is_one_launch(ctx, scenes_gpu):
    """a 1024x1024 render over 2,560 VPLs gathers in two VPL chunks (its bit
    matrix would pass 256 MiB in one); the chunked running sums give the bits of
    one gather launch over all VPLs on the same receivers"""
    import renzoku_amd as rz
    from renzoku_amd import scenes, vplgen
    sc, gs, os_ = scenes_gpu["cornell"]
    vpl, vmat, ne = vplgen.device_vpls(gs, 2560, seed=8)
    vp = rz.Vpls(ctx, vpl, vmat)
    cam = np.array(scenes.cornell_box(1024, 1024).camera, np.float32)   # same box, 1024x1024 camera
    img = rz.render_vpl(ctx, gs, vp, cam, 1, 0.01, 1e-3, ne)
    px, py = scenes.pixel_grid(1024, 1024)
    _, _, recv, m, _ = gs.camera_primary(cam, px, py)
    one = rz.gather(ctx, gs, vp, recv, m, 1, 0.01, 1e-3, ne).reshape(1024, 1024, 3)
    assert np.array_equal(img.view(np.uint32), one.view(np.uint32))
    assert img.mean() > 0
