"""Device 6-D kd-tree slicing (rzk_kdslices_*, rzk_map_vec6) against the oracle:
order, slices, axes, planes and slice tables bit-exact (boxes compared as
floats: a box's min / max of +0 and -0 may keep either sign), slice lookups and
nearest-slice lists identical."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import renzoku_amd as rz
    return rz.Ctx(0)


def _points(n, seed, ties):
    rng = np.random.default_rng(seed)
    P = rng.random((n, 6)).astype(np.float32)
    if ties:
        P = (np.round(P * 9) / 9).astype(np.float32)
    return P


def _same_tree(g, o):
    assert g.S == o.S and g.heap_used == o.heap_used
    assert np.array_equal(g.get("order"), o.order)
    assert np.array_equal(g.get("slices"), o.slices)
    assert np.array_equal(g.get("boxes"), o.boxes)
    assert np.array_equal(g.get("cl"), o.cl)
    assert np.array_equal(g.get("axis"), o.axis)
    assert np.array_equal(g.get("plane").view(np.uint32), o.plane.view(np.uint32))


# 300k points run the multi-block rounds at the top depths (nodes above 4096 points)
@pytest.mark.parametrize("n,max_size,ties,seed", [(3000, 256, False, 1), (50_000, 256, True, 2),
                                                  (300_000, 256, False, 3), (70_001, 100, True, 4),
                                                  (1000, 3, True, 5), (2, 3, False, 6), (200, 256, False, 7)])
def test_kd_build_matches_oracle(ctx, n, max_size, ties, seed):
    import renzoku_amd as rz
    print(f"kd build n={n} max={max_size} ties={ties} seed={seed}", flush=True)
    P = _points(n, seed, ties)
    g = rz.KdSlices(ctx, P, max_size, seed)
    o = oracle.KdSlices(P, max_size, seed)
    _same_tree(g, o)
    q = np.concatenate([P[:: max(1, n // 2000)], _points(500, seed + 100, ties)])
    assert np.array_equal(g.find(q), o.find(q))
    for k in (1, 16, 32):
        assert np.array_equal(g.nearest(q[:600], k), o.nearest(q[:600], k))


# the reference's pivot order: one MT stream drawn depth-first (host build)
@pytest.mark.parametrize("n,max_size,ties,seed", [(3000, 256, False, 2013), (50_000, 256, True, 2),
                                                  (250_000, 256, False, 3), (1000, 3, True, 5), (2, 3, False, 6),
                                                  (200, 256, False, 7)])
def test_kd_build_mt_matches_oracle(ctx, n, max_size, ties, seed):
    import renzoku_amd as rz
    print(f"kd build (MT pivots) n={n} max={max_size} ties={ties} seed={seed}", flush=True)
    P = _points(n, seed, ties)
    g = rz.KdSlices(ctx, P, max_size, seed, pivots="mt")
    o = oracle.KdSlices(P, max_size, seed, pivots="mt")
    _same_tree(g, o)
    assert g.draws == o.draws and (n < 2 or g.draws >= g.S)
    q = np.concatenate([P[:: max(1, n // 2000)], _points(500, seed + 100, ties)])
    assert np.array_equal(g.find(q), o.find(q))
    for k in (1, 16):
        assert np.array_equal(g.nearest(q[:600], k), o.nearest(q[:600], k))
    g.close()


def test_kd_build_mt_rejects_non_finite_points(ctx):
    import renzoku_amd as rz
    P = _points(100, 1, False)
    P[17, 3] = np.nan
    with pytest.raises(rz.RzkError):
        rz.KdSlices(ctx, P, 16, 1, pivots="mt")


def test_kd_empty_and_bad_input(ctx):
    import renzoku_amd as rz
    for n in (0, 1):
        g = rz.KdSlices(ctx, np.zeros((n, 6), np.float32), 8, 1)
        assert g.S == 0 and g.heap_used == 0
        assert (g.find(np.zeros((3, 6), np.float32)) == -1).all()
        assert (g.nearest(np.zeros((3, 6), np.float32), 4) == -1).all()
    with pytest.raises(rz.RzkError):
        rz.KdSlices(ctx, _points(10, 1, False), 2, 1)
    g = rz.KdSlices(ctx, _points(100, 1, False), 16, 1)
    with pytest.raises(rz.RzkError):
        g.nearest(np.zeros((1, 6), np.float32), 33)
    with pytest.raises(rz.RzkError):
        g.nearest(np.zeros((1, 6), np.float32), 0)


def test_map_vec6_matches_oracle(ctx):
    import renzoku_amd as rz
    from renzoku_amd import scenes
    sc = scenes.room(64, 64, seed=2013, target_tris=20_000)
    # two of the materials become Ward (gloss clamped at 4) and a sharp Phong (2.5)
    mats_ = np.array(sc.materials, np.float32).copy()
    mats_[0, 0], mats_[0, 7], mats_[0, 8] = 2, 1e-4, 0.3
    mats_[1, 0], mats_[1, 7] = 1, 2500
    sc.materials = mats_
    scene = rz.Scene(ctx, sc)
    osc = oracle.OrcScene(sc)
    px, py = np.meshgrid(np.arange(64, dtype=np.int32), np.arange(64, dtype=np.int32))
    rays = oracle.camera_rays(sc.camera, px.ravel(), py.ravel())
    t, b, g, tri = osc.trace_closest(rays)
    recv, mats, _ = osc.fill_receivers(rays, t, b, g, tri)
    assert (mats >= 0).sum() > 1000 and len(set(sc.materials[mats[mats >= 0], 0].tolist())) == 3
    diag = 17.5
    got = rz.map_vec6(ctx, scene, recv, mats, diag)
    want = oracle.map_vec6(recv, mats, osc.materials, diag)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_kd_build_from_device_points(ctx):
    import ctypes as C
    import renzoku_amd as rz
    P = _points(20_000, 9, True)
    d = C.c_void_p()
    ctx.check(ctx.lib.rzk_device_alloc(ctx.h, P.nbytes, C.byref(d)), "alloc")
    try:
        ctx.check(ctx.lib.rzk_copy_h2d(ctx.h, d, P.ctypes.data_as(C.c_void_p), P.nbytes), "h2d")
        g = rz.KdSlices(ctx, d.value, 128, 9, n=len(P))
        _same_tree(g, oracle.KdSlices(P, 128, 9))
        g.close()
    finally:
        ctx.lib.rzk_device_free(ctx.h, d)


def _rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("pivots", ["philox", "mt"])
def test_lightslice_with_kd_slices_matches_oracle(ctx, pivots):
    """LightSlice with the reference's slices: gather points, kd-tree, MT
    representatives and rows, nearest-slice neighbours, each pixel's slice,
    then clustering and per-slice refinement on the continuing MT stream and
    the gather over each pixel's own slice.  pivots="mt": the tree's pivots are
    the first numbers of the same MT stream, drawn in the reference's recursion
    order, and everything after them moves up the stream by their count."""
    import renzoku_amd as rz
    from renzoku_amd import scenes, vplgen
    W = H = 128
    sc = scenes.room(W, H, seed=11, target_tris=20_000)
    gs = rz.Scene(ctx, sc)
    vpl, vmat, ne = vplgen.device_vpls(gs, 2048, seed=5)
    vp = rz.Vpls(ctx, vpl, vmat)
    ls = rz.LightSlice(ctx, gs, vp, width=W, height=H, num_clusters=32, num_neighbor_slices=4, n_emitted=ne,
                       seed=2013, max_slice_clusters=64, kd_max_slice_size=128, kd_pivots=int(pivots == "mt"))
    ls.initialize(sc.camera)
    inf = ls.info()
    S, V = inf.num_slices, len(vmat)
    # oracle: receivers, gather points, points, tree
    osc = oracle.OrcScene(sc)
    px, py = scenes.pixel_grid(W, H)
    rays = oracle.camera_rays(sc.camera, px, py)
    t, b, g, tri = osc.trace_closest(rays)
    recv, m, _ = osc.fill_receivers(rays, t, b, g, tri)
    gp = np.nonzero(m >= 0)[0].astype(np.int32)
    assert np.array_equal(ls.get("gather_pix", np.int32, len(gp)), gp)
    p = np.asarray(sc.positions, np.float32).reshape(-1, 3)
    d = p.max(0) - p.min(0)
    diag = float(np.sqrt(np.float32(np.float32(d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]), dtype=np.float32))
    pts = oracle.map_vec6(recv[gp], m[gp], osc.materials, diag)
    okd = oracle.KdSlices(pts, 128, 2013, pivots=pivots)
    assert S == okd.S and S >= 64
    reps, ndraw = oracle.kd_representatives(okd.slices, len(gp), 2013, skip=okd.draws)
    assert np.array_equal(ls.get("row_pix", np.int32, S), gp[okd.order[reps]])
    nb = ls.get("neighbors", np.int32, S * 4).reshape(S, 4)
    assert np.array_equal(nb, okd.nearest(pts[okd.order[reps]], 4))
    psl = np.full(W * H, -1, np.int32)
    psl[gp] = okd.find(pts)
    assert np.array_equal(ls.get("pixel_slice", np.int32, W * H), psl)
    # clustering and refinement continue the MT stream after the representatives
    R = ls.get("R", np.float32, S * V * 3).reshape(V, S, 3)
    o = oracle.cluster_and_refine(R, 32, 64, nb, 2013, skip=ndraw)
    cnt = ls.get("slice_count", np.int32, S)
    sreps = ls.get("slice_reps", np.int32, S * 64).reshape(S, 64)
    rw = ls.get("slice_rep_w", np.float32, S * 64).reshape(S, 64)
    assert np.array_equal(ls.get("centers", np.int32, inf.num_centers), o["centers"])
    assert np.array_equal(cnt, o["count"])
    for s in range(S):
        assert np.array_equal(sreps[s, :cnt[s]], o["rep"][s, :cnt[s]])
        assert np.array_equal(rw[s, :cnt[s]].view(np.uint32), o["w"][s, :cnt[s]].view(np.uint32))
    # the gather: each pixel over its own slice's representatives
    img = ls.gather()
    ref = np.zeros((W * H, 3), np.float32)
    for s in range(S):
        sel = np.nonzero(psl == s)[0]
        n = cnt[s]
        ref[sel] = osc.gather_clusters(recv[sel], m[sel], vpl[sreps[s, :n]], vmat[sreps[s, :n]], rw[s, :n], 1, 0.01,
                                       1e-3, ne, nthreads=8)
    assert _rel_l2(img.reshape(-1, 3), ref) <= 1e-4
    work = ls.count_work()
    assert work["gather_pairs"] == int(cnt[psl[gp]].sum())
    # guided rendering reads each pixel's slice
    out, rays_ = ls.render_guided(sc.camera, spp=4, max_bounce=2, guide_mode=1, seed=3)
    assert out.shape == (H, W, 3) and np.isfinite(out).all() and out.sum() > 0
    with pytest.raises(rz.RzkError):
        ls.guide(sc.camera)


def test_lightslice_kd_slices_without_refinement(ctx):
    """every slice gathers the kept clusters: the image equals the shared-list
    gather, and the slice table, rows and neighbours still follow the tree"""
    import renzoku_amd as rz
    from renzoku_amd import scenes, vplgen
    W, H = 96, 80
    sc = scenes.cornell_box(W, H)
    gs = rz.Scene(ctx, sc)
    vpl, vmat, ne = vplgen.device_vpls(gs, 1024, seed=3)
    vp = rz.Vpls(ctx, vpl, vmat)
    ls = rz.LightSlice(ctx, gs, vp, width=W, height=H, num_clusters=24, num_neighbor_slices=8, n_emitted=ne,
                       seed=99, kd_max_slice_size=64)
    ls.initialize(sc.camera)
    inf = ls.info()
    S = inf.num_slices
    osc = oracle.OrcScene(sc)
    px, py = scenes.pixel_grid(W, H)
    rays = oracle.camera_rays(sc.camera, px, py)
    t, b, g, tri = osc.trace_closest(rays)
    recv, m, _ = osc.fill_receivers(rays, t, b, g, tri)
    gp = np.nonzero(m >= 0)[0].astype(np.int32)
    p = np.asarray(sc.positions, np.float32).reshape(-1, 3)
    d = p.max(0) - p.min(0)
    diag = float(np.sqrt(np.float32(np.float32(d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]), dtype=np.float32))
    pts = oracle.map_vec6(recv[gp], m[gp], osc.materials, diag)
    okd = oracle.KdSlices(pts, 64, 99)
    assert S == okd.S
    reps, ndraw = oracle.kd_representatives(okd.slices, len(gp), 99)
    assert np.array_equal(ls.get("row_pix", np.int32, S), gp[okd.order[reps]])
    assert np.array_equal(ls.get("neighbors", np.int32, S * 8).reshape(S, 8), okd.nearest(pts[okd.order[reps]], 8))
    V = len(vmat)
    R = ls.get("R", np.float32, S * V * 3).reshape(V, S, 3)
    oc = oracle.cluster_by_sampling(R, 24, 99, nthreads=8, skip=ndraw)
    creps = ls.get("reps", np.int32, inf.num_clusters)
    crw = ls.get("rep_w", np.float32, inf.num_clusters)
    assert np.array_equal(creps, oc["reps"]) and np.array_equal(crw.view(np.uint32), oc["rep_w"].view(np.uint32))
    img = ls.gather()
    ref = osc.gather_clusters(recv, m, vpl[creps], vmat[creps], crw, 1, 0.01, 1e-3, ne, nthreads=8).reshape(H, W, 3)
    assert _rel_l2(img, ref) <= 1e-4
    with pytest.raises(rz.RzkError):
        rz.LightSlice(ctx, gs, vp, width=W, height=H, num_clusters=24, kd_max_slice_size=2)
    with pytest.raises(rz.RzkError):
        rz.LightSlice(ctx, gs, vp, width=W, height=H, num_clusters=24, kd_max_slice_size=64, kd_pivots=2)
