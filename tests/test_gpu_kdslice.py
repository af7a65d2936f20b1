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
