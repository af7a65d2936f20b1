"""The scene's trees built on the device (rzk_scene_create_ex with
RZK_SCENE_BUILD_DEVICE) against the host build, which the oracle pins
(test_gpu_parity.test_bvh_topology_matches_oracle).

Bars: the binary Morton tree (sorted order, codes, node records with their
child boxes and refs, triangle records, depth, scene box) bit-equal to the host
build's; the 4-wide tree a valid collapse of it (every triangle once, every
slot box the exact union of what lies below, preorder layout); every walk on
the binary tree (closest hits, their visit counts, the receivers built from
their triangles' frames) and every
any-hit, segment and visibility answer bit-equal between the two builds, so
the VPL render's image is too.  Scenes: the Cornell box, the room, a soup whose
second half repeats the first (equal Morton codes: midpoint splits, exact
closest-hit ties), a soup in the plane z = 0 (zero scene extent on one axis),
two triangles.
"""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def soup(base, n, seed, planar=False, repeat=False):
    rng = np.random.default_rng(seed)
    c = (rng.random((n, 3)) * 4 - 2).astype(np.float32)
    tri = c[:, None, :] + ((rng.random((n, 3, 3)) - 0.5) * 0.2).astype(np.float32)
    if planar:
        tri[:, :, 2] = 0.0
    if repeat:
        tri[n // 2:] = tri[:n - n // 2]
    return dataclasses.replace(base, positions=np.ascontiguousarray(tri.reshape(n, 9), np.float32),
                               tangents=np.tile(np.float32([1, 0, 0]), (n, 1)),
                               tri_surface=np.zeros(n, np.int32), light_tris=None)


@pytest.fixture(scope="module")
def built(ctx, cornell, small_room):
    import renzoku_amd as rz
    scs = {"cornell": cornell, "room": small_room,
           "soup_repeat": soup(cornell, 6000, 5, repeat=True),
           "soup_planar": soup(cornell, 3000, 6, planar=True),
           "two": soup(cornell, 2, 7)}
    out = {}
    for name, sc in scs.items():
        print(f"[bvh_device] building {name}: {sc.ntri} triangles", flush=True)
        out[name] = (sc, rz.Scene(ctx, sc), rz.Scene(ctx, sc, build="device"))
    return out


NAMES = ["cornell", "room", "soup_repeat", "soup_planar", "two"]


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.parametrize("name", NAMES)
def test_binary_tree_bit_equal(built, name):
    sc, hs, ds = built[name]
    hi, di = hs.info(), ds.info()
    assert (hi.num_triangles, hi.num_interior_nodes, hi.depth) == (di.num_triangles, di.num_interior_nodes, di.depth)
    assert np.array_equal(bits(np.array(hi.box[:], np.float32)), bits(np.array(di.box[:], np.float32)))
    hst, hsm = hs.export_bvh()[:2]
    dst, dsm = ds.export_bvh_sorted()
    assert np.array_equal(hst, dst) and np.array_equal(hsm, dsm)
    hn, ht, _, _ = hs.export_trees()
    dn, dt, _, _ = ds.export_trees()
    assert np.array_equal(bits(ht), bits(dt)), "triangle records"
    assert np.array_equal(bits(hn), bits(dn)), f"{(bits(hn) != bits(dn)).any(1).sum()} node records differ"


def tri_boxes(sc):
    p = sc.positions.reshape(-1, 3, 3)
    return p.min(1), p.max(1)


@pytest.mark.parametrize("name", NAMES)
def test_four_wide_tree_is_a_collapse(built, name):
    sc, hs, ds = built[name]
    _, _, n4, t4 = ds.export_trees()
    refs = n4[:, 24:28].view(np.int32)
    lo = np.stack([n4[:, 0:4], n4[:, 4:8], n4[:, 8:12]], -1)     # [node, slot, axis]
    hi = np.stack([n4[:, 12:16], n4[:, 16:20], n4[:, 20:24]], -1)
    tlo, thi = tri_boxes(sc)
    orig = t4[:, 3].view(np.int32)
    assert np.array_equal(np.sort(orig), np.arange(sc.ntri))
    empty = np.int32(-2 ** 31)
    leaf_seen = np.zeros(sc.ntri, np.int64)
    parent_seen = np.zeros(len(n4), np.int64)
    for j in range(len(n4)):
        r = refs[j]
        used = r != empty
        k = int(used.sum())
        assert k >= 2 and used[:k].all() and not used[k:].any(), f"node {j} slots {r}"
        first_inner = True
        for s in range(k):
            if r[s] < 0:
                i = ~r[s]
                leaf_seen[i] += 1
                assert np.array_equal(bits(lo[j, s]), bits(tlo[orig[i]])) and \
                    np.array_equal(bits(hi[j, s]), bits(thi[orig[i]])), f"leaf box {j}/{s}"
            else:
                c = r[s]
                assert c > j and parent_seen[c] == 0
                if first_inner:
                    assert c == j + 1, "preorder: the first interior child follows its parent"
                    first_inner = False
                parent_seen[c] += 1
                cu = refs[c] != empty
                assert np.array_equal(bits(lo[j, s]), bits(lo[c, cu].min(0))), f"box {j}/{s}"
                assert np.array_equal(bits(hi[j, s]), bits(hi[c, cu].max(0))), f"box {j}/{s}"
    assert (leaf_seen == 1).all()
    assert parent_seen[0] == 0 and (parent_seen[1:] == 1).all()


def random_rays(sc, n, seed):
    import oracle
    rng = np.random.default_rng(seed)
    lo = sc.positions.reshape(-1, 3).min(0) - 0.1
    hi = sc.positions.reshape(-1, 3).max(0) + 0.1
    o = (lo + (hi - lo) * rng.random((n, 3))).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    return np.stack([oracle.ray_make(o[i], d[i]) for i in range(n)])


@pytest.mark.parametrize("name", NAMES)
def test_walks_bit_equal(built, name):
    sc, hs, ds = built[name]
    rays = random_rays(sc, 8192, 21)
    hits = hs.trace_closest(rays)
    for a, b in zip(hits, ds.trace_closest(rays)):
        assert np.array_equal(bits(a), bits(b))
    # receivers read the per-triangle frames (normal, surface, tangent) the device build computes itself
    for a, b in zip(hs.fill_receivers(rays, *hits), ds.fill_receivers(rays, *hits)):
        assert np.array_equal(bits(a), bits(b))
    for kind in (0, 1):
        assert np.array_equal(hs.trace_stats(rays, kind), ds.trace_stats(rays, kind))
    for tmax in (1e6, 0.5):
        assert np.array_equal(hs.trace_any(rays, 1e-3, tmax), ds.trace_any(rays, 1e-3, tmax))
    rng = np.random.default_rng(22)
    lo, hi = sc.positions.reshape(-1, 3).min(0) - 0.5, sc.positions.reshape(-1, 3).max(0) + 0.5
    p = (lo + (hi - lo) * rng.random((8192, 3))).astype(np.float32)
    q = (lo + (hi - lo) * rng.random((8192, 3))).astype(np.float32)
    blocked = hs.trace_segments(p, q)
    assert np.array_equal(blocked, ds.trace_segments(p, q))
    if name != "two":
        assert 0 < blocked.sum() < len(blocked)


def test_render_vpl_bit_equal(ctx, built):
    import renzoku_amd as rz
    from renzoku_amd import vplgen
    sc, hs, ds = built["cornell"]
    vpl, vmat, ne = vplgen.device_vpls(hs, 1024, seed=2013)
    dvpl, dvmat, dne = vplgen.device_vpls(ds, 1024, seed=2013)
    assert np.array_equal(bits(vpl), bits(dvpl)) and ne == dne
    vp = rz.Vpls(ctx, vpl, vmat)
    a = rz.render_vpl(ctx, hs, vp, sc.camera, 1, 0.01, 1e-3, ne)
    b = rz.render_vpl(ctx, ds, vp, sc.camera, 1, 0.01, 1e-3, ne)
    assert np.array_equal(bits(a), bits(b)) and a.mean() > 0


def test_host_node_export_refused_for_device_build(built):
    sc, hs, ds = built["cornell"]
    import ctypes as C
    import renzoku_amd._lib as L
    n = C.c_int32()
    ch = np.zeros(2 * (2 * sc.ntri - 1), np.int32)
    rc = ds.lib.rzk_scene_export_bvh(ds.h, None, None, L.ptr(ch), None, None, C.byref(n))
    assert rc == 4  # RZK_E_STATE: the reference-order node view is kept by the host build only
