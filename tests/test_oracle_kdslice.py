"""CPU checks of the oracle's 6-D kd-tree slicing (oracle/kdslice.cpp).

What pins it:
  * the tree: node ranges split at (s + e) / 2, the plane = the largest key of
    the left child and <= every key of the right one, the longest axis of the
    node's box, slices = first nodes of at most max_slice_size points, boxes =
    min / max of their points, and the order = a permutation;
  * the selection's moves: a second formulation (depth by depth, in-place
    stable partitions of each node, numpy) gives the same order, planes and
    axes as the oracle's ping-pong recursion, with many equal keys;
  * the reference's pivot order (pivots from one MT stream, kdtree_n.h:192): a
    depth-first numpy build drawing from the MT stream (itself pinned on the
    reference's mtwist golden vectors) takes the same numbers in the same order
    and gives the same order, planes and axes;
  * find: every point farther than the plane bias from the planes on its path
    lands in its own slice;
  * nearest: on integer data with queries a quarter off the grid, where the
    plane bias cannot change a pruning decision, the k distances equal the k
    smallest box distances by brute force, nearest last;
  * map_vec6: closed forms (Lambertian: the normal; Phong / Ward: the mirror
    direction of wo, whose angle to n equals wo's, scaled by the clamped gloss).
"""
import numpy as np
import pytest

import oracle


def _philox(c, k0, k1):
    c = list(c)
    for _ in range(10):
        p0, p1 = 0xD2511F53 * c[0], 0xCD9E8D57 * c[2]
        c = [(p1 >> 32) ^ c[1] ^ k0, p1 & 0xffffffff, (p0 >> 32) ^ c[3] ^ k1, p0 & 0xffffffff]
        k0, k1 = (k0 + 0x9E3779B9) & 0xffffffff, (k1 + 0xBB67AE85) & 0xffffffff
    return c


def _pivot(seed, s, e, t, lo, hi):
    c = _philox([s, e, t, 0x6b64], seed ^ 0x2545f491, 0x6b647472)
    u = np.float32(c[0] & 0xffffff) * np.float32(1.0 / 16777216.0)
    return lo + int(np.float32(u * np.float32(hi - lo + 1)))


def _node_range(n, d, j):
    s, e = 0, n - 1
    for b in range(d - 1, -1, -1):
        if s >= e:
            return None
        m = (s + e) >> 1
        s, e = (m + 1, e) if (j >> b) & 1 else (s, m)
    return s, e


def _level_build(P, max_size, seed):
    """depth by depth, each node's selection as in-place stable partitions"""
    n = len(P)
    order = np.arange(n)
    axes, planes = {}, {}
    m, d = n, 0
    while m >= 2:
        for j in range(1 << d):
            r = _node_range(n, d, j)
            if r is None or r[0] >= r[1]:
                continue
            s, e = r
            ids = order[s:e + 1].copy()
            pts = P[ids]
            ext = np.abs(pts.max(0) - pts.min(0))
            ax, best = 0, np.float32(0)
            for a in range(6):
                if ext[a] > best:
                    ax, best = a, ext[a]
            key = pts[:, ax].copy()
            k, lo, hi, t = (e - s) >> 1, 0, e - s, 0
            while True:
                v = key[_pivot(seed, s, e, t, lo, hi)]
                kk, ii = key[lo:hi + 1], ids[lo:hi + 1]
                L, E, G = kk < v, kk == v, kk > v
                key[lo:hi + 1] = np.concatenate([kk[L], kk[E], kk[G]])
                ids[lo:hi + 1] = np.concatenate([ii[L], ii[E], ii[G]])
                LT, EQ = int(L.sum()), int(E.sum())
                if k < lo + LT:
                    hi = lo + LT - 1
                elif k < lo + LT + EQ:
                    break
                else:
                    lo += LT + EQ
                t += 1
            heap = (1 << d) - 1 + j
            axes[heap], planes[heap] = ax, key[k]
            order[s:e + 1] = ids
        m, d = (m + 1) // 2, d + 1
    return order, axes, planes


def _points(n, seed, ties):
    rng = np.random.default_rng(seed)
    P = rng.random((n, 6)).astype(np.float32)
    if ties:
        P = (np.round(P * 5) / 5).astype(np.float32)
    return P


@pytest.mark.parametrize("n,max_size,ties", [(3000, 256, False), (6000, 100, True), (777, 3, True), (2, 3, False),
                                             (300, 256, True)])
def test_selection_moves_match_a_depth_by_depth_build(n, max_size, ties):
    P = _points(n, n, ties)
    kd = oracle.KdSlices(P, max_size, 99)
    order, axes, planes = _level_build(P, max_size, 99)
    assert np.array_equal(order, kd.order)
    for h in range(kd.heap_used):
        if h in axes:
            assert kd.axis[h] == axes[h] and kd.plane[h].view(np.uint32) == np.float32(planes[h]).view(np.uint32)
        else:
            assert kd.axis[h] == 127


def _dfs_build_mt(P, seed):
    """depth-first, pivots taken in order from the MT stream seeded with seed;
    returns order, axes, planes and the numbers taken"""
    n = len(P)
    stream = (oracle.mt_stream(seed, 40 * n + 64) & 0xffffff).astype(np.float32) * np.float32(1.0 / 16777216.0)
    order = np.arange(n)
    axes, planes = {}, {}
    taken = 0

    def node(heap, s, e):
        nonlocal taken
        if s >= e:
            return
        ids = order[s:e + 1].copy()
        pts = P[ids]
        ext = np.abs(pts.max(0) - pts.min(0))
        ax, best = 0, np.float32(0)
        for a in range(6):
            if ext[a] > best:
                ax, best = a, ext[a]
        key = pts[:, ax].copy()
        k, lo, hi = (e - s) >> 1, 0, e - s
        while True:
            u = stream[taken]
            taken += 1
            v = key[lo + int(np.float32(u * np.float32(hi - lo + 1)))]
            kk, ii = key[lo:hi + 1], ids[lo:hi + 1]
            L, E, G = kk < v, kk == v, kk > v
            key[lo:hi + 1] = np.concatenate([kk[L], kk[E], kk[G]])
            ids[lo:hi + 1] = np.concatenate([ii[L], ii[E], ii[G]])
            LT, EQ = int(L.sum()), int(E.sum())
            if k < lo + LT:
                hi = lo + LT - 1
            elif k < lo + LT + EQ:
                break
            else:
                lo += LT + EQ
        axes[heap], planes[heap] = ax, key[k]
        order[s:e + 1] = ids
        m = (s + e) // 2
        node(2 * heap + 1, s, m)
        node(2 * heap + 2, m + 1, e)

    node(0, 0, n - 1)
    assert taken < len(stream)
    return order, axes, planes, taken


@pytest.mark.parametrize("n,max_size,ties,seed", [(2500, 256, False, 2013), (4000, 100, True, 7), (513, 3, True, 1),
                                                  (2, 3, False, 5), (200, 256, True, 11)])
def test_mt_pivots_follow_the_reference_recursion(n, max_size, ties, seed):
    P = _points(n, 3 * n + 1, ties)
    kd = oracle.KdSlices(P, max_size, seed, pivots="mt")
    order, axes, planes, taken = _dfs_build_mt(P, seed)
    assert kd.draws == taken
    assert np.array_equal(order, kd.order)
    for h in range(kd.heap_used):
        if h in axes:
            assert kd.axis[h] == axes[h] and kd.plane[h].view(np.uint32) == np.float32(planes[h]).view(np.uint32)
        else:
            assert kd.axis[h] == 127
    # the stream goes on with the representatives: the first one draws number taken + 1
    if kd.S:
        reps, nd = oracle.kd_representatives(kd.slices, n, seed, skip=kd.draws)
        u = np.float32(oracle.mt_stream(seed, taken + 1)[taken] & 0xffffff) * np.float32(1.0 / 16777216.0)
        a, b = kd.slices[0]
        assert nd == taken + 2 * kd.S
        assert reps[0] == min(int(np.float32(np.float32(a) + np.float32(u * np.float32(b - a + 1)))), n - 1)
    # the same points under the per-node stream give another tree of the same shape
    kp = oracle.KdSlices(P, max_size, seed)
    assert np.array_equal(kp.slices, kd.slices) and kp.heap_used == kd.heap_used


@pytest.mark.parametrize("pivots", ["philox", "mt"])
@pytest.mark.parametrize("n,max_size,ties", [(5000, 256, False), (4096, 64, True), (1000, 7, True)])
def test_tree_invariants(n, max_size, ties, pivots):
    P = _points(n, 7 * n, ties)
    kd = oracle.KdSlices(P, max_size, 5, pivots=pivots)
    assert np.array_equal(np.sort(kd.order), np.arange(n))
    Q = P[kd.order]
    # slices: depth-first first nodes of at most max_size points, tiling [0, n)
    want = []

    def topo(s, e):
        if s >= e:
            return
        if e - s + 1 <= max_size:
            want.append((s, e))
            return
        m = (s + e) // 2
        topo(s, m)
        topo(m + 1, e)
    topo(0, n - 1)
    assert [tuple(x) for x in kd.slices] == want
    for sid, (s, e) in enumerate(kd.slices):
        assert np.array_equal(kd.boxes[sid, :6], Q[s:e + 1].min(0))
        assert np.array_equal(kd.boxes[sid, 6:], Q[s:e + 1].max(0))
    # every split node: the longest axis, and the plane between its children's keys
    for h in range(kd.heap_used):
        d = int(np.floor(np.log2(h + 1)))
        r = _node_range(n, d, h - (1 << d) + 1)
        if r is None or r[0] >= r[1]:
            continue
        s, e = r
        ax, pl, m = kd.axis[h], kd.plane[h], (s + e) // 2
        ext = np.abs(Q[s:e + 1].max(0) - Q[s:e + 1].min(0))
        assert ext[ax] == ext.max() and (ext[:ax] < ext[ax]).all()
        # the median key is the largest of the left child's (deeper splits move it within [s, m])
        assert Q[s:m + 1, ax].max() == pl and (Q[m + 1:e + 1, ax] >= pl).all()


def test_find_lands_in_own_slice_away_from_planes():
    n = 8000
    P = _points(n, 11, False)
    kd = oracle.KdSlices(P, 128, 3)
    own = np.empty(n, np.int32)
    for sid, (s, e) in enumerate(kd.slices):
        own[kd.order[s:e + 1]] = sid
    got = kd.find(P)
    checked = 0
    for i in range(n):
        node, clear = 0, True
        while kd.cl[node] < 0:
            dist = P[i, kd.axis[node]] - kd.plane[node]
            clear = clear and abs(dist) > 2e-3
            node = 2 * node + 1 if dist <= 0 else 2 * node + 2
        if clear:
            checked += 1
            assert got[i] == own[i]
    assert checked > n // 2


@pytest.mark.parametrize("k", [1, 4, 16])
def test_nearest_equals_brute_force_off_grid(k):
    rng = np.random.default_rng(k)
    P = rng.integers(0, 10, size=(4000, 6)).astype(np.float32)
    kd = oracle.KdSlices(P, 64, 1)
    Q = (rng.integers(0, 10, size=(200, 6)) + rng.choice([0.25, 0.75, -0.25], size=(200, 6))).astype(np.float32)
    got = kd.nearest(Q, k)
    lo, hi = kd.boxes[:, :6].astype(np.float64), kd.boxes[:, 6:].astype(np.float64)
    for q, row in zip(Q.astype(np.float64), got):
        d = (((np.clip(q, lo, hi) - q) ** 2).sum(1))
        kk = min(k, kd.S)
        assert (row[kk:] == -1).all() and len(set(row[:kk])) == kk
        assert np.array_equal(d[row[:kk]], np.sort(d)[:kk][::-1])      # furthest first


def test_nearest_returns_all_slices_when_fewer_than_k():
    P = _points(500, 2, False)
    kd = oracle.KdSlices(P, 256, 1)
    assert kd.S == 2
    out = kd.nearest(P[:3], 5)
    assert (out[:, 2:] == -1).all() and all(set(r[:2]) == {0, 1} for r in out)


def test_degenerate_sizes():
    for n in (0, 1):
        kd = oracle.KdSlices(np.zeros((n, 6), np.float32), 8, 1)
        assert kd.S == 0 and kd.heap_used == 0
        assert (kd.find(np.zeros((2, 6), np.float32)) == -1).all()
    kd = oracle.KdSlices(_points(5, 1, False), 8, 1)       # the root is the only slice
    assert kd.S == 1 and list(kd.slices[0]) == [0, 4] and kd.heap_used == 1
    with pytest.raises(ValueError):
        oracle.KdSlices(_points(5, 1, False), 2, 1)


def test_map_vec6_closed_forms():
    rng = np.random.default_rng(4)
    n = 64
    nrm = rng.normal(size=(n, 3)); nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    wo = rng.normal(size=(n, 3)); wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    wo = np.where((wo * nrm).sum(1, keepdims=True) < 0, -wo, wo)
    rec = np.zeros((n, 21), np.float32)
    rec[:, 0:3] = rng.random((n, 3)) * 4
    rec[:, 6:9] = nrm
    rec[:, 9:12] = wo
    # materials: Lambert, Phong n = 100 / 2500 / 1e4, Ward (0.1, 0.2) / (1e-4, 0.3)
    mats = np.zeros((6, 12), np.float32)
    mats[1, 0], mats[1, 7] = 1, 100
    mats[2, 0], mats[2, 7] = 1, 2500
    mats[3, 0], mats[3, 7] = 1, 1e4
    mats[4, 0], mats[4, 7], mats[4, 8] = 2, 0.1, 0.2
    mats[5, 0], mats[5, 7], mats[5, 8] = 2, 1e-4, 0.3
    gscale = {0: None, 1: 1.0, 2: 2.5, 3: 4.0, 4: 1.0, 5: 4.0}
    diag = 12.0
    ids = np.array([i % 7 - 1 for i in range(n)], np.int32)        # -1: no material
    out = oracle.map_vec6(rec, ids, mats, diag)
    w = diag / 16
    assert np.array_equal(out[:, :3], rec[:, :3])
    for i in range(n):
        m = ids[i]
        dvec = out[i, 3:].astype(np.float64) / w
        if m <= 0:
            assert np.allclose(dvec, nrm[i], atol=1e-6)
            continue
        g = gscale[int(m)]
        assert abs(np.linalg.norm(dvec) - g) < 1e-5 * g
        wr = dvec / g
        assert abs(wr @ nrm[i] - wo[i] @ nrm[i]) < 1e-5                  # same angle to n
        assert np.linalg.norm(np.cross(wr + wo[i], nrm[i])) < 1e-5       # wr + wo along n
