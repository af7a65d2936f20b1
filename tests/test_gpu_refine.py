"""LightSlice per-slice refinement on the GPU (rzk_refine_slices and
rzk_lightslice_* with max_slice_clusters > the kept clusters) against the
oracle's restatement (oracle/refine.cpp, lightslice.cc:837-902): every
slice's cluster count, representatives and weights bit for bit, on inputs
full of equal projections (columns that are zero over a slice's rows, where
std::sort's tie order decides), on degenerate ones (one- and two-member
clusters, early stops: slices that take fewer draws than the rule, so later
slices run again at guessed offsets until the chain of exact offsets reaches
them), on lines longer than the
LDS sort buffer, and through the whole LightSlice pipeline (per-slice pdfs,
maps and the final gather over each slice's own representatives)."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def _case(V, S, nnb, ncl, maxc, seed, zero_frac, nb=None):
    rng = np.random.default_rng(seed)
    R = (rng.uniform(0, 1, (V, S, 3)) * (rng.uniform(size=(V, S, 1)) > zero_frac)).astype(np.float32)
    if nb is None:
        nb = np.array([[(s + d) % S for d in range(nnb)] for s in range(S)], np.int32)
    o = oracle.cluster_and_refine(R, ncl, maxc, nb, seed)
    # the numbers that follow the clustering's in the same stream
    oracle.cluster_by_sampling(R, ncl, seed)
    L = oracle.lib()
    n = 4 * (maxc - len(o["reps"])) * S
    draws = np.array([L.orc_mt_rand() for _ in range(n)], np.float32)
    return R, nb, o, draws


def _check(ctx, R, nb, o, draws, maxc):
    import renzoku_amd as rz
    count, rep, w = rz.refine_slices(ctx, R, nb, o["off"], o["mem"], o["reps"], o["rep_w"], maxc, draws)
    assert np.array_equal(count, o["count"])
    for s in range(R.shape[1]):
        n = count[s]
        assert np.array_equal(rep[s, :n], o["rep"][s, :n]), f"slice {s} representatives"
        assert np.array_equal(w[s, :n].view(np.uint32), o["w"][s, :n].view(np.uint32)), f"slice {s} weights"
    return count


@pytest.mark.parametrize("V,S,nnb,ncl,maxc,seed,zero_frac", [
    (400, 16, 4, 12, 30, 1, 0.7),        # many columns zero over a slice: equal projections
    (300, 8, 3, 10, 24, 2, 0.0),         # continuous values
    (60, 12, 4, 20, 60, 3, 0.5),         # singletons and pairs: early stops, short draw counts
    (2000, 6, 2, 4, 9, 4, 0.3),
    (240, 320, 4, 10, 24, 5, 0.5),       # more slices than compute units: long chains of guessed offsets
])
def test_refine_slices_matches_oracle(ctx, V, S, nnb, ncl, maxc, seed, zero_frac):
    R, nb, o, draws = _case(V, S, nnb, ncl, maxc, seed, zero_frac)
    count = _check(ctx, R, nb, o, draws, maxc)
    print(f"V={V} S={S}: clusters per slice {count.min()}..{count.max()}, draws {o['draws'].min()}..{o['draws'].max()}",
          flush=True)


def test_refine_slices_short_draw_counts_are_corrected(ctx):
    """slices that stop early take fewer than 4 draws per split; the later
    slices must read the stream from where those stopped"""
    R, nb, o, draws = _case(40, 10, 3, 16, 40, 7, 0.3)
    per = 4 * (40 - len(o["reps"]))
    assert (o["draws"] < per).any()
    _check(ctx, R, nb, o, draws, 40)


def test_refine_slices_long_line_and_widest_neighbourhood(ctx):
    """a cluster of more than 4,096 members (its projections sort in HBM
    scratch) and 32 neighbour rows per slice (96 flattened rows)"""
    R, nb, o, draws = _case(9000, 3, 2, 2, 4, 8, 0.2)
    assert (np.diff(o["off"]) > 4098).any()
    _check(ctx, R, nb, o, draws, 4)
    rng = np.random.default_rng(9)
    S = 40
    nb32 = np.array([rng.permutation(S)[:32] for _ in range(S)], np.int32)
    R, nb, o, draws = _case(500, S, 32, 8, 20, 10, 0.5, nb=nb32)
    _check(ctx, R, nb, o, draws, 20)


def test_refine_slices_rejects_bad_input(ctx):
    import renzoku_amd as rz
    R, nb, o, draws = _case(100, 4, 2, 5, 10, 11, 0.0)
    with pytest.raises(rz.RzkError):
        rz.refine_slices(ctx, R, nb, o["off"], o["mem"], o["reps"], o["rep_w"], 10, draws[:5])   # too few draws
    with pytest.raises(rz.RzkError):
        rz.refine_slices(ctx, R, nb, o["off"], o["mem"], o["reps"], o["rep_w"], 3, draws)       # max < clusters
    bad = nb.copy()
    bad[0, 0] = 99
    with pytest.raises(rz.RzkError):
        rz.refine_slices(ctx, R, bad, o["off"], o["mem"], o["reps"], o["rep_w"], 10, draws)


@pytest.fixture(scope="module")
def room_refined(ctx):
    import renzoku_amd as rz
    from renzoku_amd import scenes, vplgen
    sc = scenes.room(128, 128, seed=11, target_tris=20_000)
    gs = rz.Scene(ctx, sc)
    vpl, vmat, ne = vplgen.device_vpls(gs, 2048, seed=5)
    vp = rz.Vpls(ctx, vpl, vmat)
    ls = rz.LightSlice(ctx, gs, vp, width=128, height=128, tile=16, num_clusters=32, num_neighbor_slices=4,
                       n_emitted=ne, seed=2013, max_slice_clusters=64)
    ls.initialize(sc.camera)
    return sc, gs, vpl, vmat, ne, ls


def test_lightslice_refined_pipeline(ctx, room_refined):
    from renzoku_amd import scenes
    sc, gs, vpl, vmat, ne, ls = room_refined
    W = H = 128
    T, tx = 16, 8
    inf = ls.info()
    S, V = inf.num_slices, len(vmat)
    R = ls.get("R", np.float32, S * V * 3).reshape(V, S, 3)
    nb = ls.get("neighbors", np.int32, S * 4).reshape(S, 4)
    o = oracle.cluster_and_refine(R, 32, 64, nb, 2013)
    cnt = ls.get("slice_count", np.int32, S)
    reps = ls.get("slice_reps", np.int32, S * 64).reshape(S, 64)
    rw = ls.get("slice_rep_w", np.float32, S * 64).reshape(S, 64)
    assert np.array_equal(cnt, o["count"]) and cnt.max() == 64
    for s in range(S):
        assert np.array_equal(reps[s, :cnt[s]], o["rep"][s, :cnt[s]])
        assert np.array_equal(rw[s, :cnt[s]].view(np.uint32), o["w"][s, :cnt[s]].view(np.uint32))
    # per-slice cluster pdfs over the slice's own representatives
    upper = ls.get("slice_upper", np.int32, S * 64).reshape(S, 64)
    gpdf = ls.get("slice_pdf", np.float32, S * 64).reshape(S, 64)
    for s in range(0, S, 7):
        n = cnt[s]
        w = oracle.slice_cluster_pdf(R, nb[s], reps[s, :n], rw[s, :n], upper[s, :n])
        ok, pdf, cdf = oracle.pdf_build(w)
        assert np.array_equal(gpdf[s, :n].view(np.uint32), pdf.view(np.uint32))
    # the gather: every pixel sums over its own slice's representatives
    img = ls.gather()
    osc = oracle.OrcScene(sc)
    px, py = scenes.pixel_grid(W, H)
    rays = oracle.camera_rays(sc.camera, px, py)
    t, b, g, tri = osc.trace_closest(rays)
    recv, m, _ = osc.fill_receivers(rays, t, b, g, tri)
    ref = np.zeros((W * H, 3), np.float32)
    sl = (py // T) * tx + px // T
    for s in range(S):
        sel = np.nonzero(sl == s)[0]
        n = cnt[s]
        ref[sel] = osc.gather_clusters(recv[sel], m[sel], vpl[reps[s, :n]], vmat[reps[s, :n]], rw[s, :n], 1, 0.01,
                                       1e-3, ne, nthreads=8)
    assert rel_l2(img.reshape(-1, 3), ref) <= 1e-4
    work = ls.count_work()
    assert work["gather_pairs"] == int(sum(cnt[sl[i]] for i in np.nonzero(m >= 0)[0]))
