"""Batched guided path tracing (rzk_batch_*): many uneven jobs in one
persistent, graph-launched kernel.  Every job's image must be bit-equal to the
plain per-request render (rzk_lightslice_render_guided) of the same window,
whatever else the batch holds, across graph replays, launch shapes and
autotuning; the plain render is itself pinned to the oracle (test_gpu_pt)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def two_guides(ctx):
    """the room scene with two LightSlices built from 1,024 and 3,072 VPLs"""
    import renzoku_amd as rz
    from renzoku_amd import scenes, vplgen
    sc = scenes.room(96, 80, seed=11, target_tris=20_000)
    gs = rz.Scene(ctx, sc)
    out = []
    for nv, seed in [(1024, 3), (3072, 4)]:
        vpl, vmat, ne = vplgen.device_vpls(gs, nv, seed=seed)
        vp = rz.Vpls(ctx, vpl, vmat)
        ls = rz.LightSlice(ctx, gs, vp, width=96, height=80, tile=16, num_clusters=48, num_neighbor_slices=4,
                           n_emitted=ne, seed=2013)
        ls.initialize(sc.camera)
        out.append((vp, ls, ls.guide(sc.camera)))
    return sc, gs, out


def _specs(guides, rng, n):
    specs = []
    for i in range(n):
        g = guides[i % len(guides)]
        w, h = int(rng.integers(1, 60)), int(rng.integers(1, 50))
        x0, y0 = int(rng.integers(0, 96 - w + 1)), int(rng.integers(0, 80 - h + 1))
        specs.append(dict(guide=g[2], ls=g[1], window=(x0, y0, w, h), spp=int(rng.choice([1, 2, 3, 4, 9])),
                          max_bounce=int(rng.integers(1, 6)), direct_lighting=int(rng.integers(0, 2)),
                          guide_mode=int(rng.integers(0, 2)), seed=int(rng.integers(0, 2 ** 31))))
    return specs


def _plain(sc, spec):
    img, rays = spec["ls"].render_guided(sc.camera, spp=spec["spp"], max_bounce=spec["max_bounce"],
                                         direct_lighting=spec["direct_lighting"], guide_mode=spec["guide_mode"],
                                         seed=spec["seed"], window=spec["window"])
    x0, y0, w, h = spec["window"]
    return img[y0:y0 + h, x0:x0 + w], rays


def test_batch_equals_plain_per_job(ctx, two_guides):
    import renzoku_amd as rz
    sc, gs, guides = two_guides
    specs = _specs(guides, np.random.default_rng(1), 24)
    b = rz.Batch(ctx, gs)
    imgs, rays, ms = b.render(specs, count_rays=True)
    tot_c = tot_s = 0
    for k, spec in enumerate(specs):
        ref, r = _plain(sc, spec)
        assert imgs[k].shape == ref.shape
        assert np.array_equal(imgs[k].view(np.uint32), ref.view(np.uint32)), f"job {k} {spec['window']}"
        tot_c += r["closest_rays"]
        tot_s += r["shadow_rays"]
    assert rays == {"closest_rays": tot_c, "shadow_rays": tot_s}
    assert ms > 0
    print(f"batch of {len(specs)} jobs: {ms:.3f} ms, rays {rays}", flush=True)


def test_batch_graph_replays_and_shapes(ctx, two_guides):
    """one captured graph replayed for batches of different sizes with eager
    work between the replays; other launch shapes give the same bits"""
    import renzoku_amd as rz
    sc, gs, guides = two_guides
    big = _specs(guides, np.random.default_rng(2), 20)
    small = big[3:7]
    b = rz.Batch(ctx, gs, dict(block=128, blocks_per_cu=2, items_per_dequeue=1, samples_per_item=2, engine=0))
    first, _, _ = b.render(big)
    for rep in range(3):
        _plain(sc, big[rep])                         # eager work on the same stream between replays
        again, _, _ = b.render(big)
        for a, c in zip(first, again):
            assert np.array_equal(a.view(np.uint32), c.view(np.uint32))
        part, _, _ = b.render(small)
        for a, c in zip(first[3:7], part):
            assert np.array_equal(a.view(np.uint32), c.view(np.uint32))
    for cfg in [dict(block=64, blocks_per_cu=8, items_per_dequeue=4, samples_per_item=1, engine=0),
                dict(block=256, blocks_per_cu=1, items_per_dequeue=2, samples_per_item=16, engine=0),
                dict(block=256, blocks_per_cu=8, engine=1, path_slots=1 << 20),
                dict(block=128, blocks_per_cu=4, engine=1, path_slots=4096),     # far fewer slots than paths
                dict(block=64, blocks_per_cu=2, engine=1, path_slots=20000)]:
        b.config = cfg
        assert b.config == dict(rz._lib.BatchConfig.DEFAULTS, **cfg)
        other, _, _ = b.render(big)
        for a, c in zip(first, other):
            assert np.array_equal(a.view(np.uint32), c.view(np.uint32))


def test_batch_autotune(ctx, two_guides):
    import renzoku_amd as rz
    sc, gs, guides = two_guides
    specs = _specs(guides, np.random.default_rng(3), 12)
    b = rz.Batch(ctx, gs)
    before, _, _ = b.render(specs)
    best, tried = b.autotune(specs)
    assert len(tried) >= 6
    assert best == b.config
    assert min(ms for _, ms in tried) == pytest.approx([ms for c, ms in tried if c == best][0])
    after, _, _ = b.render(specs)
    for a, c in zip(before, after):
        assert np.array_equal(a.view(np.uint32), c.view(np.uint32))


def test_wavefront_equals_persistent_kernel_counts(ctx, two_guides):
    """the wavefront engine traces exactly the rays the one-lane-per-pixel
    kernel traces and gives the same bits"""
    import renzoku_amd as rz
    sc, gs, guides = two_guides
    specs = _specs(guides, np.random.default_rng(4), 16)
    a = rz.Batch(ctx, gs, dict(engine=0))
    w = rz.Batch(ctx, gs, dict(engine=1, path_slots=8192))
    ia, ra, _ = a.render(specs, count_rays=True)
    iw, rw, _ = w.render(specs, count_rays=True)
    assert ra == rw
    for x, y in zip(ia, iw):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


def test_batch_rejects_bad_jobs(ctx, two_guides):
    import renzoku_amd as rz
    sc, gs, guides = two_guides
    b = rz.Batch(ctx, gs)
    bad = dict(guide=guides[0][2], window=(90, 0, 10, 10), spp=1)
    with pytest.raises(rz.RzkError):
        b.render([bad])
    with pytest.raises(rz.RzkError):
        b.render([dict(guide=guides[0][2], window=(0, 0, 8, 8), spp=0)])
    ok, _, _ = b.render([dict(guide=guides[0][2], window=(0, 0, 8, 8), spp=1)])
    assert ok[0].shape == (8, 8, 3)
