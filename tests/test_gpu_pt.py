"""GPU parity of guided path tracing (rzk_lightslice_render_guided) against the
oracle's restatement (vpl.cc:934-1041, area_light.cc:48-142) on the same
scene, the same per-slice maps read back from the device and the same Philox
streams.

Bar: per-image radiance rel-L2 <= 1e-6 (fp32).  Ray arithmetic is restated
bit-for-bit, and the transcendentals that pick a sampled direction are the
correctly rounded float of the double function on both sides, so both trace
the same rays; what remains is the ulp-level difference of the float expf /
powf inside the Phong and Ward evaluations (glibc vs the device library).  The
test reports how many pixels differ at all."""
import numpy as np
import pytest

import oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _setup(ctx, sc, nvpl, tile, clusters, seed=5):
    import renzoku_amd as rz
    from renzoku_amd import vplgen
    gs = rz.Scene(ctx, sc)
    vpl, vmat, ne = vplgen.device_vpls(gs, nvpl, seed=seed)
    vp = rz.Vpls(ctx, vpl, vmat)
    W, H = int(sc.camera[17]), int(sc.camera[18])
    ls = rz.LightSlice(ctx, gs, vp, width=W, height=H, tile=tile, num_clusters=clusters, num_neighbor_slices=4,
                       n_emitted=ne, seed=2013)
    ls.initialize(sc.camera)
    return gs, vp, ls


def _guide(ls):
    inf = ls.info()
    S = inf.num_slices
    res = ls.res
    return dict(res=res, maps=ls.get("maps", np.float32, S * res * res), row_cdf=ls.get("row_cdf", np.float32, S * res),
                col_cdf=ls.get("col_cdf", np.float32, S * res * res), totals=ls.get("map_totals", np.float32, S))


def _compare(sc, ls, osc, tile, **kw):
    img, rays = ls.render_guided(sc.camera, **kw)
    okw = dict(kw)
    window = okw.pop("window", None)
    okw["direct"] = okw.pop("direct_lighting")
    ref = oracle.guided_pt(osc, sc.camera, tile=tile, window=window, nthreads=8, **_guide(ls), **okw)
    err = rel_l2(img, ref)
    differ = int((np.abs(img - ref).max(2) > 1e-5 * np.maximum(np.abs(ref).max(2), 1e-3)).sum())
    print(f"guided pt {sc.name} {kw}: rel-L2 {err:.3e}, {differ} of {img.shape[0] * img.shape[1]} pixels differ, "
          f"rays {rays}", flush=True)
    assert np.isfinite(img).all()
    assert ref.sum() > 0
    assert err <= 1e-6
    return img, ref, rays


@pytest.mark.parametrize("guide_mode", [0, 1])
def test_guided_pt_cornell(ctx, guide_mode):
    from renzoku_amd import scenes
    sc = scenes.cornell_box(48, 48)
    gs, vp, ls = _setup(ctx, sc, 512, 16, 32)
    osc = oracle.OrcScene(sc)
    img, ref, rays = _compare(sc, ls, osc, 16, spp=4, max_bounce=5, direct_lighting=1, guide_mode=guide_mode,
                              seed=2013, tmin=1e-3)
    assert rays["closest_rays"] >= 48 * 48 * 4 and rays["shadow_rays"] > 0


def test_guided_pt_room_phong(ctx, small_room):
    sc = small_room
    gs, vp, ls = _setup(ctx, sc, 2048, 16, 64)
    osc = oracle.OrcScene(sc)
    _compare(sc, ls, osc, 16, spp=4, max_bounce=4, direct_lighting=1, guide_mode=1, seed=7, tmin=1e-3)


def test_guided_pt_ward(ctx):
    from renzoku_amd import scenes
    sc = scenes.instanced_meshes(64, 48, target_tris=30_000)
    gs, vp, ls = _setup(ctx, sc, 2048, 16, 64)
    osc = oracle.OrcScene(sc)
    _compare(sc, ls, osc, 16, spp=4, max_bounce=4, direct_lighting=1, guide_mode=1, seed=3, tmin=1e-3)


def test_guided_pt_window_is_the_same_pixels(ctx):
    """a pixel's radiance depends only on (seed, pixel, sample): a window
    renders the same bits as the full frame, the rest reads back as 0"""
    from renzoku_amd import scenes
    sc = scenes.cornell_box(40, 40)
    gs, vp, ls = _setup(ctx, sc, 256, 8, 16)
    kw = dict(spp=2, max_bounce=5, direct_lighting=1, guide_mode=1, seed=1, tmin=1e-3)
    full, _ = ls.render_guided(sc.camera, **kw)
    win, _ = ls.render_guided(sc.camera, window=(7, 5, 20, 30), **kw)
    assert np.array_equal(full[5:35, 7:27].view(np.uint32), win[5:35, 7:27].view(np.uint32))
    mask = np.ones((40, 40), bool)
    mask[5:35, 7:27] = False
    assert not win[mask].any()
    again, _ = ls.render_guided(sc.camera, **kw)
    assert np.array_equal(full.view(np.uint32), again.view(np.uint32))
