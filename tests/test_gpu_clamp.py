"""GPU parity for every clamp mode of VplRadianceClamp (vpl_interface.h:50-123,
vpl.cc:214-230): NONE, FORM_FACTOR, REFLECTIVITY and RADIANCE, in the gather
(rzk_gather), the reduced matrix's incident radiance (rzk_reduced_matrix) and
the LightSlice pipeline (rzk_lightslice_*), against the oracle on the same
inputs.  Thresholds sit near the median of each clamped quantity on these
scenes, so both branches of every clamp run; the reduced-matrix rows include
receivers whose shading normal is tilted away from the geometric one, where
incident radiance must use the geometric normal (vpl.cc:413).

Bars as in test_gpu_parity: visibility exact, Lambertian-only radiance
<= 1e-6 rel-L2 (summation order), Phong radiance <= 1e-4."""
import numpy as np
import pytest

import oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu

# (mode, threshold) pairs near the medians of G (form factor), |G fs fl|
# (reflectivity) and |P G fs fl| (radiance) of gather pairs in the Cornell box
CORNELL_GATHER = [(0, 0.0), (1, 3e-6), (2, 8e-8), (3, 0.04)]
# the same for incident radiance (fs white) in the 20k-triangle room
ROOM_INCIDENT = [(0, 0.0), (1, 1.5e-6), (2, 2.25e-7), (3, 0.85)]
# (mode, gather clamp, incident clamp) for the room's LightSlice
ROOM_LIGHTSLICE = [(0, 0.0, 0.0), (2, 5.75e-8, 2.25e-7), (3, 0.217, 0.85)]


@pytest.fixture(scope="module")
def cornell_setup(ctx, cornell):
    import renzoku_amd as rz
    from renzoku_amd import scenes, vplgen
    gs = rz.Scene(ctx, cornell)
    osc = oracle.OrcScene(cornell)
    vpl, vmat, ne = vplgen.device_vpls(gs, 1024, seed=2013)
    px, py = scenes.pixel_grid(256, 256)
    sel = np.arange(0, 256 * 256, 5)
    rays = oracle.camera_rays(cornell.camera, px[sel], py[sel])
    t, b, g, tri = osc.trace_closest(rays)
    recv, m, _ = osc.fill_receivers(rays, t, b, g, tri)
    return gs, osc, vpl, vmat, ne, rz.Vpls(ctx, vpl, vmat), recv, m


@pytest.mark.parametrize("mode,thr", CORNELL_GATHER)
def test_gather_clamp_modes_cornell(ctx, cornell_setup, mode, thr):
    import renzoku_amd as rz
    gs, osc, vpl, vmat, ne, vp, recv, m = cornell_setup
    out = rz.gather(ctx, gs, vp, recv, m, mode, thr, 1e-3, ne)
    ref = osc.gather(recv, m, vpl, vmat, mode, thr, 1e-3, ne, nthreads=8)
    assert rel_l2(out, ref) <= 1e-6
    if mode != 0:
        # the clamp took effect: the unclamped image differs
        free = osc.gather(recv, m, vpl, vmat, 0, 0.0, 1e-3, ne, nthreads=8)
        assert rel_l2(ref, free) > 1e-3


def _tilted(recv, m, seed):
    """every second lit receiver with its shading normal (and frame) turned 25-50
    degrees away from its geometric normal"""
    rng = np.random.default_rng(seed)
    recv = recv.copy()
    for i in range(0, len(recv), 2):
        if m[i] < 0:
            continue
        ng = recv[i, 3:6].astype(np.float64)
        t = np.cross(ng, rng.normal(size=3))
        t /= np.linalg.norm(t)
        a = np.radians(rng.uniform(25, 50))
        ns = np.cos(a) * ng + np.sin(a) * t
        u = np.cross(ns, [1.0, 0, 0] if abs(ns[0]) < 0.9 else [0, 1.0, 0])
        u /= np.linalg.norm(u)
        v = np.cross(ns, u)
        recv[i, 6:9] = ns
        recv[i, 12:15], recv[i, 15:18], recv[i, 18:21] = u, v, ns
    return recv


@pytest.fixture(scope="module")
def room_setup(ctx, small_room):
    import renzoku_amd as rz
    from renzoku_amd import vplgen
    gs = rz.Scene(ctx, small_room)
    osc = oracle.OrcScene(small_room)
    vpl, vmat, ne = vplgen.device_vpls(gs, 2048, seed=5)
    return gs, osc, vpl, vmat, ne, rz.Vpls(ctx, vpl, vmat)


@pytest.mark.parametrize("mode,thr", ROOM_INCIDENT)
def test_reduced_matrix_clamp_modes_room(ctx, small_room, room_setup, mode, thr):
    import renzoku_amd as rz
    from renzoku_amd import vplgen
    gs, osc, vpl, vmat, ne, vp = room_setup
    px, py = vplgen.tile_centre_pixels(128, 128, 16)
    rays = oracle.camera_rays(small_room.camera, px, py)
    t, b, g, tri = osc.trace_closest(rays)
    recv, m, _ = osc.fill_receivers(rays, t, b, g, tri)
    recv = _tilted(recv, m, 7 + mode)
    R, V = rz.reduced_matrix(ctx, gs, vp, recv, m, mode, thr, 1e-3)
    oR, oV = osc.reduced_matrix(recv, m, vpl, vmat, mode, thr, 1e-3, nthreads=8)
    assert np.array_equal(V, oV)
    assert rel_l2(R, oR) <= 1e-4
    assert np.abs(oR).sum() > 0
    if mode != 0:
        fR, _ = osc.reduced_matrix(recv, m, vpl, vmat, 0, 0.0, 1e-3, nthreads=8)
        assert rel_l2(oR, fR) > 1e-3


@pytest.mark.parametrize("mode,clamp,inc", ROOM_LIGHTSLICE)
def test_lightslice_clamp_modes_room(ctx, small_room, room_setup, mode, clamp, inc):
    import renzoku_amd as rz
    from renzoku_amd import scenes, vplgen
    gs, osc, vpl, vmat, ne, vp = room_setup
    W = H = 128
    ls = rz.LightSlice(ctx, gs, vp, width=W, height=H, tile=16, num_clusters=64, num_neighbor_slices=4,
                       n_emitted=ne, seed=2013, clamp_mode=mode, clamp=clamp, incident_clamp=inc)
    ls.initialize(small_room.camera)
    img = ls.gather()
    inf = ls.info()
    S, V = inf.num_slices, len(vmat)
    R = ls.get("R", np.float32, S * V * 3).reshape(V, S, 3)
    px, py = vplgen.tile_centre_pixels(W, H, 16)
    rays = oracle.camera_rays(small_room.camera, px, py)
    t, b, g, tri = osc.trace_closest(rays)
    rrecv, rm, _ = osc.fill_receivers(rays, t, b, g, tri)
    oR, _ = osc.reduced_matrix(rrecv, rm, vpl, vmat, mode, inc, 1e-3, nthreads=8)
    assert rel_l2(R, oR) <= 1e-4
    # clustering on the device's R, replayed by the oracle on the same MT stream
    oc = oracle.cluster_by_sampling(R, 64, 2013, nthreads=8)
    reps = ls.get("reps", np.int32, inf.num_clusters)
    rep_w = ls.get("rep_w", np.float32, inf.num_clusters)
    assert np.array_equal(reps, oc["reps"])
    pxa, pya = scenes.pixel_grid(W, H)
    prays = oracle.camera_rays(small_room.camera, pxa, pya)
    t, b, g, tri = osc.trace_closest(prays)
    recv, m, _ = osc.fill_receivers(prays, t, b, g, tri)
    ref = osc.gather_clusters(recv, m, vpl[reps], vmat[reps], rep_w, mode, clamp, 1e-3, ne,
                              nthreads=8).reshape(H, W, 3)
    assert rel_l2(img, ref) <= 1e-4
    # the per-slice upper-hemisphere flags use the incident clamp of this mode
    upper = ls.get("slice_upper", np.int32, S * len(reps)).reshape(S, -1)
    Vd = ls.get("V", np.float32, S * V).reshape(V, S)
    for s in range(0, S, 9):
        _, up = osc.slice_cluster_info(rrecv[s], reps, vpl, vmat, Vd[:, s].copy(), mode, inc)
        assert np.array_equal(up, upper[s]), s
