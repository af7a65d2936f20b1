"""rzk_gather_dev_noalloc (the gather with its visibility bits in a static
device array, no allocation) against rzk_gather on the same inputs, bit for
bit, over two calls in a row, the second with a smaller input (fewer receivers
and a smaller VPL set), plus the size limit it must refuse."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(ctx, arr):
    p = C.c_void_p()
    ctx.check(ctx.lib.rzk_device_alloc(ctx.h, arr.nbytes, C.byref(p)), "rzk_device_alloc")
    ctx.check(ctx.lib.rzk_copy_h2d(ctx.h, p, arr.ctypes.data_as(C.c_void_p), arr.nbytes), "rzk_copy_h2d")
    return p


def _noalloc(ctx, scene, vpls, recv, rmat, mode, clamp, ne):
    import renzoku_amd as rz
    recv = rz._lib.f32(recv)
    rmat = rz._lib.i32(rmat)
    d_recv, d_mat = _dev(ctx, recv), _dev(ctx, rmat)
    out = np.full((len(rmat), 3), np.nan, np.float32)
    d_out = _dev(ctx, out)
    try:
        print(f"noalloc gather: {len(rmat)} receivers x {vpls.n} VPLs, clamp mode {mode}", flush=True)
        ctx.check(ctx.lib.rzk_gather_dev_noalloc(ctx.h, scene.h, vpls.h, d_recv, d_mat, len(rmat), mode, clamp,
                                                 1e-3, ne, d_out), "rzk_gather_dev_noalloc")
        ctx.synchronize()
        ctx.check(ctx.lib.rzk_copy_d2h(ctx.h, out.ctypes.data_as(C.c_void_p), d_out, out.nbytes), "rzk_copy_d2h")
    finally:
        for p in (d_recv, d_mat, d_out):
            ctx.lib.rzk_device_free(ctx.h, p)
    return out


def test_noalloc_gather_matches_gather(ctx, cornell):
    import renzoku_amd as rz
    from renzoku_amd import scenes, vplgen
    gs = rz.Scene(ctx, cornell)
    vpl, vmat, ne = vplgen.device_vpls(gs, 512, seed=2013)
    big = rz.Vpls(ctx, vpl, vmat)
    small = rz.Vpls(ctx, vpl[:77], vmat[:77])
    px, py = scenes.pixel_grid(256, 256)
    sel = np.arange(0, 256 * 256, 3)
    _, _, recv, m, _ = gs.camera_primary(cornell.camera, px[sel], py[sel])
    assert (m >= 0).sum() > 0 and (m < 0).sum() > 0   # rays onto the light and lit surfaces both present
    # first call: 21,846 receivers x 512 VPLs; second: 1,001 receivers (a ragged
    # last group) x 77 VPLs, so the second call leaves most of the first call's
    # bits in the shared scratch
    for n, vp, mode, thr in ((len(sel), big, 1, 0.01), (1001, small, 1, 0.01)):
        a = _noalloc(ctx, gs, vp, recv[:n], m[:n], mode, thr, ne)
        b = rz.gather(ctx, gs, vp, recv[:n], m[:n], mode, thr, 1e-3, ne)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f"{n} receivers"
        assert np.abs(a).sum() > 0


def test_noalloc_gather_refuses_oversize(ctx, cornell):
    import renzoku_amd as rz
    gs = rz.Scene(ctx, cornell)
    vpl = np.zeros((4097, rz._lib.VPL_FLOATS), np.float32)
    vpl[:, 4] = 1.0
    vp = rz.Vpls(ctx, vpl, np.zeros(4097, np.int32))
    # 65,536 receivers = 1,024 groups; x 4,097 VPLs is past the 4,194,304-word
    # scratch, so the call must refuse before any launch (null device pointers
    # are never read)
    n = 65536
    rc = ctx.lib.rzk_gather_dev_noalloc(ctx.h, gs.h, vp.h, C.c_void_p(16), C.c_void_p(16), n, 1, 0.01, 1e-3, 1.0,
                                        C.c_void_p(16))
    assert rc == 1
    assert b"no-allocation scratch" in ctx.lib.rzk_ctx_last_error(ctx.h)
