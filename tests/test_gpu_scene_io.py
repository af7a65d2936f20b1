"""A scene read from the reference's JSON + OBJ files renders on the device
like the oracle renders it, and the frame goes out through the progressive
frame buffer to EXR / PPM files."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def test_json_scene_renders_and_saves(tmp_path):
    import renzoku_amd as rz
    from renzoku_amd import frame_io, scene_io, scenes
    src = scenes.cornell_box(48, 40)
    sc = scene_io.load_json_scene(scene_io.save_json_scene(src, str(tmp_path), "cornell"))
    ctx = rz.Ctx(0)
    gs = rz.Scene(ctx, sc)
    vpl, vmat, used = rz.generate_vpls(ctx, gs, 512, seed=7)
    vp = rz.Vpls(ctx, vpl, vmat)
    img = rz.render_vpl(ctx, gs, vp, sc.camera, 1, 0.01, 1e-3, float(used))
    osc = oracle.OrcScene(sc)
    ov, om, ou = osc.generate_vpls(512, seed=7)
    assert ou == used and np.array_equal(ov.view(np.uint32), vpl.view(np.uint32))
    px, py = scenes.pixel_grid(48, 40)
    rays = oracle.camera_rays(sc.camera, px, py)
    t, b, g, tri = osc.trace_closest(rays)
    recv, m, _ = osc.fill_receivers(rays, t, b, g, tri)
    ref = osc.gather(recv, m, vpl, vmat, 1, 0.01, 1e-3, float(used), nthreads=8).reshape(40, 48, 3)
    assert np.linalg.norm(img - ref) <= 1e-4 * np.linalg.norm(ref) and img.sum() > 0
    fb = frame_io.FrameBuffer(48, 40)
    fb.update(img)
    fb.complete_frame()
    fb.update(img)
    frame_io.save_exr(tmp_path / "f.exr", fb.frame)
    frame_io.save_ppm(tmp_path / "f.ppm", fb.bytes)
    back = frame_io.load_exr(tmp_path / "f.exr")
    assert np.allclose(back[..., :3], img, rtol=1e-3, atol=1e-6) and fb.bytes.max() > 0
