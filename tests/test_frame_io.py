"""Frame output (renzoku_amd/frame_io.py): tone maps against their closed
forms, the progressive frame buffer's switch from the gamma map to the
Reinhard map, and the PFM / EXR / PPM files."""
import struct

import numpy as np

from renzoku_amd import frame_io


def test_gamma_map_closed_form():
    v = np.array([[[0.0, 0.5, 1.0], [2.0, -1.0, np.nan]]], np.float32)
    out = frame_io.gamma_map(v)
    assert np.allclose(out[0, 0], [0.0, 0.5 ** (1 / 2.2), 1.0], atol=1e-6)
    # pow of a negative value is NaN, and clamp01 keeps NaN (tonemap.h:14-16)
    assert out[0, 1, 0] == 1.0 and np.isnan(out[0, 1, 1]) and np.isnan(out[0, 1, 2])
    assert list(frame_io.to_bytes(out)[0, 1]) == [255, 0, 0]


def test_reinhard_constant_image_closed_form():
    c = np.array([0.3, 0.6, 0.9])
    img = np.tile(c.astype(np.float32), (8, 8, 1))
    L = 0.27 * c[0] + 0.67 * c[1] + 0.06 * c[2]
    s = frame_io.reinhard_scale(img)
    assert s == np.float32(s) and abs(s - 0.18 / (1e-4 + L)) < 1e-6 * s
    out = frame_io.reinhard_map(img, s)
    want = c * (s * L / (1 + s * L)) / L
    assert np.allclose(out, want, rtol=1e-6)
    # a NaN pixel has zero luminance: its other channels map to zero, the NaN stays
    img[0, 0, 1] = np.nan
    o = frame_io.reinhard_map(img, s)[0, 0]
    assert o[0] == 0 and o[2] == 0 and np.isnan(o[1])


def test_frame_buffer_progression():
    rng = np.random.default_rng(1)
    img = rng.uniform(0, 3, (6, 5, 3)).astype(np.float32)
    fb = frame_io.FrameBuffer(5, 6)
    fb.update(img)
    assert np.array_equal(fb.tonemap, frame_io.gamma_map(img))
    assert np.array_equal(fb.bytes, np.floor(fb.tonemap.astype(np.float64) * 255).astype(np.uint8)) or \
        np.array_equal(fb.bytes, (fb.tonemap * np.float32(255)).astype(np.uint8))
    fb.complete_frame()
    assert fb.index == 1 and np.array_equal(fb.frame, img)
    img2 = (img * 0.5).astype(np.float32)
    fb.update(img2, 1, 2, 4, 5)                          # a window, with the first frame's scale
    win = frame_io.reinhard_map(img2[2:5, 1:4], frame_io.reinhard_scale(img))
    assert np.array_equal(fb.tonemap[2:5, 1:4], win)
    assert np.array_equal(fb.tonemap[0], frame_io.gamma_map(img)[0])


def test_pfm_ppm_exr_files(tmp_path):
    rng = np.random.default_rng(2)
    img = rng.uniform(0, 5, (7, 9, 3)).astype(np.float32)
    frame_io.save_pfm(tmp_path / "a.pfm", img)
    assert np.array_equal(frame_io.load_pfm(tmp_path / "a.pfm"), img)
    b = (rng.uniform(0, 255, (7, 9, 3))).astype(np.uint8)
    frame_io.save_ppm(tmp_path / "a.ppm", b)
    raw = (tmp_path / "a.ppm").read_bytes()
    assert raw.startswith(b"P6\n9 7\n255\n") and raw[len(b"P6\n9 7\n255\n"):] == b.tobytes()
    img[0, 0, 0] = 1e6                                   # beyond half range
    frame_io.save_exr(tmp_path / "a.exr", img)
    raw = (tmp_path / "a.exr").read_bytes()
    assert struct.unpack_from("<ii", raw, 0) == (20000630, 2)
    assert b"channels\0chlist\0" in raw and b"compression\0compression\0" in raw
    back = frame_io.load_exr(tmp_path / "a.exr")
    assert back.shape == (7, 9, 4) and (back[..., 3] == 1).all()
    assert np.array_equal(back[..., :3], img.astype(np.float16).astype(np.float32))
    assert np.isinf(back[0, 0, 0])
