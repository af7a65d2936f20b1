"""Frame output: the reference's progressive frame buffer and its image files.

  * FrameBuffer (frame.h:55-101, frame.cc:29-82): the latest progressive
    image, its tone-mapped copy and an 8-bit copy (tone-mapped value x 255,
    truncated).  Before the first complete frame the tone map is
    GammaToneMap(2.2) (clamp01(v^(1/2.2)), tonemap.cc:9-27); after it, the
    Reinhard map (tonemap.cc:33-160) with the log-average luminance of the
    last complete frame (precompute): L = 0.27 R + 0.67 G + 0.06 B (0 for a
    NaN pixel), avg = exp(sum(log(1e-4 + L)) / N) summed in pixel order in
    float, scale = 0.18 / avg, each colour times (sL / (1 + sL)) / L (0 where
    L <= 0), clamped to [0, 1].
  * save_ppm (image.h:303-342): P6 bytes; save_pfm / load_pfm: float RGB;
    save_exr (image.h:344-395): half-float RGBA with alpha 1, written as an
    uncompressed scanline OpenEXR file (the reference calls OpenEXR, which
    this image does not have, so the file is laid out here: magic, version 2,
    the required header attributes, the offset table, one line per chunk).

Arithmetic follows the reference's types: the luminance weights are double
literals, so L is formed in double and stored as float; the log sum is a
float accumulation in pixel order.
"""
from __future__ import annotations

import struct

import numpy as np

F = np.float32


def clamp01(a):
    """ToneMap::clamp01 (tonemap.h:14-16): a > 1 ? 1 : (a < 0 ? 0 : a); NaN stays NaN"""
    a = np.asarray(a, F)
    with np.errstate(invalid="ignore"):
        return np.where(a > 1, F(1), np.where(a < 0, F(0), a)).astype(F)


def to_bytes(t, scale=255.0):
    """(Byte)(t * scale) (frame.cc:55-61); a NaN converts as x86-64 does (to 0)"""
    v = (np.asarray(t, F) * F(scale)).astype(F)
    return np.where(np.isnan(v), F(0), v).astype(np.uint8)


def gamma_map(img, gamma=2.2):
    inv = F(1.0) / F(gamma)
    with np.errstate(invalid="ignore"):
        v = np.power(np.asarray(img, F), inv, dtype=F)
    return clamp01(v)


def _luminance(img):
    x = np.asarray(img, F).reshape(-1, 3)
    bad = np.isnan(x).any(1)
    L = (0.27 * x[:, 0].astype(np.float64) + 0.67 * x[:, 1].astype(np.float64)
         + 0.06 * x[:, 2].astype(np.float64)).astype(F)
    L[bad] = 0.0
    return L


def reinhard_scale(img, key=0.18, delta=1e-4):
    """ReinhardToneMap::precompute: key / exp(mean log(delta + L)), float"""
    L = _luminance(img)
    logs = np.log(F(delta) + L, dtype=F)
    acc = F(0.0)
    for v in logs:                                  # float accumulation in pixel order
        acc = F(acc + v)
    avg = F(np.exp(F(acc / F(L.size)), dtype=F))
    return F(F(key) / avg)


def reinhard_map(img, scale):
    """ReinhardToneMap::map (region form) with a precomputed scale"""
    x = np.asarray(img, F).reshape(-1, 3)
    L = _luminance(x)
    ls = (L * F(scale)).astype(F)
    r = (ls.astype(np.float64) / (1.0 + ls.astype(np.float64))).astype(F)
    with np.errstate(divide="ignore", invalid="ignore"):
        ratio = np.where(L > 0, (r / np.where(L > 0, L, F(1))).astype(F), F(0))
    with np.errstate(invalid="ignore"):
        out = (x * ratio[:, None]).astype(F)
    return clamp01(out).reshape(np.shape(img))


class FrameBuffer:
    def __init__(self, width, height):
        self.w, self.h = int(width), int(height)
        self.reset()

    def reset(self):
        z = np.zeros((self.h, self.w, 3), F)
        self.buffer, self.tonemap, self.frame = z.copy(), z.copy(), z.copy()
        self.bytes = np.zeros((self.h, self.w, 3), np.uint8)
        self.index, self.scale, self.rscale = 0, F(255.0), F(1.0)

    def update(self, img, x0=0, y0=0, x1=None, y1=None):
        """on_update_image: copy the window, tone-map it, convert to bytes"""
        x1 = self.w if x1 is None else x1
        y1 = self.h if y1 is None else y1
        win = np.asarray(img, F).reshape(self.h, self.w, 3)[y0:y1, x0:x1]
        self.buffer[y0:y1, x0:x1] = win
        t = reinhard_map(win, self.rscale) if self.index > 0 else gamma_map(win)
        self.tonemap[y0:y1, x0:x1] = t
        self.bytes[y0:y1, x0:x1] = to_bytes(t, self.scale)

    def complete_frame(self):
        """on_complete_frame: keep the frame, precompute the Reinhard scale"""
        self.index += 1
        self.frame = self.buffer.copy()
        self.rscale = reinhard_scale(self.frame)


def save_ppm(path, rgb_bytes):
    a = np.ascontiguousarray(rgb_bytes, np.uint8)
    h, w = a.shape[:2]
    with open(path, "wb") as f:
        f.write(b"P6\n%d %d\n255\n" % (w, h))
        f.write(a.tobytes())


def save_pfm(path, img):
    a = np.asarray(img, F)
    h, w = a.shape[:2]
    with open(path, "wb") as f:
        f.write(b"PF\n%d %d\n-1.0\n" % (w, h))
        f.write(np.ascontiguousarray(a[::-1]).astype("<f4").tobytes())


def load_pfm(path):
    with open(path, "rb") as f:
        if f.readline().strip() != b"PF":
            raise ValueError("not an RGB PFM file")
        w, h = (int(x) for x in f.readline().split())
        scale = float(f.readline())
        data = np.frombuffer(f.read(), "<f4" if scale < 0 else ">f4").astype(F)
    return data.reshape(h, w, 3)[::-1].copy()


def _attr(name, kind, payload):
    return name.encode() + b"\0" + kind.encode() + b"\0" + struct.pack("<i", len(payload)) + payload


def save_exr(path, img):
    """RGBA half, no compression, increasing y; alpha = 1"""
    a = np.asarray(img, F).reshape(np.shape(img)[0], np.shape(img)[1], -1)
    h, w = a.shape[:2]
    rgb = a[..., :3] if a.shape[2] >= 3 else np.repeat(a[..., :1], 3, 2)
    chans = {"A": np.ones((h, w), np.float16), "B": rgb[..., 2].astype(np.float16),
             "G": rgb[..., 1].astype(np.float16), "R": rgb[..., 0].astype(np.float16)}
    chlist = b"".join(n.encode() + b"\0" + struct.pack("<iB3xii", 1, 0, 1, 1) for n in sorted(chans)) + b"\0"
    box = struct.pack("<iiii", 0, 0, w - 1, h - 1)
    header = (_attr("channels", "chlist", chlist) + _attr("compression", "compression", b"\0")
              + _attr("dataWindow", "box2i", box) + _attr("displayWindow", "box2i", box)
              + _attr("lineOrder", "lineOrder", b"\0") + _attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
              + _attr("screenWindowCenter", "v2f", struct.pack("<ff", 0.0, 0.0))
              + _attr("screenWindowWidth", "float", struct.pack("<f", 1.0)) + b"\0")
    head = struct.pack("<ii", 20000630, 2) + header
    line_bytes = 4 + 4 + 2 * w * len(chans)
    start = len(head) + 8 * h
    offsets = struct.pack("<%dQ" % h, *[start + y * line_bytes for y in range(h)])
    with open(path, "wb") as f:
        f.write(head)
        f.write(offsets)
        for y in range(h):
            f.write(struct.pack("<ii", y, 2 * w * len(chans)))
            for n in sorted(chans):
                f.write(chans[n][y].astype("<f2").tobytes())


def load_exr(path):
    """the files save_exr writes (uncompressed scanline half RGBA) -> (h, w, 4) f32"""
    with open(path, "rb") as f:
        data = f.read()
    magic, ver = struct.unpack_from("<ii", data, 0)
    if magic != 20000630 or (ver & 0xff) != 2:
        raise ValueError("not a scanline OpenEXR file")
    p, attrs = 8, {}
    while data[p] != 0:
        e = data.index(b"\0", p)
        name = data[p:e].decode()
        e2 = data.index(b"\0", e + 1)
        size = struct.unpack_from("<i", data, e2 + 1)[0]
        attrs[name] = data[e2 + 5:e2 + 5 + size]
        p = e2 + 5 + size
    p += 1
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    if attrs["compression"] != b"\0":
        raise ValueError("compressed EXR")
    names, q, cl = [], 0, attrs["channels"]
    while cl[q] != 0:
        e = cl.index(b"\0", q)
        names.append(cl[q:e].decode())
        q = e + 1 + 16
    offsets = struct.unpack_from("<%dQ" % h, data, p)
    out = np.zeros((h, w, len(names)), F)
    for y in range(h):
        o = offsets[y] + 8
        for c in range(len(names)):
            out[y, :, c] = np.frombuffer(data, "<f2", w, o + 2 * w * c).astype(F)
    order = [names.index(c) for c in ("R", "G", "B", "A")]
    return out[..., order]
