"""Seeded synthetic scenes for the hot-path workloads.

A scene is plain arrays in the layout of include/rzk.h: triangles in surface
order with their tangent (quad.cc:22-24: both triangles of a quad share
unit(p0 - p1); a lone triangle uses its own unit(p0 - p1), triangle.cc:17),
surfaces (material, light, flags), materials (12 floats each) and lights.
Quads split as (p0, p1, p2), (p0, p2, p3) (quad.cc:5-27).  All arithmetic is
float32.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

LAMBERT, PHONG, WARD = 0, 1, 2
TWO_SIDED = 2
F = np.float32


def unit(v):
    v = np.asarray(v, F)
    inv = F(1.0) / np.sqrt(v[..., 0] * v[..., 0] + v[..., 1] * v[..., 1] + v[..., 2] * v[..., 2])
    return (v * inv[..., None]).astype(F)


def cross(a, b):
    a, b = np.asarray(a, F), np.asarray(b, F)
    return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
                     a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                     a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], -1).astype(F)


@dataclass
class SceneData:
    positions: np.ndarray                 # (N, 9)
    tangents: np.ndarray                  # (N, 3)
    tri_surface: np.ndarray               # (N,)
    surf_material: np.ndarray             # (S,)
    surf_light: np.ndarray                # (S,)
    surf_flags: np.ndarray                # (S,)
    materials: np.ndarray                 # (M, 12)
    light_radiance: np.ndarray            # (L, 3)
    camera: np.ndarray                    # (19,)
    light_tris: np.ndarray = field(default=None)  # indices of emitting triangles
    name: str = ""

    @property
    def ntri(self):
        return len(self.tri_surface)


class _Builder:
    def __init__(self):
        self.pos, self.tan, self.tsurf = [], [], []
        self.smat, self.slight, self.sflag = [], [], []
        self.mats, self.lights = [], []

    def material(self, kind, kd, ks=(0, 0, 0), p0=0.0, p1=0.0):
        m = np.zeros(12, F)
        m[0] = kind
        m[1:4], m[4:7] = kd, ks
        m[7], m[8] = p0, p1
        self.mats.append(m)
        return len(self.mats) - 1

    def surface(self, mat=-1, light=-1, flags=TWO_SIDED):
        self.smat.append(mat)
        self.slight.append(light)
        self.sflag.append(flags)
        return len(self.smat) - 1

    def quads(self, surf, q):
        """q: (K, 4, 3) quads p0..p3"""
        q = np.asarray(q, F).reshape(-1, 4, 3)
        t0 = np.concatenate([q[:, 0], q[:, 1], q[:, 2]], 1)
        t1 = np.concatenate([q[:, 0], q[:, 2], q[:, 3]], 1)
        tan = unit(q[:, 0] - q[:, 1])
        tris = np.stack([t0, t1], 1).reshape(-1, 9)
        self.pos.append(tris)
        self.tan.append(np.repeat(tan, 2, 0))
        self.tsurf.append(np.full(len(tris), surf, np.int32))

    def triangles(self, surf, t):
        t = np.asarray(t, F).reshape(-1, 3, 3)
        self.pos.append(t.reshape(-1, 9))
        self.tan.append(unit(t[:, 0] - t[:, 1]))
        self.tsurf.append(np.full(len(t), surf, np.int32))

    def build(self, camera, name):
        pos = np.concatenate(self.pos).astype(F)
        tsurf = np.concatenate(self.tsurf).astype(np.int32)
        slight = np.asarray(self.slight, np.int32)
        light_tris = np.nonzero(slight[tsurf] >= 0)[0].astype(np.int32)
        return SceneData(pos, np.concatenate(self.tan).astype(F), tsurf, np.asarray(self.smat, np.int32), slight,
                         np.asarray(self.sflag, np.int32), np.asarray(self.mats, F).reshape(-1, 12),
                         np.asarray(self.lights, F).reshape(-1, 3), camera, light_tris, name)


def camera(eye, lookat, up, fov_deg, width, height, film=0.025):
    """Camera::Camera + Sensor (camera.cc:12-41, sensor.h:176-182): onb from
    (eye - lookat, up) with init_from_nv; cop = image/2; scale = image/film."""
    eye, lookat, up = (np.asarray(x, F) for x in (eye, lookat, up))
    nvec = eye - lookat
    c = unit(cross(up, nvec))
    nn = unit(nvec)
    vv = cross(nn, c)
    focal = F(film / (2.0 * math.tan(math.radians(fov_deg) / 2.0)))
    cam = np.zeros(19, F)
    cam[0:3], cam[3:6], cam[6:9], cam[9:12] = eye, c, vv, nn
    cam[12], cam[13] = F(width / 2.0), F(height / 2.0)
    cam[14], cam[15] = F(width / film), F(height / film)
    cam[16], cam[17], cam[18] = focal, width, height
    return cam


def _box_quads(center, size, angle, grid=1):
    """6 faces of a box rotated about +y, each split into grid x grid quads."""
    cx, cy, cz = center
    sx, sy, sz = (s / 2.0 for s in size)
    ca, sa = math.cos(angle), math.sin(angle)

    def xf(p):
        x, y, z = p
        return (cx + ca * x + sa * z, cy + y, cz - sa * x + ca * z)

    faces = [  # (origin, edge u, edge v) in box space, outward by right-hand rule of (p0-p1)x(p0-p2)
        ((-sx, -sy, -sz), (0, 0, 2 * sz), (2 * sx, 0, 0)),   # bottom
        ((-sx, sy, -sz), (2 * sx, 0, 0), (0, 0, 2 * sz)),    # top
        ((-sx, -sy, -sz), (2 * sx, 0, 0), (0, 2 * sy, 0)),   # front
        ((-sx, -sy, sz), (0, 2 * sy, 0), (2 * sx, 0, 0)),    # back
        ((-sx, -sy, -sz), (0, 2 * sy, 0), (0, 0, 2 * sz)),   # left
        ((sx, -sy, -sz), (0, 0, 2 * sz), (0, 2 * sy, 0)),    # right
    ]
    out = []
    for o, eu, ev in faces:
        for i in range(grid):
            for j in range(grid):
                a = [o[k] + eu[k] * i / grid + ev[k] * j / grid for k in range(3)]
                b = [o[k] + eu[k] * (i + 1) / grid + ev[k] * j / grid for k in range(3)]
                c = [o[k] + eu[k] * (i + 1) / grid + ev[k] * (j + 1) / grid for k in range(3)]
                d = [o[k] + eu[k] * i / grid + ev[k] * (j + 1) / grid for k in range(3)]
                out.append([xf(a), xf(b), xf(c), xf(d)])
    return np.asarray(out, np.float64)


def _uv_sphere(center, radius, nseg, nring):
    """UV sphere: pole caps as triangles, the bands as quads split in two."""
    cx, cy, cz = center
    th = np.linspace(0.0, math.pi, nring + 1)
    ph = np.linspace(0.0, 2 * math.pi, nseg + 1)
    P = np.stack([cx + radius * np.sin(th)[:, None] * np.cos(ph)[None, :],
                  cy + radius * np.cos(th)[:, None] * np.ones_like(ph)[None, :],
                  cz + radius * np.sin(th)[:, None] * np.sin(ph)[None, :]], -1)
    tris = []
    for j in range(nseg):
        tris.append([P[0, j], P[1, j + 1], P[1, j]])
        tris.append([P[nring, j], P[nring - 1, j], P[nring - 1, j + 1]])
    for i in range(1, nring - 1):
        for j in range(nseg):
            a, b, c, d = P[i, j], P[i, j + 1], P[i + 1, j + 1], P[i + 1, j]
            tris.append([a, b, c])
            tris.append([a, c, d])
    return np.asarray(tris, np.float64)


def cornell_box(width=256, height=256, seed=2013) -> SceneData:
    """Cornell box in world units 0..555: 5 wall quads + 2 boxes (34 triangles)
    + a 1-quad area light (2 triangles, radiance (17, 12, 4)); Lambertian kd
    ~ U[0.2, 0.8] per channel per surface."""
    rng = np.random.default_rng(seed)
    b = _Builder()
    L = 555.0
    walls = [
        [(0, 0, 0), (L, 0, 0), (L, 0, L), (0, 0, L)],        # floor
        [(0, L, 0), (0, L, L), (L, L, L), (L, L, 0)],        # ceiling
        [(0, 0, L), (L, 0, L), (L, L, L), (0, L, L)],        # back
        [(L, 0, 0), (L, L, 0), (L, L, L), (L, 0, L)],        # left (x = 555)
        [(0, 0, 0), (0, 0, L), (0, L, L), (0, L, 0)],        # right (x = 0)
    ]
    for w in walls:
        m = b.material(LAMBERT, rng.uniform(0.2, 0.8, 3))
        b.quads(b.surface(m), [w])
    for center, size, ang in [((185.5, 82.5, 169.0), (165, 165, 165), -0.314),
                              ((368.5, 165.0, 351.0), (165, 330, 165), 0.262)]:
        m = b.material(LAMBERT, rng.uniform(0.2, 0.8, 3))
        b.quads(b.surface(m), _box_quads(center, size, ang))
    b.lights.append((17.0, 12.0, 4.0))
    y = 554.0
    b.quads(b.surface(-1, 0), [[(213, y, 227), (343, y, 227), (343, y, 332), (213, y, 332)]])
    cam = camera((278, 273, -800), (278, 273, 0), (0, 1, 0), 39.3077, width, height)
    return b.build(cam, "cornell")


def room(width=512, height=512, seed=2013, target_tris=200_000) -> SceneData:
    """A 10 x 10 x 3 m room at scale x100 (1000 x 300 x 1000 units) filled with
    seeded tessellated boxes and UV spheres up to ~target_tris triangles.  Half
    the objects are Lambertian, half modified Phong (exponent log-U[10, 200],
    ks ~ U[0.1, 0.4]); a 200 x 200 ceiling area light."""
    rng = np.random.default_rng(seed)
    b = _Builder()
    X, Y, Z = 1000.0, 300.0, 1000.0
    walls = [
        [(0, 0, 0), (X, 0, 0), (X, 0, Z), (0, 0, Z)],
        [(0, Y, 0), (0, Y, Z), (X, Y, Z), (X, Y, 0)],
        [(0, 0, Z), (X, 0, Z), (X, Y, Z), (0, Y, Z)],
        [(0, 0, 0), (0, Y, 0), (X, Y, 0), (X, 0, 0)],
        [(X, 0, 0), (X, Y, 0), (X, Y, Z), (X, 0, Z)],
        [(0, 0, 0), (0, 0, Z), (0, Y, Z), (0, Y, 0)],
    ]
    for w in walls:
        m = b.material(LAMBERT, rng.uniform(0.3, 0.7, 3))
        b.quads(b.surface(m), [w])

    def obj_material():
        kd = rng.uniform(0.2, 0.6, 3)
        if rng.random() < 0.5:
            return b.material(LAMBERT, kd)
        ks = rng.uniform(0.1, 0.4)
        n = math.exp(rng.uniform(math.log(10.0), math.log(200.0)))
        return b.material(PHONG, kd, (ks, ks, ks), n)

    count = 14
    sph_tris = 2 * 32 * 16  # nseg=32, nring=17 -> 2*32 caps + 2*32*15 band
    box_tris = 12 * 10 * 10
    while count < target_tris - 2:
        remaining = target_tris - 2 - count
        if rng.random() < 0.66 and remaining >= sph_tris:
            r = rng.uniform(15.0, 60.0)
            c = (rng.uniform(80, X - 80), r, rng.uniform(150, Z - 80))
            b.triangles(b.surface(obj_material()), _uv_sphere(c, r, 32, 17))
            count += sph_tris
        elif remaining >= box_tris:
            s = rng.uniform(30.0, 120.0, 3)
            s[1] = rng.uniform(30.0, 200.0)
            c = (rng.uniform(80, X - 80), s[1] / 2.0, rng.uniform(150, Z - 80))
            b.quads(b.surface(obj_material()), _box_quads(c, s, rng.uniform(0, math.pi), grid=10))
            count += box_tris
        else:
            break
    b.lights.append((30.0, 30.0, 30.0))
    y = Y - 1.0
    b.quads(b.surface(-1, 0), [[(400, y, 400), (600, y, 400), (600, y, 600), (400, y, 600)]])
    cam = camera((500, 160, 20), (500, 110, 700), (0, 1, 0), 60.0, width, height)
    return b.build(cam, "room")


def instanced_meshes(width=1280, height=720, seed=2013, target_tris=1_000_000) -> SceneData:
    """~target_tris triangles of instanced meshes (sphere and box prototypes)
    flattened to world space in the room, Ward BRDFs with isotropic alpha ~
    U[0.05, 0.5]."""
    rng = np.random.default_rng(seed)
    b = _Builder()
    X, Y, Z = 2000.0, 600.0, 2000.0
    walls = [
        [(0, 0, 0), (X, 0, 0), (X, 0, Z), (0, 0, Z)],
        [(0, Y, 0), (0, Y, Z), (X, Y, Z), (X, Y, 0)],
        [(0, 0, Z), (X, 0, Z), (X, Y, Z), (0, Y, Z)],
        [(0, 0, 0), (0, Y, 0), (X, Y, 0), (X, 0, 0)],
        [(X, 0, 0), (X, Y, 0), (X, Y, Z), (X, 0, Z)],
        [(0, 0, 0), (0, 0, Z), (0, Y, Z), (0, Y, 0)],
    ]
    for w in walls:
        m = b.material(LAMBERT, rng.uniform(0.3, 0.7, 3))
        b.quads(b.surface(m), [w])
    proto_sphere = _uv_sphere((0, 0, 0), 1.0, 48, 25)
    proto_box = _box_quads((0, 0, 0), (2, 2, 2), 0.0, grid=12)
    count = 14
    while count < target_tris - 2:
        kd = rng.uniform(0.2, 0.6, 3)
        ks = rng.uniform(0.1, 0.4)
        a = rng.uniform(0.05, 0.5)
        m = b.material(WARD, kd, (ks, ks, ks), a, a)
        s = b.surface(m)
        scale = rng.uniform(20.0, 80.0)
        ang = rng.uniform(0, 2 * math.pi)
        ca, sa = math.cos(ang), math.sin(ang)
        rot = np.array([[ca, 0, sa], [0, 1, 0], [-sa, 0, ca]])
        c = np.array([rng.uniform(100, X - 100), scale, rng.uniform(200, Z - 100)])
        if rng.random() < 0.5:
            t = proto_sphere @ rot.T * scale + c
            if count + len(t) > target_tris - 2:
                break
            b.triangles(s, t)
            count += len(t)
        else:
            q = proto_box @ rot.T * scale + c
            if count + 2 * len(q) > target_tris - 2:
                break
            b.quads(s, q)
            count += 2 * len(q)
    b.lights.append((40.0, 40.0, 40.0))
    y = Y - 1.0
    b.quads(b.surface(-1, 0), [[(800, y, 800), (1200, y, 800), (1200, y, 1200), (800, y, 1200)]])
    cam = camera((1000, 300, 30), (1000, 200, 1200), (0, 1, 0), 60.0, width, height)
    return b.build(cam, "instanced")


def pixel_grid(width, height):
    """(px, py) of every pixel, row-major (py = row i, px = column j)."""
    py, px = np.divmod(np.arange(width * height, dtype=np.int32), width)
    return px.astype(np.int32), py.astype(np.int32)
