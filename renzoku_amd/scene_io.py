"""Scene files: the reference's JSON scene over OBJ meshes, read into (and
written from) the SceneData arrays the engine takes.

Read side, following the reference's importers:
  * OBJ (obj_parser.cpp:396-560, obj_importer.cc:296-420): `v` positions
    (atof, stored as float), `f` faces of 3 or 4 vertices (`i`, `i/t`, `i//n`,
    `i/t/n`; negative indices count back from the current vertex list; a face
    with more than 4 vertices keeps its first 4, one with fewer than 3 is
    skipped), `o` object names, `usemtl` names.  Every face is a shape of its
    own: a Triangle, or a Quad split as (p0, p1, p2), (p0, p2, p3) with both
    halves sharing the tangent unit(p0 - p1) (quad.cc:5-27, triangle.cc:17).
    A shape is named `<geometry key>.<object name>` (just the key outside any
    object, json_importer.cc:137-153).
  * JSON (json_importer.cc:155-898): `geometry` (key -> OBJ file under
    `<scene dir>/mesh/`), `materials` (lambertian kd; phong kd ks n; ward kd
    ks ax ay), `light_specs` (area radiance), `cameras` (pinhole pos lookat up
    near far image film, focal or fov), `scene.world.surfaces` (mesh + bsdf
    per key, `visible: false` skips; a shape bound again takes the new
    material in place), `scene.world.lights` (mesh + area light spec: every
    piece becomes an emitting surface, appended after the surfaces),
    `scene.views` (the first view's camera), `scene.options.geometry.
    two_sided_surfaces`.  get_shape matches a name exactly or as a prefix
    followed by '.' (json_importer.cc:98-111).

Not read (the path has no use for them): vertex normals and smoothing groups
(the engine shades with the geometric normal), texture coordinates, spheres
(the reference's BVH skips them), glass / mirror materials, point / spot /
directional lights, OBJ-declared surfaces and MTL area lights.  These raise
ValueError rather than loading a different scene silently.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

from .scenes import LAMBERT, PHONG, TWO_SIDED, WARD, SceneData, _Builder, cross, unit

F = np.float32


def _atoi(tok):
    """C atoi of the leading integer of tok ('' -> 0)"""
    j = 1 if tok[:1] in "+-" else 0
    while j < len(tok) and tok[j].isdigit():
        j += 1
    return int(tok[:j]) if j > (1 if tok[:1] in "+-" else 0) else 0


def parse_obj(text):
    """-> (vertices (N, 3) f32, faces [(vertex indices, object index, material name)], object names)"""
    verts, faces, names = [], [], []
    cur_obj, cur_mtl = -1, None
    for line in text.splitlines():
        tok = line.split()
        if not tok or tok[0].startswith("#"):
            continue
        k = tok[0]
        if k == "v":
            verts.append([float(tok[1]), float(tok[2]), float(tok[3])])
        elif k == "f":
            idx = []
            for t in tok[1:5]:                      # MAX_VERTEX_COUNT = 4
                i = _atoi(t)
                idx.append(-1 if i == 0 else (len(verts) + i if i < 0 else i - 1))
            faces.append((tuple(idx), cur_obj, cur_mtl))
        elif k == "o":
            names.append(tok[1] if len(tok) > 1 else "")
            cur_obj = len(names) - 1
        elif k == "usemtl":
            cur_mtl = tok[1] if len(tok) > 1 else None
    return np.asarray(verts, F).reshape(-1, 3), faces, names


def _pinhole(spec):
    """Camera::set_perspective + Sensor (camera.cc:23-31, camera.h:228-230, sensor.h:170-180)"""
    pos = np.asarray(spec["pos"], F)
    lookat = np.asarray(spec["lookat"], F)
    up = np.asarray(spec["up"], F)
    img = [F(x) for x in spec["image"]]
    film = [F(x) for x in spec["film"]]
    if "focal" in spec:
        focal = F(spec["focal"])
    elif "fov" in spec:
        fov = F(spec["fov"])
        focal = F(film[1] / (F(2.0) * F(math.tan(F(fov / F(2)) * F(math.pi / 180.0)))))
    else:
        raise ValueError("pinhole camera without focal or fov")
    nvec = pos - lookat
    nn = unit(nvec)
    c = unit(cross(up, nn))
    vv = cross(nn, c)
    cam = np.zeros(19, F)
    cam[0:3], cam[3:6], cam[6:9], cam[9:12] = pos, c, vv, nn
    cam[12], cam[13] = img[0] / F(2.0), img[1] / F(2.0)
    cam[14], cam[15] = img[0] / film[0], img[1] / film[1]
    cam[16], cam[17], cam[18] = focal, img[0], img[1]
    return cam


def load_json_scene(path) -> SceneData:
    with open(path) as f:
        d = json.load(f)
    scene_dir = os.path.dirname(os.path.abspath(path))
    shapes = []                                     # [name, points (k, 3)]
    for key, v in d.get("geometry", {}).items():
        if not isinstance(v, str):
            raise ValueError(f"geometry '{key}': only OBJ files are read (spheres are not traced)")
        with open(os.path.join(scene_dir, "mesh", v)) as f:
            verts, faces, names = parse_obj(f.read())
        for idx, obj, _ in faces:
            n = len(idx)
            if n < 3 or min(idx) < 0:
                continue                            # polygons of fewer than 3 vertices are skipped
            name = key + "." + names[obj] if obj >= 0 else key
            shapes.append([name, verts[list(idx)]])
    mats, mat_index = [], {}
    for key, m in d.get("materials", {}).items():
        t = m.get("type")
        row = np.zeros(12, F)
        if t == "lambertian":
            row[0] = LAMBERT
            row[1:4] = m["kd"]
        elif t == "phong":
            row[0] = PHONG
            row[1:4], row[4:7], row[7] = m["kd"], m["ks"], m["n"]
        elif t == "ward":
            row[0] = WARD
            row[1:4], row[4:7], row[7], row[8] = m["kd"], m["ks"], m["ax"], m["ay"]
        else:
            raise ValueError(f"material '{key}': type '{t}' is not on the path")
        mat_index[key] = len(mats)
        mats.append(row)
    specs = {}
    for key, s in d.get("light_specs", {}).items():
        if s.get("type") != "area":
            raise ValueError(f"light spec '{key}': only area lights are on the path")
        specs[key] = np.asarray(s["radiance"], F)
    cams = {k: _pinhole(c) for k, c in d.get("cameras", {}).items() if c.get("type") == "pinhole"}
    scene = d["scene"]
    world = scene["world"]

    def get_shape(name):
        return [i for i, (nm, _) in enumerate(shapes) if nm == name or nm.startswith(name + ".")]

    surf = []                                       # [shape index, material, light]
    bound = {}
    for key, s in world.get("surfaces", {}).items():
        if isinstance(s, str):
            raise ValueError(f"surface '{key}': OBJ-declared surfaces are not read")
        if s.get("visible", True) is False:
            continue
        m = mat_index[s["bsdf"]]
        for i in get_shape(s["mesh"]):
            if i in bound:
                surf[bound[i]][1] = m
            else:
                bound[i] = len(surf)
                surf.append([i, m, -1])
    radiance = []
    for key, s in world.get("lights", {}).items():
        if isinstance(s, str):
            raise ValueError(f"light '{key}': only area lights bound to a mesh are on the path")
        rad = specs[s["light"]]
        for i in get_shape(s["mesh"]):
            surf.append([i, -1, len(radiance)])
            radiance.append(rad)
    two = bool(scene.get("options", {}).get("geometry", {}).get("two_sided_surfaces", False))
    views = scene.get("views", {})
    view = next(iter(views.values())) if views else None
    cam = cams[view["camera"]] if view and "camera" in view else next(iter(cams.values()))
    b = _Builder()
    b.mats = mats
    b.lights = radiance
    for i, m, lt in surf:
        sid = b.surface(m, lt, TWO_SIDED if two else 0)
        pts = shapes[i][1]
        if len(pts) == 4:
            b.quads(sid, pts[None])
        else:
            b.triangles(sid, pts[None])
    return b.build(cam, os.path.splitext(os.path.basename(path))[0])


def _num(x):
    return float(np.float32(x))                     # the shortest decimal that reads back as this float


def save_json_scene(sc: SceneData, directory, name="scene"):
    """Write sc as <directory>/<name>.json + <directory>/mesh/<name>.obj.  Each
    surface is an OBJ object; two consecutive triangles (p0, p1, p2), (p0, p2,
    p3) of a surface sharing the tangent unit(p0 - p1) are written as one quad,
    so reading back gives the same triangles and tangents.  Emitting surfaces
    come after the others on reading (the reference appends lights last)."""
    os.makedirs(os.path.join(directory, "mesh"), exist_ok=True)
    pos = np.asarray(sc.positions, F).reshape(-1, 3, 3)
    tan = np.asarray(sc.tangents, F)
    ts = np.asarray(sc.tri_surface)
    lines, nv = [], 0
    for s in range(len(sc.surf_material)):
        lines.append(f"o s{s}")
        tris = np.nonzero(ts == s)[0]
        k = 0
        while k < len(tris):
            a = tris[k]
            p = pos[a]
            if k + 1 < len(tris):
                q = pos[tris[k + 1]]
                quad = (np.array_equal(q[0], p[0]) and np.array_equal(q[1], p[2])
                        and np.array_equal(tan[tris[k + 1]], tan[a]) and np.array_equal(tan[a], unit(p[0] - p[1])))
                if quad:
                    for v in (p[0], p[1], p[2], q[2]):
                        lines.append("v " + " ".join(repr(_num(x)) for x in v))
                    lines.append(f"f {nv + 1} {nv + 2} {nv + 3} {nv + 4}")
                    nv += 4
                    k += 2
                    continue
            for v in p:
                lines.append("v " + " ".join(repr(_num(x)) for x in v))
            lines.append(f"f {nv + 1} {nv + 2} {nv + 3}")
            nv += 3
            k += 1
    with open(os.path.join(directory, "mesh", name + ".obj"), "w") as f:
        f.write("\n".join(lines) + "\n")
    mats = {}
    for i, m in enumerate(np.asarray(sc.materials, F).reshape(-1, 12)):
        t = int(m[0])
        if t == LAMBERT:
            mats[f"m{i}"] = {"type": "lambertian", "kd": [_num(x) for x in m[1:4]]}
        elif t == PHONG:
            mats[f"m{i}"] = {"type": "phong", "kd": [_num(x) for x in m[1:4]], "ks": [_num(x) for x in m[4:7]],
                             "n": _num(m[7])}
        else:
            mats[f"m{i}"] = {"type": "ward", "kd": [_num(x) for x in m[1:4]], "ks": [_num(x) for x in m[4:7]],
                             "ax": _num(m[7]), "ay": _num(m[8])}
    specs = {f"L{j}": {"type": "area", "radiance": [_num(x) for x in r]}
             for j, r in enumerate(np.asarray(sc.light_radiance, F).reshape(-1, 3))}
    surfaces, lights = {}, {}
    for s in range(len(sc.surf_material)):
        if sc.surf_light[s] >= 0:
            lights[f"light{s}"] = {"light": f"L{int(sc.surf_light[s])}", "mesh": f"{name}.s{s}"}
        elif sc.surf_material[s] >= 0:
            surfaces[f"surf{s}"] = {"mesh": f"{name}.s{s}", "bsdf": f"m{int(sc.surf_material[s])}"}
    cam = np.asarray(sc.camera, F)
    eye, n = cam[0:3], cam[9:12]
    # a look-at point far along -n keeps eye - lookat's direction to float precision
    far = F(1024.0 * max(1.0, float(np.abs(eye).max())))
    w, h = float(cam[17]), float(cam[18])
    cameras = {"cam": {"type": "pinhole", "pos": [_num(x) for x in eye], "lookat": [_num(x) for x in eye - far * n],
                       "up": [_num(x) for x in cam[6:9]], "near": 1e-3, "far": 1e6, "image": [w, h],
                       "film": [_num(w / cam[14]), _num(h / cam[15])], "focal": _num(cam[16])}}
    two = bool(len(sc.surf_flags) and (np.asarray(sc.surf_flags) & TWO_SIDED).all())
    doc = {"geometry": {name: name + ".obj"}, "materials": mats, "light_specs": specs, "cameras": cameras,
           "scene": {"world": {"surfaces": surfaces, "lights": lights},
                     "views": {"view0": {"camera": "cam"}},
                     "options": {"geometry": {"traversal": "bvh", "two_sided_surfaces": two}}}}
    path = os.path.join(directory, name + ".json")
    with open(path, "w") as f:
        json.dump(doc, f, indent=1)
    return path
