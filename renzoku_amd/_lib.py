"""ctypes binding of librzk.so (the C ABI in include/rzk.h).

The library is the product: there is no CPU fallback.  Loading fails loudly
when librzk.so has not been built, and every compute call raises RzkError when
the GPU path fails.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librzk.so")

RZK_OK = 0
RAY_FLOATS, RECV_FLOATS, VPL_FLOATS, MAT_FLOATS, CAM_FLOATS = 6, 21, 21, 12, 19
CLAMP_NONE, CLAMP_FORM_FACTOR, CLAMP_REFLECTIVITY, CLAMP_RADIANCE = 0, 1, 2, 3
SURF_FLIPPED, SURF_TWO_SIDED = 1, 2

# every symbol include/rzk.h declares
EXPORTS = [
    "rzk_ctx_create", "rzk_ctx_destroy", "rzk_ctx_last_error", "rzk_ctx_stream", "rzk_ctx_synchronize",
    "rzk_device_alloc", "rzk_device_free", "rzk_copy_h2d", "rzk_copy_d2h",
    "rzk_scene_create", "rzk_scene_create_ex", "rzk_scene_destroy", "rzk_scene_get_info", "rzk_scene_export_bvh",
    "rzk_scene_export_trees",
    "rzk_trace_closest", "rzk_trace_any", "rzk_trace_segments",
    "rzk_trace_closest_dev", "rzk_trace_any_dev", "rzk_trace_segments_dev", "rzk_trace_stats",
    "rzk_fill_receivers", "rzk_camera_primary",
    "rzk_vpls_create", "rzk_vpls_destroy", "rzk_generate_vpls",
    "rzk_gather", "rzk_gather_dev_noalloc", "rzk_visibility", "rzk_render_vpl",
    "rzk_reduced_matrix", "rzk_transport_matrix", "rzk_alphas", "rzk_cluster_assign", "rzk_refine_slices",
    "rzk_lightslice_create", "rzk_lightslice_destroy", "rzk_lightslice_initialize",
    "rzk_lightslice_gather", "rzk_lightslice_get_info", "rzk_lightslice_count_work", "rzk_lightslice_get",
    "rzk_lightslice_render_guided",
    "rzk_lightslice_guide", "rzk_guide_destroy",
    "rzk_batch_create", "rzk_batch_destroy", "rzk_batch_set_config", "rzk_batch_get_config", "rzk_batch_render",
    "rzk_batch_autotune",
    "rzk_lightslice_clusters", "rzk_clusters_destroy", "rzk_guide_create", "rzk_guide_get", "rzk_render_guided", "rzk_render_guided_rays",
    "rzk_server_create", "rzk_server_destroy", "rzk_server_submit", "rzk_server_wait", "rzk_server_get_stats",
    "rzk_server_last_error",
    "rzk_kdslices_build", "rzk_kdslices_build_mt", "rzk_kdslices_build_dev", "rzk_kdslices_destroy", "rzk_kdslices_info", "rzk_kdslices_get", "rzk_kdslices_find",
    "rzk_kdslices_nearest", "rzk_map_vec6",
]


class RzkError(RuntimeError):
    pass


class SceneDesc(C.Structure):
    _fields_ = [
        ("num_triangles", C.c_int32), ("positions", C.c_void_p), ("tangents", C.c_void_p),
        ("tri_surface", C.c_void_p), ("num_surfaces", C.c_int32), ("surf_material", C.c_void_p),
        ("surf_light", C.c_void_p), ("surf_flags", C.c_void_p), ("num_materials", C.c_int32),
        ("materials", C.c_void_p), ("num_lights", C.c_int32), ("light_radiance", C.c_void_p),
    ]


class SceneInfo(C.Structure):
    _fields_ = [
        ("num_triangles", C.c_int32), ("num_interior_nodes", C.c_int32), ("depth", C.c_int32),
        ("box", C.c_float * 6), ("device_bytes", C.c_int64),
    ]


class LightSliceParams(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32), ("tile", C.c_int32), ("num_clusters", C.c_int32),
        ("num_neighbor_slices", C.c_int32), ("map_resolution", C.c_int32), ("map_epsilon", C.c_float),
        ("clamp_mode", C.c_int32), ("clamp", C.c_float), ("incident_clamp", C.c_float), ("tmin", C.c_float),
        ("n_emitted", C.c_float), ("seed", C.c_uint32), ("map_gate", C.c_int32), ("max_slice_clusters", C.c_int32),
        ("kd_max_slice_size", C.c_int32), ("kd_pivots", C.c_int32),
    ]


class PtParams(C.Structure):
    _fields_ = [
        ("spp", C.c_int32), ("max_bounce", C.c_int32), ("direct_lighting", C.c_int32), ("guide_mode", C.c_int32),
        ("seed", C.c_uint32), ("tmin", C.c_float), ("x0", C.c_int32), ("y0", C.c_int32), ("w", C.c_int32),
        ("h", C.c_int32),
    ]


class RenderJob(C.Structure):
    _fields_ = [
        ("guide", C.c_void_p), ("x0", C.c_int32), ("y0", C.c_int32), ("w", C.c_int32), ("h", C.c_int32),
        ("spp", C.c_int32), ("max_bounce", C.c_int32), ("direct_lighting", C.c_int32), ("guide_mode", C.c_int32),
        ("seed", C.c_uint32), ("tmin", C.c_float),
    ]


class BatchConfig(C.Structure):
    _fields_ = [("block", C.c_int32), ("blocks_per_cu", C.c_int32), ("items_per_dequeue", C.c_int32),
                ("samples_per_item", C.c_int32), ("engine", C.c_int32), ("path_slots", C.c_int32)]
    DEFAULTS = dict(block=256, blocks_per_cu=8, items_per_dequeue=1, samples_per_item=1, engine=1,
                    path_slots=1 << 20)

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchTrial(C.Structure):
    _fields_ = [("cfg", BatchConfig), ("ms", C.c_float)]


class GuideParams(C.Structure):
    _fields_ = [("tile", C.c_int32), ("map_resolution", C.c_int32), ("map_epsilon", C.c_float),
                ("clamp_mode", C.c_int32), ("incident_clamp", C.c_float), ("tmin", C.c_float), ("map_gate", C.c_int32)]


class ServerParams(C.Structure):
    _fields_ = [("memory_budget", C.c_int64), ("step_iterations", C.c_int32), ("launch", BatchConfig),
                ("use_cluster_params", C.c_int32), ("guide", GuideParams)]


class Request(C.Structure):
    _fields_ = [("camera", C.c_float * 19), ("spp", C.c_int32), ("max_bounce", C.c_int32),
                ("direct_lighting", C.c_int32), ("guide_mode", C.c_int32), ("seed", C.c_uint32), ("tmin", C.c_float)]


class RequestStats(C.Structure):
    _fields_ = [("submitted", C.c_double), ("admitted", C.c_double), ("finished", C.c_double),
                ("rounds", C.c_int32), ("pad", C.c_int32), ("device_bytes", C.c_int64)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("finished", C.c_int64), ("rounds", C.c_int64), ("items", C.c_int64),
                ("peak_bytes", C.c_int64), ("busy_seconds", C.c_double), ("now", C.c_double)]


class LightSliceInfo(C.Structure):
    _fields_ = [
        ("num_slices", C.c_int32), ("num_centers", C.c_int32), ("num_clusters", C.c_int32),
        ("draws", C.c_int32), ("reduced_pairs", C.c_int64), ("gather_pairs", C.c_int64), ("refine_passes", C.c_int32),
    ]


_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RzkError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(LIB_PATH)
    P, I32, I64, F = C.c_void_p, C.c_int32, C.c_int64, C.c_float
    sig = {
        "rzk_ctx_create": [C.c_int, C.POINTER(P)],
        "rzk_ctx_destroy": [P],
        "rzk_ctx_last_error": [P],
        "rzk_ctx_stream": [P],
        "rzk_ctx_synchronize": [P],
        "rzk_device_alloc": [P, I64, C.POINTER(P)],
        "rzk_device_free": [P, P],
        "rzk_copy_h2d": [P, P, P, I64],
        "rzk_copy_d2h": [P, P, P, I64],
        "rzk_scene_create": [P, C.POINTER(SceneDesc), C.POINTER(P)],
        "rzk_scene_destroy": [P],
        "rzk_scene_get_info": [P, C.POINTER(SceneInfo)],
        "rzk_scene_export_bvh": [P, P, P, P, P, P, C.POINTER(I32)],
        "rzk_scene_create_ex": [P, C.POINTER(SceneDesc), I32, C.POINTER(P)],
        "rzk_scene_export_trees": [P, P, P, P, P, C.POINTER(I32)],
        "rzk_trace_closest": [P, P, P, I64, F, F, P],
        "rzk_trace_any": [P, P, P, I64, F, F, P],
        "rzk_trace_segments": [P, P, P, P, I64, F, P],
        "rzk_trace_closest_dev": [P, P, P, I64, F, F, P, P],
        "rzk_trace_any_dev": [P, P, P, I64, F, F, P, P],
        "rzk_trace_segments_dev": [P, P, P, P, I64, F, P, P],
        "rzk_trace_stats": [P, P, C.c_int, P, I64, F, F, P],
        "rzk_fill_receivers": [P, P, P, P, I64, P, P, P],
        "rzk_camera_primary": [P, P, P, I64, P, P, P, P, P, P, P],
        "rzk_vpls_create": [P, P, P, I64, C.POINTER(P)],
        "rzk_generate_vpls": [P, P, I64, C.c_uint32, I64, F, P, P, C.POINTER(I64), P],
        "rzk_vpls_destroy": [P],
        "rzk_kdslices_build": [P, P, C.c_int32, C.c_int32, C.c_uint32, P],
        "rzk_kdslices_build_dev": [P, P, C.c_int32, C.c_int32, C.c_uint32, P],
        "rzk_kdslices_build_mt": [P, P, C.c_int32, C.c_int32, C.c_uint32, P, P],
        "rzk_kdslices_destroy": [P],
        "rzk_kdslices_info": [P, P, P],
        "rzk_kdslices_get": [P, C.c_char_p, P, I64],
        "rzk_kdslices_find": [P, P, C.c_int32, P],
        "rzk_kdslices_nearest": [P, P, C.c_int32, C.c_int32, P],
        "rzk_map_vec6": [P, P, P, P, I64, F, P],
        "rzk_gather": [P, P, P, P, P, I64, C.c_int, F, F, F, P],
        "rzk_gather_dev_noalloc": [P, P, P, P, P, I64, C.c_int, F, F, F, P],
        "rzk_visibility": [P, P, P, P, P, I64, F, P],
        "rzk_render_vpl": [P, P, P, P, C.c_int, F, F, F, P],
        "rzk_reduced_matrix": [P, P, P, P, P, I32, C.c_int, F, F, P, P],
        "rzk_transport_matrix": [P, P, P, P, P, I32, C.c_int, C.c_int, F, F, P, P],
        "rzk_alphas": [P, I32, I64, P, P, P],
        "rzk_cluster_assign": [P, I32, I64, P, I32, P, P, P],
        "rzk_refine_slices": [P, I32, I64, P, I32, P, I32, P, P, P, P, I32, P, I64, P, P, P],
        "rzk_lightslice_create": [P, P, P, C.POINTER(LightSliceParams), C.POINTER(P)],
        "rzk_lightslice_destroy": [P],
        "rzk_lightslice_initialize": [P, P],
        "rzk_lightslice_gather": [P, P],
        "rzk_lightslice_get_info": [P, C.POINTER(LightSliceInfo)],
        "rzk_lightslice_count_work": [P, P],
        "rzk_lightslice_render_guided": [P, P, C.POINTER(PtParams), P, P],
        "rzk_lightslice_get": [P, C.c_char_p, P, I64],
        "rzk_lightslice_guide": [P, P, C.POINTER(P)],
        "rzk_guide_destroy": [P],
        "rzk_batch_create": [P, P, C.POINTER(BatchConfig), C.POINTER(P)],
        "rzk_batch_destroy": [P],
        "rzk_batch_set_config": [P, C.POINTER(BatchConfig)],
        "rzk_batch_get_config": [P, C.POINTER(BatchConfig)],
        "rzk_batch_render": [P, C.POINTER(RenderJob), I32, P, P, C.POINTER(C.c_float)],
        "rzk_batch_autotune": [P, C.POINTER(RenderJob), I32, C.POINTER(BatchConfig), C.POINTER(BatchTrial), I32,
                               C.POINTER(I32)],
        "rzk_lightslice_clusters": [P, C.POINTER(P)],
        "rzk_clusters_destroy": [P],
        "rzk_guide_create": [P, P, P, P, C.POINTER(GuideParams), C.POINTER(P)],
        "rzk_guide_get": [P, C.c_char_p, P, I64],
        "rzk_render_guided": [P, P, C.POINTER(PtParams), P, P],
        "rzk_render_guided_rays": [P, P, C.POINTER(PtParams), P, P, P, I64, P],
        "rzk_server_create": [C.c_int, P, P, C.POINTER(ServerParams), C.POINTER(P)],
        "rzk_server_destroy": [P],
        "rzk_server_submit": [P, C.POINTER(Request), C.POINTER(I64)],
        "rzk_server_wait": [P, I64, C.c_double, P, C.POINTER(RequestStats)],
        "rzk_server_get_stats": [P, C.POINTER(ServerStats)],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.rzk_ctx_last_error.restype = C.c_char_p
    lib.rzk_server_last_error.argtypes = [P]
    lib.rzk_server_last_error.restype = C.c_char_p
    lib.rzk_ctx_stream.restype = C.c_void_p
    _lib = lib
    return lib


def ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays passed to librzk must be C-contiguous"
    return a.ctypes.data_as(C.c_void_p)


def f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


class Ctx:
    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.rzk_ctx_create(device, C.byref(h))
        if rc != RZK_OK:
            raise RzkError(f"rzk_ctx_create({device}) failed with status {rc} (no usable GPU?)")
        self.h = h

    def check(self, rc: int, what: str):
        if rc != RZK_OK:
            msg = self.lib.rzk_ctx_last_error(self.h)
            raise RzkError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    @property
    def stream(self) -> int:
        return self.lib.rzk_ctx_stream(self.h)

    def synchronize(self):
        self.check(self.lib.rzk_ctx_synchronize(self.h), "rzk_ctx_synchronize")

    def close(self):
        if getattr(self, "h", None):
            self.lib.rzk_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


SCENE_BUILD_DEVICE = 1


class Scene:
    """The Aggregate: scene + Morton LBVH on the device.  build="device" builds
    both trees on the GPU (rzk_scene_create_ex with RZK_SCENE_BUILD_DEVICE)."""

    def __init__(self, ctx: Ctx, sc, build="host"):
        if build not in ("host", "device"):
            raise ValueError(f"build must be 'host' or 'device', not {build!r}")
        self.ctx, self.lib = ctx, ctx.lib
        self._keep = [f32(sc.positions), f32(sc.tangents), i32(sc.tri_surface), i32(sc.surf_material),
                      i32(sc.surf_light), i32(sc.surf_flags), f32(sc.materials), f32(sc.light_radiance)]
        pos, tan, ts, sm, sl, sf, mat, lr = self._keep
        d = SceneDesc(len(ts), ptr(pos), ptr(tan), ptr(ts), len(sm), ptr(sm), ptr(sl), ptr(sf),
                      mat.shape[0], ptr(mat), lr.shape[0], ptr(lr))
        h = C.c_void_p()
        flags = SCENE_BUILD_DEVICE if build == "device" else 0
        ctx.check(self.lib.rzk_scene_create_ex(ctx.h, C.byref(d), flags, C.byref(h)), "rzk_scene_create_ex")
        self.h = h
        self.ntri = len(ts)

    def info(self) -> SceneInfo:
        inf = SceneInfo()
        self.lib.rzk_scene_get_info(self.h, C.byref(inf))
        return inf

    def export_bvh(self):
        n = self.ntri
        nn = max(2 * n - 1, 0)
        st, sm = np.zeros(n, np.int32), np.zeros(n, np.uint32)
        ch, rg, bx = np.zeros(2 * nn, np.int32), np.zeros(2 * nn, np.int32), np.zeros(6 * nn, np.float32)
        cnt = C.c_int32()
        self.ctx.check(self.lib.rzk_scene_export_bvh(self.h, ptr(st), ptr(sm), ptr(ch), ptr(rg), ptr(bx),
                                                     C.byref(cnt)), "rzk_scene_export_bvh")
        return st, sm, ch.reshape(-1, 2), rg.reshape(-1, 2), bx.reshape(-1, 6)

    def export_bvh_sorted(self):
        """(sorted position -> triangle, sorted Morton codes); also for a device-built scene"""
        st, sm = np.zeros(self.ntri, np.int32), np.zeros(self.ntri, np.uint32)
        cnt = C.c_int32()
        self.ctx.check(self.lib.rzk_scene_export_bvh(self.h, ptr(st), ptr(sm), None, None, None, C.byref(cnt)),
                       "rzk_scene_export_bvh")
        return st, sm

    def export_trees(self):
        """(binary nodes [num_interior, 16] f32, Morton-order triangles [ntri, 12] f32,
        4-wide nodes [n4, 32] f32, 4-wide triangles [ntri, 12] f32) as the walks read them"""
        inf = self.info()
        n4 = C.c_int32()
        self.ctx.check(self.lib.rzk_scene_export_trees(self.h, None, None, None, None, C.byref(n4)),
                       "rzk_scene_export_trees")
        nodes = np.zeros((inf.num_interior_nodes, 16), np.float32)
        tris = np.zeros((self.ntri, 12), np.float32)
        nodes4 = np.zeros((n4.value, 32), np.float32)
        tris4 = np.zeros((self.ntri, 12), np.float32)
        self.ctx.check(self.lib.rzk_scene_export_trees(self.h, ptr(nodes), ptr(tris), ptr(nodes4), ptr(tris4),
                                                       C.byref(n4)), "rzk_scene_export_trees")
        return nodes, tris, nodes4, tris4

    def trace_closest(self, rays, tmin=1e-3, tmax=1e6):
        rays = f32(rays).reshape(-1, 6)
        out = np.zeros((len(rays), 4), np.float32)
        self.ctx.check(self.lib.rzk_trace_closest(self.ctx.h, self.h, ptr(rays), len(rays), tmin, tmax, ptr(out)),
                       "rzk_trace_closest")
        return out[:, 0].copy(), out[:, 1].copy(), out[:, 2].copy(), out[:, 3].view(np.int32).copy()

    def trace_any(self, rays, tmin=1e-3, tmax=1e6):
        rays = f32(rays).reshape(-1, 6)
        out = np.zeros(len(rays), np.uint8)
        self.ctx.check(self.lib.rzk_trace_any(self.ctx.h, self.h, ptr(rays), len(rays), tmin, tmax, ptr(out)),
                       "rzk_trace_any")
        return out

    def trace_segments(self, p, q, tmin=1e-3):
        p, q = f32(p).reshape(-1, 3), f32(q).reshape(-1, 3)
        out = np.zeros(len(p), np.uint8)
        self.ctx.check(self.lib.rzk_trace_segments(self.ctx.h, self.h, ptr(p), ptr(q), len(p), tmin, ptr(out)),
                       "rzk_trace_segments")
        return out

    def trace_stats(self, rays, kind=0, tmin=1e-3, tmax=1e6):
        rays = f32(rays).reshape(-1, 6)
        out = np.zeros((len(rays), 2), np.int32)
        self.ctx.check(self.lib.rzk_trace_stats(self.ctx.h, self.h, kind, ptr(rays), len(rays), tmin, tmax, ptr(out)),
                       "rzk_trace_stats")
        return out

    def fill_receivers(self, rays, t, beta, gamma, tri):
        rays = f32(rays).reshape(-1, 6)
        n = len(rays)
        hits = np.zeros((n, 4), np.float32)
        hits[:, 0], hits[:, 1], hits[:, 2] = t, beta, gamma
        hits[:, 3] = np.asarray(tri, np.int32).view(np.float32)
        recv = np.zeros((n, RECV_FLOATS), np.float32)
        m, l = np.zeros(n, np.int32), np.zeros(n, np.int32)
        self.ctx.check(self.lib.rzk_fill_receivers(self.ctx.h, self.h, ptr(rays), ptr(hits), n, ptr(recv), ptr(m),
                                                   ptr(l)), "rzk_fill_receivers")
        return recv, m, l

    def camera_primary(self, cam, px, py):
        cam, px, py = f32(cam), i32(px), i32(py)
        n = len(px)
        rays = np.zeros((n, 6), np.float32)
        hits = np.zeros((n, 4), np.float32)
        recv = np.zeros((n, RECV_FLOATS), np.float32)
        m, l = np.zeros(n, np.int32), np.zeros(n, np.int32)
        self.ctx.check(self.lib.rzk_camera_primary(self.ctx.h, self.h, ptr(cam), n, ptr(px), ptr(py), ptr(rays),
                                                   ptr(hits), ptr(recv), ptr(m), ptr(l)), "rzk_camera_primary")
        return rays, hits, recv, m, l

    def close(self):
        if getattr(self, "h", None):
            self.lib.rzk_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Vpls:
    def __init__(self, ctx: Ctx, vpl, vmat):
        self.ctx, self.lib = ctx, ctx.lib
        self.vpl, self.vmat = f32(vpl).reshape(-1, VPL_FLOATS), i32(vmat)
        h = C.c_void_p()
        ctx.check(self.lib.rzk_vpls_create(ctx.h, ptr(self.vpl), ptr(self.vmat), len(self.vmat), C.byref(h)),
                  "rzk_vpls_create")
        self.h = h
        self.n = len(self.vmat)

    def close(self):
        if getattr(self, "h", None):
            self.lib.rzk_vpls_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class KdSlices:
    """rzk_kdslices: LightSlice's slicing of 6-D gather points by a kd-tree
    (max_slice_size points per slice at most); arrays order, slices (S x 2),
    boxes (S x 12) and the heap arrays axis / plane / cl of the query tree"""

    def __init__(self, ctx: Ctx, points, max_slice_size=256, seed=2013, n=None, pivots="philox"):
        """points: host array (n x 6), or a device address (int) of n x 6 floats.
        pivots="philox": per-node pivot numbers, the parallel device build;
        pivots="mt": one MT stream in the reference's recursion order, the host
        build (host points only); draws = the numbers it took"""
        self.ctx, self.lib = ctx, ctx.lib
        h = C.c_void_p()
        self.draws = 0
        if pivots not in ("philox", "mt"):
            raise ValueError(f"unknown pivot stream {pivots!r}")
        if pivots == "mt":
            if isinstance(points, int):
                raise ValueError("the MT-order build takes host points")
            pts = f32(points).reshape(-1, 6)
            self.n = len(pts)
            draws = C.c_int64(0)
            ctx.check(self.lib.rzk_kdslices_build_mt(ctx.h, ptr(pts), self.n, int(max_slice_size),
                                                     int(seed) & 0xffffffff, C.byref(draws), C.byref(h)),
                      "rzk_kdslices_build_mt")
            self.draws = draws.value
        elif isinstance(points, int):
            self.n = int(n)
            ctx.check(self.lib.rzk_kdslices_build_dev(ctx.h, C.c_void_p(points), self.n, int(max_slice_size),
                                                      int(seed) & 0xffffffff, C.byref(h)), "rzk_kdslices_build_dev")
        else:
            pts = f32(points).reshape(-1, 6)
            self.n = len(pts)
            ctx.check(self.lib.rzk_kdslices_build(ctx.h, ptr(pts), self.n, int(max_slice_size),
                                                  int(seed) & 0xffffffff, C.byref(h)), "rzk_kdslices_build")
        self.h = h
        S, H = C.c_int32(0), C.c_int32(0)
        ctx.check(self.lib.rzk_kdslices_info(self.h, C.byref(S), C.byref(H)), "rzk_kdslices_info")
        self.S, self.heap_used = S.value, H.value

    def get(self, name):
        shape, dt = {"order": ((self.n,), np.int32), "slices": ((self.S, 2), np.int32),
                     "boxes": ((self.S, 12), np.float32), "axis": ((self.heap_used,), np.int32),
                     "plane": ((self.heap_used,), np.float32), "cl": ((self.heap_used,), np.int32)}[name]
        out = np.zeros(shape, dt)
        self.ctx.check(self.lib.rzk_kdslices_get(self.h, name.encode(), ptr(out), out.nbytes), "rzk_kdslices_get")
        return out

    def find(self, q):
        q = f32(q).reshape(-1, 6)
        out = np.zeros(len(q), np.int32)
        self.ctx.check(self.lib.rzk_kdslices_find(self.h, ptr(q), len(q), ptr(out)), "rzk_kdslices_find")
        return out

    def nearest(self, q, k):
        q = f32(q).reshape(-1, 6)
        out = np.zeros((len(q), int(k)), np.int32)
        self.ctx.check(self.lib.rzk_kdslices_nearest(self.h, ptr(q), len(q), int(k), ptr(out)), "rzk_kdslices_nearest")
        return out

    def close(self):
        if getattr(self, "h", None):
            self.lib.rzk_kdslices_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def map_vec6(ctx, scene, recv, mats, diagonal):
    """rzk_map_vec6: LightSlice's 6-D point of each receiver record"""
    recv = f32(recv).reshape(-1, RECV_FLOATS)
    mats = i32(mats)
    out = np.zeros((len(recv), 6), np.float32)
    ctx.check(ctx.lib.rzk_map_vec6(ctx.h, scene.h, ptr(recv), ptr(mats), len(recv), float(diagonal), ptr(out)),
              "rzk_map_vec6")
    return out


def generate_vpls(ctx, scene, num_vpls, seed=2013, max_samples=1 << 26, tmin=1e-3, count_work=False):
    """rzk_generate_vpls: (vpl (num_vpls, 21) f32, vmat (num_vpls,) i32, light samples used
    [, work dict with closest_rays / nodes / tris])"""
    vpl = np.zeros((int(num_vpls), VPL_FLOATS), np.float32)
    vmat = np.zeros(int(num_vpls), np.int32)
    used = C.c_int64()
    work = np.zeros(3, np.int64) if count_work else None
    ctx.check(ctx.lib.rzk_generate_vpls(ctx.h, scene.h, int(num_vpls), int(seed) & 0xffffffff, int(max_samples),
                                        float(tmin), ptr(vpl), ptr(vmat), C.byref(used), ptr(work)),
              "rzk_generate_vpls")
    if count_work:
        return vpl, vmat, int(used.value), {"closest_rays": int(work[0]), "nodes": int(work[1]), "tris": int(work[2])}
    return vpl, vmat, int(used.value)


def gather(ctx, scene, vpls, recv, rmat, clamp_mode, clamp, tmin, n_emitted):
    recv, rmat = f32(recv).reshape(-1, RECV_FLOATS), i32(rmat)
    out = np.zeros((len(rmat), 3), np.float32)
    ctx.check(ctx.lib.rzk_gather(ctx.h, scene.h, vpls.h, ptr(recv), ptr(rmat), len(rmat), clamp_mode, clamp, tmin,
                                 n_emitted, ptr(out)), "rzk_gather")
    return out


def visibility(ctx, scene, vpls, recv, rmat, tmin):
    recv, rmat = f32(recv).reshape(-1, RECV_FLOATS), i32(rmat)
    ng = (len(rmat) + 63) // 64
    out = np.zeros(ng * vpls.n, np.uint64)
    ctx.check(ctx.lib.rzk_visibility(ctx.h, scene.h, vpls.h, ptr(recv), ptr(rmat), len(rmat), tmin, ptr(out)),
              "rzk_visibility")
    return out.reshape(vpls.n, ng)


def render_vpl(ctx, scene, vpls, cam, clamp_mode, clamp, tmin, n_emitted):
    cam = f32(cam)
    w, h = int(cam[17]), int(cam[18])
    out = np.zeros((h, w, 3), np.float32)
    ctx.check(ctx.lib.rzk_render_vpl(ctx.h, scene.h, vpls.h, ptr(cam), clamp_mode, clamp, tmin, n_emitted, ptr(out)),
              "rzk_render_vpl")
    return out


def reduced_matrix(ctx, scene, vpls, recv, rmat, clamp_mode, clamp, tmin):
    recv, rmat = f32(recv).reshape(-1, RECV_FLOATS), i32(rmat)
    rows = len(rmat)
    R = np.zeros((vpls.n, rows, 3), np.float32)
    V = np.zeros((vpls.n, rows), np.float32)
    ctx.check(ctx.lib.rzk_reduced_matrix(ctx.h, scene.h, vpls.h, ptr(recv), ptr(rmat), rows, clamp_mode, clamp, tmin,
                                         ptr(R), ptr(V)), "rzk_reduced_matrix")
    return R, V


TERM_RADIANCE, TERM_INCIDENT = 0, 1


def transport_matrix(ctx, scene, vpls, recv, rmat, clamp_mode, clamp, tmin, term=TERM_RADIANCE):
    """rzk_transport_matrix: (M (V, rows, 3) per pair: outgoing radiance with
    visibility, or incident radiance without; Vis (V, rows))"""
    recv, rmat = f32(recv).reshape(-1, RECV_FLOATS), i32(rmat)
    rows = len(rmat)
    M = np.zeros((vpls.n, rows, 3), np.float32)
    V = np.zeros((vpls.n, rows), np.float32)
    ctx.check(ctx.lib.rzk_transport_matrix(ctx.h, scene.h, vpls.h, ptr(recv), ptr(rmat), rows, int(term), clamp_mode,
                                           clamp, tmin, ptr(M), ptr(V)), "rzk_transport_matrix")
    return M, V


def alphas(ctx, R):
    R = f32(R)
    cols, rows = R.shape[0], R.shape[1]
    a, m = np.zeros(cols, np.float32), np.zeros(cols, np.float32)
    ctx.check(ctx.lib.rzk_alphas(ctx.h, rows, cols, ptr(R), ptr(a), ptr(m)), "rzk_alphas")
    return a, m


def cluster_assign(ctx, R, centers, weights):
    R, centers, weights = f32(R), i32(centers), f32(weights)
    cols, rows = R.shape[0], R.shape[1]
    out = np.zeros(cols, np.int32)
    ctx.check(ctx.lib.rzk_cluster_assign(ctx.h, rows, cols, ptr(R), len(centers), ptr(centers), ptr(weights),
                                         ptr(out)), "rzk_cluster_assign")
    return out


def refine_slices(ctx, R, nb, cl_off, cl_mem, cl_rep, cl_w, max_slice_clusters, draws):
    """rzk_refine_slices: per-slice refinement of the kept clusters.
    R: (V, S, 3); nb: (S, nnb); returns (count (S,), reps (S, M), weights (S, M))."""
    R, nb = f32(R), i32(nb)
    V, S = R.shape[0], R.shape[1]
    cl_off, cl_mem, cl_rep, cl_w, draws = i32(cl_off), i32(cl_mem), i32(cl_rep), f32(cl_w), f32(draws)
    M = int(max_slice_clusters)
    count, rep, w = np.zeros(S, np.int32), np.zeros((S, M), np.int32), np.zeros((S, M), np.float32)
    ctx.check(ctx.lib.rzk_refine_slices(ctx.h, S, V, ptr(R), nb.shape[1], ptr(nb), len(cl_rep), ptr(cl_off),
                                        ptr(cl_mem), ptr(cl_rep), ptr(cl_w), M, ptr(draws), len(draws), ptr(count),
                                        ptr(rep), ptr(w)), "rzk_refine_slices")
    return count, rep, w


class LightSlice:
    def __init__(self, ctx, scene, vpls, **kw):
        self.ctx, self.lib = ctx, ctx.lib
        p = LightSliceParams()
        defaults = dict(width=512, height=512, tile=32, num_clusters=300, num_neighbor_slices=16,
                        map_resolution=32, map_epsilon=1e-3, clamp_mode=CLAMP_FORM_FACTOR, clamp=0.01,
                        incident_clamp=0.01, tmin=1e-3, n_emitted=1.0, seed=2013, map_gate=0, max_slice_clusters=0,
                        kd_max_slice_size=0, kd_pivots=0)
        defaults.update(kw)
        for k, v in defaults.items():
            setattr(p, k, v)
        self.params = defaults
        h = C.c_void_p()
        ctx.check(self.lib.rzk_lightslice_create(ctx.h, scene.h, vpls.h, C.byref(p), C.byref(h)),
                  "rzk_lightslice_create")
        self.h = h
        # the native object points at both: they must outlive it
        self.scene, self.vpls = scene, vpls
        self.w, self.hgt = defaults["width"], defaults["height"]
        self.res = defaults["map_resolution"]

    def initialize(self, cam):
        cam = f32(cam)
        self.ctx.check(self.lib.rzk_lightslice_initialize(self.h, ptr(cam)), "rzk_lightslice_initialize")

    def gather(self, to_host=True):
        out = np.zeros((self.hgt, self.w, 3), np.float32) if to_host else None
        self.ctx.check(self.lib.rzk_lightslice_gather(self.h, ptr(out)), "rzk_lightslice_gather")
        return out

    def info(self) -> LightSliceInfo:
        inf = LightSliceInfo()
        self.lib.rzk_lightslice_get_info(self.h, C.byref(inf))
        return inf

    def render_guided(self, cam, spp=16, max_bounce=5, direct_lighting=0, guide_mode=1, seed=2013, tmin=1e-3,
                      window=None, to_host=True, count_rays=True):
        p = PtParams(spp, max_bounce, direct_lighting, guide_mode, seed, tmin, 0, 0, 0, 0)
        if window is not None:
            p.x0, p.y0, p.w, p.h = window
        cam = f32(cam)
        out = np.zeros((self.hgt, self.w, 3), np.float32) if to_host else None
        rays = np.zeros(2, np.int64) if count_rays else None
        self.ctx.check(self.lib.rzk_lightslice_render_guided(self.h, ptr(cam), C.byref(p), ptr(out), ptr(rays)),
                       "rzk_lightslice_render_guided")
        if rays is None:
            return out, None
        return out, {"closest_rays": int(rays[0]), "shadow_rays": int(rays[1])}

    def guide(self, cam):
        """borrow this LightSlice's guiding data (valid until it is initialised again)"""
        return Guide.from_lightslice(self, cam)

    def clusters(self):
        """the kept clusters as a resident set (rzk_lightslice_clusters)"""
        return Clusters(self)

    def count_work(self):
        out = np.zeros(8, np.int64)
        self.ctx.check(self.lib.rzk_lightslice_count_work(self.h, ptr(out)), "rzk_lightslice_count_work")
        keys = ["primary_rays", "reduced_pairs", "reduced_nodes", "reduced_tris", "gather_pairs", "gather_nodes",
                "gather_tris", "clusters"]
        return dict(zip(keys, (int(x) for x in out)))

    def get(self, name, dtype, count):
        out = np.zeros(count, dtype)
        self.ctx.check(self.lib.rzk_lightslice_get(self.h, name.encode(), ptr(out), out.nbytes),
                       f"rzk_lightslice_get({name})")
        return out

    def close(self):
        if getattr(self, "h", None):
            self.lib.rzk_lightslice_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Guide:
    """rzk_guide: the per-slice guiding data of one camera"""

    def __init__(self, ctx, h, width, height, keep=()):
        self.ctx, self.lib, self.h = ctx, ctx.lib, h
        self.width, self.height = width, height
        self._keep = keep

    @classmethod
    def from_lightslice(cls, ls, cam):
        cam = f32(cam)
        h = C.c_void_p()
        ls.ctx.check(ls.lib.rzk_lightslice_guide(ls.h, ptr(cam), C.byref(h)), "rzk_lightslice_guide")
        return cls(ls.ctx, h, int(cam[17]), int(cam[18]), keep=(ls,))

    @classmethod
    def create(cls, ctx, scene, clusters, cam, params=None):
        """a guide for `cam` from resident clusters (rzk_guide_create)"""
        cam = f32(cam)
        h = C.c_void_p()
        gp = GuideParams(**params) if params else None
        ctx.check(ctx.lib.rzk_guide_create(ctx.h, scene.h, clusters.h, ptr(cam), C.byref(gp) if gp else None,
                                           C.byref(h)), "rzk_guide_create")
        return cls(ctx, h, int(cam[17]), int(cam[18]), keep=(scene, clusters))

    def get(self, name, count):
        out = np.zeros(count, np.float32)
        self.ctx.check(self.lib.rzk_guide_get(self.h, name.encode(), ptr(out), out.nbytes), f"rzk_guide_get({name})")
        return out

    def render(self, spp=16, max_bounce=5, direct_lighting=0, guide_mode=1, seed=2013, tmin=1e-3, window=None,
               count_rays=False):
        """the plain per-request drop-in (rzk_render_guided): full [H][W][3] image"""
        p = PtParams(spp, max_bounce, direct_lighting, guide_mode, seed, tmin, 0, 0, 0, 0)
        if window is not None:
            p.x0, p.y0, p.w, p.h = window
        out = np.zeros((self.height, self.width, 3), np.float32)
        rays = np.zeros(2, np.int64) if count_rays else None
        self.ctx.check(self.lib.rzk_render_guided(self.ctx.h, self.h, C.byref(p), ptr(out), ptr(rays)),
                       "rzk_render_guided")
        return out, ({"closest_rays": int(rays[0]), "shadow_rays": int(rays[1])} if count_rays else None)

    def close(self):
        if getattr(self, "h", None):
            self.lib.rzk_guide_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_jobs(specs):
    """specs: dicts with guide, window (x0, y0, w, h), spp, max_bounce, direct_lighting, guide_mode, seed, tmin"""
    arr = (RenderJob * len(specs))()
    for i, s in enumerate(specs):
        x0, y0, w, h = s.get("window") or (0, 0, s["guide"].width, s["guide"].height)
        arr[i] = RenderJob(s["guide"].h, x0, y0, w, h, s.get("spp", 16), s.get("max_bounce", 5),
                           s.get("direct_lighting", 0), s.get("guide_mode", 1), s.get("seed", 2013),
                           s.get("tmin", 1e-3))
    return arr


class Batch:
    """rzk_batch: many guided path tracing jobs per persistent, graph-launched kernel"""

    def __init__(self, ctx, scene, config=None):
        self.ctx, self.lib = ctx, ctx.lib
        h = C.c_void_p()
        cfg = BatchConfig(**dict(BatchConfig.DEFAULTS, **config)) if config else None
        ctx.check(self.lib.rzk_batch_create(ctx.h, scene.h, C.byref(cfg) if cfg else None, C.byref(h)),
                  "rzk_batch_create")
        self.h = h
        self.scene = scene

    @property
    def config(self):
        c = BatchConfig()
        self.ctx.check(self.lib.rzk_batch_get_config(self.h, C.byref(c)), "rzk_batch_get_config")
        return c.as_dict()

    @config.setter
    def config(self, d):
        c = BatchConfig(**dict(BatchConfig.DEFAULTS, **d))
        self.ctx.check(self.lib.rzk_batch_set_config(self.h, C.byref(c)), "rzk_batch_set_config")

    def render(self, specs, to_host=True, count_rays=False):
        """returns (list of [h][w][3] images or None, rays dict or None, device ms)"""
        jobs = make_jobs(specs)
        npx = sum(j.w * j.h for j in jobs)
        out = np.zeros(npx * 3, np.float32) if to_host else None
        rays = np.zeros(2, np.int64) if count_rays else None
        ms = C.c_float()
        self.ctx.check(self.lib.rzk_batch_render(self.h, jobs, len(jobs), ptr(out), ptr(rays), C.byref(ms)),
                       "rzk_batch_render")
        imgs = None
        if to_host:
            imgs, o = [], 0
            for j in jobs:
                n = j.w * j.h * 3
                imgs.append(out[o:o + n].reshape(j.h, j.w, 3))
                o += n
        rd = {"closest_rays": int(rays[0]), "shadow_rays": int(rays[1])} if count_rays else None
        return imgs, rd, ms.value

    def autotune(self, specs, max_tried=64):
        jobs = make_jobs(specs)
        best = BatchConfig()
        tried = (BatchTrial * max_tried)()
        n = C.c_int32()
        self.ctx.check(self.lib.rzk_batch_autotune(self.h, jobs, len(jobs), C.byref(best), tried, max_tried,
                                                   C.byref(n)), "rzk_batch_autotune")
        return best.as_dict(), [(tried[i].cfg.as_dict(), tried[i].ms) for i in range(min(n.value, max_tried))]

    def close(self):
        if getattr(self, "h", None):
            self.lib.rzk_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Clusters:
    """rzk_clusters: a LightSlice's kept clusters, resident on the device"""

    def __init__(self, ls):
        self.ctx, self.lib = ls.ctx, ls.lib
        h = C.c_void_p()
        self.ctx.check(self.lib.rzk_lightslice_clusters(ls.h, C.byref(h)), "rzk_lightslice_clusters")
        self.h = h
        self.scene = ls.scene              # referred to by the native object

    def close(self):
        if getattr(self, "h", None):
            self.lib.rzk_clusters_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """rzk_server: continuous-batching request loop over a resident scene and clusters.
    launch: the engine and its launch shape; by default the persistent path kernel
    (engine 0, 128-lane blocks, 16 per compute unit, one 64-pixel item per wave
    dequeue); engine=1 selects the wavefront engine and its path slots."""

    LAUNCH = dict(BatchConfig.DEFAULTS, engine=0, block=128, blocks_per_cu=16, items_per_dequeue=1)

    def __init__(self, device, scene, clusters, memory_budget=8 << 30, step_iterations=8, launch=None, guide=None):
        self.lib = load()
        self.scene, self.clusters = scene, clusters
        p = ServerParams()
        p.memory_budget = int(memory_budget)
        p.step_iterations = int(step_iterations)
        p.launch = BatchConfig(**dict(self.LAUNCH, **(launch or {})))
        p.use_cluster_params = 0 if guide else 1
        if guide:
            p.guide = GuideParams(**guide)
        h = C.c_void_p()
        rc = self.lib.rzk_server_create(device, scene.h, clusters.h, C.byref(p), C.byref(h))
        if rc != RZK_OK:
            raise RzkError(f"rzk_server_create failed ({rc})")
        self.h = h
        self.sizes = {}

    def _check(self, rc, what):
        if rc != RZK_OK:
            msg = self.lib.rzk_server_last_error(self.h)
            raise RzkError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def submit(self, cam, spp=4, max_bounce=5, direct_lighting=0, guide_mode=1, seed=2013, tmin=1e-3):
        r = Request()
        cam = f32(cam)
        for i in range(19):
            r.camera[i] = float(cam[i])
        r.spp, r.max_bounce, r.direct_lighting, r.guide_mode, r.seed, r.tmin = (spp, max_bounce, direct_lighting,
                                                                               guide_mode, seed, tmin)
        t = C.c_int64()
        self._check(self.lib.rzk_server_submit(self.h, C.byref(r), C.byref(t)), "rzk_server_submit")
        self.sizes[t.value] = (int(cam[18]), int(cam[17]))
        return t.value

    def wait(self, ticket, timeout=-1.0, to_host=True):
        H, W = self.sizes.pop(ticket)
        out = np.zeros((H, W, 3), np.float32) if to_host else None
        st = RequestStats()
        self._check(self.lib.rzk_server_wait(self.h, ticket, float(timeout), ptr(out), C.byref(st)),
                    "rzk_server_wait")
        return out, {"submitted": st.submitted, "admitted": st.admitted, "finished": st.finished,
                     "rounds": st.rounds, "device_bytes": st.device_bytes}

    def stats(self):
        st = ServerStats()
        self._check(self.lib.rzk_server_get_stats(self.h, C.byref(st)), "rzk_server_get_stats")
        return {k: getattr(st, k) for k, _ in st._fields_}

    def close(self):
        if getattr(self, "h", None):
            self.lib.rzk_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
