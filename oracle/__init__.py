"""ctypes binding of liboracle.so — the CPU restatement of Renzoku's hot path.

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it as the checker / CPU baseline.  The product
(renzoku_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_MT = os.path.join(HERE, "_ref", "libmtwist_ref.so")

RAY, RECV, VPL, MAT, CAM = 6, 21, 21, 12, 19


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def _f(a):
    return np.ascontiguousarray(a, np.float32)


def _i(a):
    return np.ascontiguousarray(a, np.int32)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P, I, F = C.c_void_p, C.c_int, C.c_float
        sig = {
            "orc_mt_seed32": ([C.c_uint32], None), "orc_mt_lrand": ([], C.c_uint32),
            "orc_mt_rand": ([], C.c_float), "orc_mt_fill": ([I, P], None),
            "orc_ray_make": ([P, P, P], None), "orc_box_hit": ([P, P, P, F, F], I),
            "orc_tri_hit": ([P, P, P, P, F, F, P, P, P], I), "orc_morton3d": ([F, F, F], C.c_uint32),
            "orc_find_split": ([P, I, I], I),
            "orc_scene_create": ([I, P, P, P, I, P, P, P, I, P], P), "orc_scene_destroy": ([P], None),
            "orc_bvh_num_nodes": ([P], I), "orc_bvh_depth": ([P], I), "orc_bvh_export": ([P, P, P, P, P, P], None),
            "orc_scene_box": ([P, P], None),
            "orc_trace_closest": ([P, I, P, F, F, P, P, P, P, P], None),
            "orc_trace_any": ([P, I, P, F, F, P, P], None),
            "orc_trace_segment": ([P, I, P, P, F, P, P], None),
            "orc_trace_closest_brute": ([P, I, P, F, F, P, P], None),
            "orc_fill_receivers": ([P, I, P, P, P, P, P, P, P, P], None),
            "orc_camera_rays": ([P, I, P, P, P], None),
            "orc_bsdf_eval": ([P, P, P, P, P], None), "orc_form_factor_abs": ([P, P, P, P], F),
            "orc_clamp_term": ([I, F, P, P, P, F, P], None),
            "orc_gather_point": ([P, P, I, P, I, I, F, F, P], I),
            "orc_incident_radiance": ([P, P, P, I, I, F, P], None),
            "orc_gather": ([P, I, P, P, I, P, P, I, F, F, F, P, I], None),
            "orc_gather_weighted": ([P, I, P, P, I, P, P, P, I, F, F, F, P, I], None),
            "orc_reduced_matrix": ([P, I, P, P, I, P, P, I, F, F, P, P, I], None),
            "orc_alphas": ([I, I, P, P], None), "orc_column_magnitudes": ([I, I, P, P], None),
            "orc_cluster_assign": ([I, I, P, I, P, P, P, I], None),
            "orc_pdf_build": ([I, P, P, P], I), "orc_pdf_sample": ([I, P, P, F, P, P], None),
            "orc_cluster_by_sampling": ([I, I, P, I, P, P, P, P, P, P, I], I),
            "orc_refine_slices": ([I, I, P, I, P, I, P, P, P, P, P, I, P, P, P, P, P, P], I),
            "orc_refine_slices_n": ([I, I, I, P, I, P, I, P, P, P, P, P, I, P, P, P, P, P, P], I),
            "orc_slice_cost": ([I, P, I, P, I, P], F),
            "orc_map_vec6": ([P, P, F, P], None),
            "orc_kd_build": ([P, I, I, C.c_uint32, P, P, P, I, P, P, P, I, P], I),
            "orc_kd_build_mt": ([P, I, I, P, P, P, I, P, P, P, I, P, P], I),
            "orc_kd_find": ([P, P, P, I, P, I, P], None),
            "orc_kd_nearest": ([P, P, P, I, P, I, P, I, I, P], None),
            "orc_direction_to_square": ([P, P], None), "orc_square_to_direction": ([P, P], None),
            "orc_slice_cluster_info": ([P, P, I, P, P, P, P, I, F, P, P], None),
            "orc_slice_cluster_pdf": ([I, P, I, P, I, P, P, P, P], None),
            "orc_build_map": ([I, I, P, P, P, P, P, F, I, P, P, P, P, P], I),
            "orc_scene_set_lights": ([P, I, P], None), "orc_philox": ([P, C.c_uint32, C.c_uint32], None),
            "orc_sincos_turn": ([C.c_float, P, P], None),
            "orc_generate_vpls": ([P, C.c_longlong, C.c_uint32, C.c_longlong, C.c_float, P, P,
                                   C.POINTER(C.c_longlong)], I),
            "orc_bsdf_pdf": ([P, P, P, P, P], None), "orc_bsdf_sample": ([P, P, P, P, P, P, P], None),
            "orc_guided_pt": ([P, P, I, I, I, I, C.c_uint32, F, I, I, P, P, P, P, I, I, I, I, P, I], None),
            "orc_guided_pt_rays": ([P, I, I, I, I, I, C.c_uint32, F, I, I, P, P, P, P, I, P, P, P, P, I], None),
            "orc_camera_sample_rays": ([P, I, C.c_uint32, I, P, P, P, P], None),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


class OrcScene:
    def __init__(self, sc):
        self.L = lib()
        self._keep = [_f(sc.positions), _f(sc.tangents), _i(sc.tri_surface), _i(sc.surf_material),
                      _i(sc.surf_light), _i(sc.surf_flags), _f(sc.materials)]
        pos, tan, ts, sm, sl, sf, mat = self._keep
        self.h = self.L.orc_scene_create(len(ts), _p(pos), _p(tan), _p(ts), len(sm), _p(sm), _p(sl), _p(sf),
                                         mat.shape[0], _p(mat))
        self.ntri = len(ts)
        self.materials = mat
        rad = _f(sc.light_radiance).reshape(-1, 3)
        self._keep.append(rad)
        if len(rad) > 0:
            self.L.orc_scene_set_lights(self.h, len(rad), _p(rad))

    def __del__(self):
        try:
            self.L.orc_scene_destroy(self.h)
        except Exception:
            pass

    def generate_vpls(self, num_vpls, seed=2013, max_samples=1 << 26, tmin=1e-3):
        """orc_generate_vpls: (vpl (n, 21) f32, vmat (n,) i32, light samples used)"""
        vpl = np.zeros((int(num_vpls), 21), np.float32)
        vmat = np.zeros(int(num_vpls), np.int32)
        used = C.c_longlong()
        rc = self.L.orc_generate_vpls(self.h, int(num_vpls), int(seed) & 0xffffffff, int(max_samples), float(tmin),
                                      _p(vpl), _p(vmat), C.byref(used))
        if rc != 0:
            raise RuntimeError(f"only part of {num_vpls} VPLs after {used.value} light samples")
        return vpl, vmat, int(used.value)

    def export_bvh(self):
        n = self.ntri
        nn = self.L.orc_bvh_num_nodes(self.h)
        st, sm = np.zeros(n, np.int32), np.zeros(n, np.uint32)
        ch, rg, bx = np.zeros(2 * nn, np.int32), np.zeros(2 * nn, np.int32), np.zeros(6 * nn, np.float32)
        self.L.orc_bvh_export(self.h, _p(st), _p(sm), _p(ch), _p(rg), _p(bx))
        return st, sm, ch.reshape(-1, 2), rg.reshape(-1, 2), bx.reshape(-1, 6)

    def depth(self):
        return self.L.orc_bvh_depth(self.h)

    def trace_closest(self, rays, tmin=1e-3, tmax=1e6, stats=False):
        rays = _f(rays).reshape(-1, 6)
        n = len(rays)
        t, b, g = np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32)
        tri = np.zeros(n, np.int32)
        st = np.zeros((n, 2), np.int32) if stats else None
        self.L.orc_trace_closest(self.h, n, _p(rays), tmin, tmax, _p(t), _p(b), _p(g), _p(tri), _p(st))
        return (t, b, g, tri, st) if stats else (t, b, g, tri)

    def trace_any(self, rays, tmin=1e-3, tmax=1e6, stats=False):
        rays = _f(rays).reshape(-1, 6)
        out = np.zeros(len(rays), np.uint8)
        st = np.zeros((len(rays), 2), np.int32) if stats else None
        self.L.orc_trace_any(self.h, len(rays), _p(rays), tmin, tmax, _p(out), _p(st))
        return (out, st) if stats else out

    def trace_segments(self, p, q, tmin=1e-3, stats=False):
        p, q = _f(p).reshape(-1, 3), _f(q).reshape(-1, 3)
        out = np.zeros(len(p), np.uint8)
        st = np.zeros((len(p), 2), np.int32) if stats else None
        self.L.orc_trace_segment(self.h, len(p), _p(p), _p(q), tmin, _p(out), _p(st))
        return (out, st) if stats else out

    def trace_brute(self, rays, tmin=1e-3, tmax=1e6):
        rays = _f(rays).reshape(-1, 6)
        t, tri = np.zeros(len(rays), np.float32), np.zeros(len(rays), np.int32)
        self.L.orc_trace_closest_brute(self.h, len(rays), _p(rays), tmin, tmax, _p(t), _p(tri))
        return t, tri

    def fill_receivers(self, rays, t, b, g, tri):
        rays = _f(rays).reshape(-1, 6)
        n = len(rays)
        recv = np.zeros((n, RECV), np.float32)
        m, l = np.zeros(n, np.int32), np.zeros(n, np.int32)
        self.L.orc_fill_receivers(self.h, n, _p(rays), _p(_f(t)), _p(_f(b)), _p(_f(g)), _p(_i(tri)), _p(recv), _p(m),
                                  _p(l))
        return recv, m, l

    def gather(self, recv, rmat, vpl, vmat, clamp_mode, clamp, tmin, n_emitted, nthreads=1):
        recv, rmat, vpl, vmat = _f(recv), _i(rmat), _f(vpl), _i(vmat)
        out = np.zeros((len(rmat), 3), np.float32)
        self.L.orc_gather(self.h, len(rmat), _p(recv), _p(rmat), len(vmat), _p(vpl), _p(vmat), clamp_mode, clamp,
                          tmin, n_emitted, _p(out), nthreads)
        return out

    def gather_clusters(self, recv, rmat, vpl, vmat, weights, clamp_mode, clamp, tmin, n_emitted, nthreads=1):
        """Σ_c w_c gather_point(rep_c) / n_emitted per receiver (lightslice.cc:1325-1339): vpl/vmat are
        the representatives' records, the weight multiplies after the clamp"""
        recv, rmat, vpl, vmat, w = _f(recv), _i(rmat), _f(vpl), _i(vmat), _f(weights)
        out = np.zeros((len(rmat), 3), np.float32)
        self.L.orc_gather_weighted(self.h, len(rmat), _p(recv), _p(rmat), len(vmat), _p(vpl), _p(vmat), _p(w),
                                   clamp_mode, clamp, tmin, n_emitted, _p(out), nthreads)
        return out

    def gather_point(self, recv, rmat, vpl, vmat, clamp_mode, clamp, tmin):
        out = np.zeros(3, np.float32)
        vis = self.L.orc_gather_point(self.h, _p(_f(recv)), int(rmat), _p(_f(vpl)), int(vmat), clamp_mode, clamp,
                                      tmin, _p(out))
        return out, vis

    def incident(self, recv, vpl, vmat, clamp_mode, clamp):
        out = np.zeros(3, np.float32)
        self.L.orc_incident_radiance(self.h, _p(_f(recv)), _p(_f(vpl)), int(vmat), clamp_mode, clamp, _p(out))
        return out

    def reduced_matrix(self, recv, rmat, vpl, vmat, clamp_mode, clamp, tmin, nthreads=1):
        recv, rmat, vpl, vmat = _f(recv), _i(rmat), _f(vpl), _i(vmat)
        rows, V = len(rmat), len(vmat)
        R = np.zeros((V, rows, 3), np.float32)
        Vm = np.zeros((V, rows), np.float32)
        self.L.orc_reduced_matrix(self.h, rows, _p(recv), _p(rmat), V, _p(vpl), _p(vmat), clamp_mode, clamp, tmin,
                                  _p(R), _p(Vm), nthreads)
        return R, Vm

    def slice_cluster_info(self, recv_row, reps, vpl, vmat, Vrow, clamp_mode, clamp):
        reps = _i(reps)
        wi = np.zeros((len(reps), 3), np.float32)
        up = np.zeros(len(reps), np.int32)
        self.L.orc_slice_cluster_info(self.h, _p(_f(recv_row)), len(reps), _p(reps), _p(_f(vpl)), _p(_i(vmat)),
                                      _p(_f(Vrow)), clamp_mode, clamp, _p(wi), _p(up))
        return wi, up


def guided_pt(osc, cam, spp, max_bounce, direct, guide_mode, seed, tmin, tile, res=0, maps=None, row_cdf=None,
              col_cdf=None, totals=None, window=None, nthreads=8):
    """guided path tracing of a pixel window; returns the full [H][W][3] image
    (pixels outside the window stay 0)"""
    cam = _f(cam)
    W, H = int(cam[17]), int(cam[18])
    x0, y0, w, h = window if window is not None else (0, 0, W, H)
    out = np.zeros((H, W, 3), np.float32)
    arrs = [None if a is None else _f(a) for a in (maps, row_cdf, col_cdf, totals)]
    lib().orc_guided_pt(osc.h, _p(cam), spp, max_bounce, direct, guide_mode, seed, tmin, tile, res,
                        *[_p(a) for a in arrs], x0, y0, w, h, _p(out), nthreads)
    return out


def guided_pt_rays(osc, width, height, max_bounce, direct, guide_mode, seed, tmin, tile, rays, pix, smp, res=0,
                   maps=None, row_cdf=None, col_cdf=None, totals=None, nthreads=8):
    """the guided path sampler from explicit primary rays (ray records): ray i is sample smp[i] of
    pixel pix[i] = (x, y); returns each ray's radiance (n, 3)"""
    rays, pix, smp = _f(rays).reshape(-1, 6), _i(pix).reshape(-1, 2), _i(smp)
    out = np.zeros((len(rays), 3), np.float32)
    arrs = [None if a is None else _f(a) for a in (maps, row_cdf, col_cdf, totals)]
    lib().orc_guided_pt_rays(osc.h, width, height, max_bounce, direct, guide_mode, seed, tmin, tile, res,
                             *[_p(a) for a in arrs], len(rays), _p(rays), _p(pix), _p(smp), _p(out), nthreads)
    return out


def camera_sample_rays(cam, spp, seed, px, py, smp):
    """the primary ray records guided_pt draws for sample smp of pixel (px, py)"""
    px, py, smp = _i(px), _i(py), _i(smp)
    out = np.zeros((len(px), 6), np.float32)
    lib().orc_camera_sample_rays(_p(_f(cam)), spp, seed, len(px), _p(px), _p(py), _p(smp), _p(out))
    return out


def philox(counter, k0, k1):
    c = np.ascontiguousarray(counter, np.uint32).copy()
    lib().orc_philox(_p(c), k0, k1)
    return c


def bsdf_pdf(mat, uvn, wo, wi):
    out = np.zeros(1, np.float32)
    lib().orc_bsdf_pdf(_p(_f(mat)), _p(_f(uvn)), _p(_f(wo)), _p(_f(wi)), _p(out))
    return float(out[0])


def bsdf_sample(mat, uvn, wo, u):
    wi, pdf, f = np.zeros(3, np.float32), np.zeros(1, np.float32), np.zeros(3, np.float32)
    lib().orc_bsdf_sample(_p(_f(mat)), _p(_f(uvn)), _p(_f(wo)), _p(np.ascontiguousarray(u, np.uint32)), _p(wi),
                          _p(pdf), _p(f))
    return wi, float(pdf[0]), f


def camera_rays(cam, px, py):
    px, py = _i(px), _i(py)
    rays = np.zeros((len(px), 6), np.float32)
    lib().orc_camera_rays(_p(_f(cam)), len(px), _p(px), _p(py), _p(rays))
    return rays


def ray_make(o, d):
    out = np.zeros(6, np.float32)
    lib().orc_ray_make(_p(_f(o)), _p(_f(d)), _p(out))
    return out


def box_hit(bmin, bmax, ray, tmin, tmax):
    return lib().orc_box_hit(_p(_f(bmin)), _p(_f(bmax)), _p(_f(ray)), tmin, tmax)


def tri_hit(p0, p1, p2, ray, tmin, tmax):
    t, b, g = (np.zeros(1, np.float32) for _ in range(3))
    h = lib().orc_tri_hit(_p(_f(p0)), _p(_f(p1)), _p(_f(p2)), _p(_f(ray)), tmin, tmax, _p(t), _p(b), _p(g))
    return h, float(t[0]), float(b[0]), float(g[0])


def morton3d(x, y, z):
    return lib().orc_morton3d(x, y, z)


def find_split(codes, first, last):
    codes = np.ascontiguousarray(codes, np.uint32)
    return lib().orc_find_split(_p(codes), first, last)


def mt_stream(seed, n):
    L = lib()
    L.orc_mt_seed32(seed)
    out = np.zeros(n, np.uint32)
    L.orc_mt_fill(n, _p(out))
    return out


def ref_mt_stream(seed, n):
    """the reference's own mtwist.c (built from its source into oracle/_ref)"""
    if not os.path.exists(REF_MT):
        return None
    R = C.CDLL(REF_MT)
    R.mt_seed32.argtypes = [C.c_uint32]
    R.mt_lrand.restype = C.c_uint32
    R.mt_seed32(seed)
    return np.array([R.mt_lrand() for _ in range(n)], np.uint32)


def bsdf_eval(mat, uvn, wo, wi):
    out = np.zeros(3, np.float32)
    lib().orc_bsdf_eval(_p(_f(mat)), _p(_f(uvn)), _p(_f(wo)), _p(_f(wi)), _p(out))
    return out


def form_factor_abs(p, pn, q, qn):
    return lib().orc_form_factor_abs(_p(_f(p)), _p(_f(pn)), _p(_f(q)), _p(_f(qn)))


def clamp_term(mode, thr, power, fs, fl, G):
    out = np.zeros(3, np.float32)
    lib().orc_clamp_term(mode, thr, _p(_f(power)), _p(_f(fs)), _p(_f(fl)), G, _p(out))
    return out


def alphas(R):
    R = _f(R)
    cols, rows = R.shape[0], R.shape[1]
    out = np.zeros(cols, np.float32)
    lib().orc_alphas(rows, cols, _p(R), _p(out))
    return out


def column_magnitudes(R):
    R = _f(R)
    cols, rows = R.shape[0], R.shape[1]
    out = np.zeros(cols, np.float32)
    lib().orc_column_magnitudes(rows, cols, _p(R), _p(out))
    return out


def cluster_assign(R, centers, weights, nthreads=1):
    R = _f(R)
    cols, rows = R.shape[0], R.shape[1]
    out = np.zeros(cols, np.int32)
    lib().orc_cluster_assign(rows, cols, _p(R), len(centers), _p(_i(centers)), _p(_f(weights)), _p(out), nthreads)
    return out


def pdf_build(arr):
    arr = _f(arr)
    pdf, cdf = np.zeros(len(arr), np.float32), np.zeros(len(arr), np.float32)
    ok = lib().orc_pdf_build(len(arr), _p(arr), _p(pdf), _p(cdf))
    return ok, pdf, cdf


def pdf_sample(pdf, cdf, u):
    k, pk = np.zeros(1, np.int32), np.zeros(1, np.float32)
    lib().orc_pdf_sample(len(pdf), _p(_f(pdf)), _p(_f(cdf)), u, _p(k), _p(pk))
    return int(k[0]), float(pk[0])


def cluster_by_sampling(R, num_clusters, seed, nthreads=1, skip=0):
    """ClusterBuilder::cluster_by_sampling on a freshly seeded MT stream
    (after `skip` draws)."""
    R = _f(R)
    cols, rows = R.shape[0], R.shape[1]
    L = lib()
    L.orc_mt_seed32(seed)
    for _ in range(int(skip)):
        L.orc_mt_rand()
    centers, cw = np.zeros(num_clusters, np.int32), np.zeros(num_clusters, np.float32)
    assign = np.zeros(cols, np.int32)
    reps, rw, kept = (np.zeros(num_clusters, np.int32), np.zeros(num_clusters, np.float32),
                      np.zeros(num_clusters, np.int32))
    code = L.orc_cluster_by_sampling(rows, cols, _p(R), num_clusters, _p(centers), _p(cw), _p(assign), _p(reps),
                                     _p(rw), _p(kept), nthreads)
    ncl = code // 65536
    return dict(centers=centers[:ncl], center_w=cw[:ncl], assign=assign, reps=reps[:ncl][kept[:ncl] > 0],
                rep_w=rw[:ncl][kept[:ncl] > 0], ncl=ncl)


def kept_cluster_members(assign, ncl, kept):
    """CSR (offsets, members in increasing VPL order) of the kept clusters,
    renumbered 0.. in cluster order (cluster_builder.cc:223, :264-272)"""
    assign = np.asarray(assign)
    off, mem = [0], []
    for c in range(ncl):
        if not kept[c]:
            continue
        m = np.nonzero(assign == c)[0]
        mem.extend(m.tolist())
        off.append(len(mem))
    return np.array(off, np.int32), np.array(mem, np.int32)


def slice_cost(R, nb_rows, idx):
    """compute_distance (lightslice.cc:336-357) of columns idx over rows nb_rows"""
    R, nb_rows, idx = _f(R), _i(nb_rows), _i(idx)
    return float(lib().orc_slice_cost(R.shape[1], _p(R), len(nb_rows), _p(nb_rows), len(idx), _p(idx)))


def cluster_and_refine(R, num_clusters, max_slice_clusters, nb, seed, nthreads=1, with_perm=False, slices=None,
                       skip=0):
    """cluster_by_sampling then the per-slice refinement on one MT stream seeded
    with `seed` (lightslice.cc:757-902), after `skip` draws.  R: (V, S, 3);
    nb: (S, nnb) rows; slices: refine only the first `slices` slices (default: all)."""
    R = _f(R)
    V, S = R.shape[0], R.shape[1]
    nb = _i(nb)
    L = lib()
    L.orc_mt_seed32(seed)
    for _ in range(int(skip)):
        L.orc_mt_rand()
    centers, cw = np.zeros(num_clusters, np.int32), np.zeros(num_clusters, np.float32)
    assign = np.zeros(V, np.int32)
    reps, rw, kept = (np.zeros(num_clusters, np.int32), np.zeros(num_clusters, np.float32),
                      np.zeros(num_clusters, np.int32))
    code = L.orc_cluster_by_sampling(S, V, _p(R), num_clusters, _p(centers), _p(cw), _p(assign), _p(reps), _p(rw),
                                     _p(kept), nthreads)
    ncl = code // 65536
    off, mem = kept_cluster_members(assign, ncl, kept[:ncl])
    k = kept[:ncl] > 0
    creps, crw = _i(reps[:ncl][k]), _f(rw[:ncl][k])
    mags = column_magnitudes(R)
    M = max_slice_clusters
    count, rep, w = np.zeros(S, np.int32), np.zeros(S * M, np.int32), np.zeros(S * M, np.float32)
    size, draws = np.zeros(S * M, np.int32), np.zeros(S, np.int32)
    perm = np.zeros(S * V, np.int32) if with_perm else None
    ns = S if slices is None else int(slices)
    rc = L.orc_refine_slices_n(S, ns, V, _p(R), nb.shape[1], _p(nb), len(creps), _p(off), _p(mem), _p(creps),
                               _p(crw), _p(mags), M, _p(count), _p(rep), _p(w), _p(size), _p(perm), _p(draws))
    assert rc == 0, rc
    out = dict(centers=centers[:ncl], assign=assign, reps=creps, rep_w=crw, off=off, mem=mem, count=count,
               rep=rep.reshape(S, M), w=w.reshape(S, M), size=size.reshape(S, M), draws=draws)
    if with_perm:
        out["perm"] = perm.reshape(S, V)
    return out


def sincos_turn(u):
    """the engine's cos/sin of 2 pi u, restated: (sin, cos)"""
    a, b = C.c_float(), C.c_float()
    lib().orc_sincos_turn(float(u), C.byref(a), C.byref(b))
    return a.value, b.value


def direction_to_square(wi):
    out = np.zeros(2, np.float32)
    lib().orc_direction_to_square(_p(_f(wi)), _p(out))
    return out


def square_to_direction(s):
    out = np.zeros(3, np.float32)
    lib().orc_square_to_direction(_p(_f(s)), _p(out))
    return out


def slice_cluster_pdf(R, nb_rows, reps, rep_w, upper):
    R = _f(R)
    rows = R.shape[1]
    out = np.zeros(len(reps), np.float32)
    lib().orc_slice_cluster_pdf(rows, _p(R), len(nb_rows), _p(_i(nb_rows)), len(reps), _p(_i(reps)), _p(_f(rep_w)),
                                _p(_i(upper)), _p(out))
    return out


def build_map(res, slice_wi, upper, rep_w, R_rep_rgb, uvn, eps, valid):
    n = len(upper)
    m = np.zeros(res * res, np.float32)
    rc, rp = np.zeros(res, np.float32), np.zeros(res, np.float32)
    cc = np.zeros(res * res, np.float32)
    tot = np.zeros(1, np.float32)
    ok = lib().orc_build_map(res, n, _p(_f(slice_wi)), _p(_i(upper)), _p(_f(rep_w)), _p(_f(R_rep_rgb)), _p(_f(uvn)),
                             eps, int(valid), _p(m), _p(rc), _p(rp), _p(cc), _p(tot))
    return dict(ok=ok, map=m.reshape(res, res), row_cdf=rc, row_pdf=rp, col_cdf=cc.reshape(res, res),
                total=float(tot[0]))


# ---- 6-D kd-tree slicing (kdslice.cpp) ----------------------------------------

def map_vec6(recv, mats, materials, diagonal):
    """LightSlice map_vec6 of receiver records (n x RECV) with material ids
    (-1: none) into n x 6 points"""
    L = lib()
    recv = _f(recv).reshape(-1, 21)
    mats = _i(mats)
    materials = _f(materials).reshape(-1, 12)
    out = np.zeros((len(recv), 6), np.float32)
    for i in range(len(recv)):
        m = int(mats[i])
        mp = _p(materials[m]) if m >= 0 else None
        L.orc_map_vec6(_p(recv[i]), mp, float(diagonal), _p(out[i]))
    return out


class KdSlices:
    """cluster_kdtree over n 6-D points: order, slices (S x 2), boxes (S x 12),
    heap arrays axis / plane / cl of length heap_used.  pivots="philox": round t
    of node [s, e] keyed by (seed, s, e, t); pivots="mt": the global MT stream
    seeded with seed, drawn in recursion order as the reference does (draws =
    numbers taken; the stream is left where the build stopped)."""

    def __init__(self, pts, max_slice_size=256, seed=2013, pivots="philox"):
        L = lib()
        pts = _f(pts).reshape(-1, 6)
        n = len(pts)
        cap_heap = 4 * max(n, 1) + 4
        self.order = np.zeros(n, np.int32)
        self.slices_buf = np.zeros((max(n, 1), 2), np.int32)
        self.boxes_buf = np.zeros((max(n, 1), 12), np.float32)
        axis = np.zeros(cap_heap, np.int32)
        plane = np.zeros(cap_heap, np.float32)
        cl = np.zeros(cap_heap, np.int32)
        used = C.c_int(0)
        self.draws = 0
        if pivots == "mt":
            draws = C.c_int64(0)
            L.orc_mt_seed32(int(seed) & 0xffffffff)
            S = L.orc_kd_build_mt(_p(pts), n, int(max_slice_size), _p(self.order), _p(self.slices_buf),
                                  _p(self.boxes_buf), max(n, 1), _p(axis), _p(plane), _p(cl), cap_heap,
                                  C.byref(used), C.byref(draws))
            self.draws = draws.value
        elif pivots == "philox":
            S = L.orc_kd_build(_p(pts), n, int(max_slice_size), int(seed) & 0xffffffff, _p(self.order),
                               _p(self.slices_buf), _p(self.boxes_buf), max(n, 1), _p(axis), _p(plane), _p(cl),
                               cap_heap, C.byref(used))
        else:
            raise ValueError(f"unknown pivot stream {pivots!r}")
        if S < 0:
            raise ValueError("orc_kd_build: bad arguments")
        self.n, self.S, self.heap_used = n, S, used.value
        self.slices = self.slices_buf[:S].copy()
        self.boxes = self.boxes_buf[:S].copy()
        self.axis = axis[:used.value].copy()
        self.plane = plane[:used.value].copy()
        self.cl = cl[:used.value].copy()
        self.axis_all, self.plane_all = axis, plane

    def find(self, q):
        q = _f(q).reshape(-1, 6)
        out = np.zeros(len(q), np.int32)
        lib().orc_kd_find(_p(self.axis), _p(self.plane), _p(self.cl), self.heap_used, _p(q), len(q), _p(out))
        return out

    def nearest(self, q, k):
        q = _f(q).reshape(-1, 6)
        out = np.zeros((len(q), k), np.int32)
        lib().orc_kd_nearest(_p(self.axis), _p(self.plane), _p(self.cl), self.heap_used, _p(self.boxes), self.S,
                             _p(q), len(q), int(k), _p(out))
        return out


def kd_representatives(slices, n_points, seed, skip=0):
    """LightSlice::generate_slices' representatives (lightslice.cc:293-301) on
    the MT stream seeded with seed after its first skip numbers (the tree's
    pivots when they came from the same stream): per slice start + rd() * size
    in float, truncated (kept inside the point array), then the colour's draw.
    Returns the kd positions and the number of draws taken from the seed on
    (skip + 2 per slice)."""
    L = lib()
    L.orc_mt_seed32(seed)
    for _ in range(int(skip)):
        L.orc_mt_rand()
    reps = np.zeros(len(slices), np.int32)
    for s, (a, b) in enumerate(np.asarray(slices)):
        u = np.float32(L.orc_mt_rand())
        r = int(np.float32(np.float32(a) + np.float32(u * np.float32(b - a + 1))))
        reps[s] = min(r, n_points - 1)
        L.orc_mt_rand()
    return reps, int(skip) + 2 * len(slices)
