"""Benchmark: one LightSlice gather step on one MI355X.

Workload (the 200k-triangle LightSlice configuration): seeded room of ~200k
triangles (tessellated boxes and UV spheres, half Lambertian, half Phong),
512x512 pixels, 65,536 bounce-1 VPLs, 32x32-pixel slices (256 rows), 300
cluster centres.  A step is the whole hot path with the scene BVH and the VPL
set resident in HBM: primary closest hits for every pixel, the reduced matrix
(256 x 65,536 shadow segments + incident radiance), alphas, centre sampling,
assignment, representatives, per-slice cluster pdfs and luminance maps, and
the final per-pixel gather over the cluster representatives (one shadow
segment per (pixel, representative) pair).

value = shadow Mrays/s over the step (reduced-matrix + gather segments / step
time).  The dominant kernel (the fused gather) is also timed on its own with
HIP events on the context stream, and its roofline is reported against HBM
peak with the algorithmic bytes of the traversal byte model (36 B of segment
I/O, 128 B per visited 4-wide node, 48 B per tested triangle).  cpu_baseline is
the oracle (a C restatement of the reference's gather) on a bounded sample of
the same pairs, on the host's cores.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per gather launch (k_slice_visibility + k_slice_shade) measured by rocprofv3 --pmc FETCH_SIZE
# (x2, the 512 MiB copy calibration) and WRITE_SIZE (x1) on this workload; the passes and the calibration
# are recorded in profiles/slice_gather_pmc.md
GATHER_HBM_BYTES = (30397 + 49212) * 2 * 1024 + (19206 + 3072) * 1024
# MI355X vector issue: 256 CUs x 4 SIMD x 16 lanes x 2.4 GHz lane-operations per second
VALU_LANE_OPS_PEAK = 256 * 64 * 2.4e9
# algorithmic bytes per shadow segment: ray in, result out, every 4-wide node
# fetched (128 B: four child boxes + refs) and every triangle record tested
RAY_IN_BYTES, SHADOW_OUT_BYTES, NODE4_BYTES, TRI_BYTES = 32, 4, 128, 48


class HipEvents:
    """hipEvent timing on an arbitrary stream (the engine's context stream)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [C.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.hip.hipEventDestroy.argtypes = [C.c_void_p]

    def create(self):
        e = C.c_void_p()
        assert self.hip.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.hip.hipEventRecord(e, C.c_void_p(stream)) == 0

    def elapsed_ms(self, a, b):
        assert self.hip.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.hip.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return ms.value


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(sc, ls, vpl, vmat, reps, rep_w, params, target_s=8.0):
    """oracle gather_clusters on a seeded sample of pixels, all host cores (<= 16)"""
    import oracle
    W, H = params["width"], params["height"]
    recv = ls.get("recv", np.float32, W * H * 21).reshape(-1, 21)
    rmat = ls.get("recv_mat", np.int32, W * H)
    osc = oracle.OrcScene(sc)
    vsub = vpl[reps].copy()
    vsub[:, 18:21] = vpl[reps, 18:21] * rep_w[:, None]
    vm = vmat[reps].copy()
    cores = max(1, min(16, os.cpu_count() or 1))
    rng = np.random.default_rng(1)
    ok = np.nonzero(rmat >= 0)[0]
    n = 16 * cores
    while True:
        sel = np.sort(rng.choice(ok, min(n, len(ok)), replace=False))
        t0 = time.perf_counter()
        osc.gather(recv[sel], rmat[sel], vsub, vm, params["clamp_mode"], params["clamp"], params["tmin"],
                   params["n_emitted"], nthreads=cores)
        dt = time.perf_counter() - t0
        if dt >= target_s * 0.5 or n >= len(ok):
            break
        n = int(min(len(ok), n * max(2.0, target_s / max(dt, 1e-3))))
    pairs = len(sel) * len(reps)
    return {"value": pairs / dt / 1e6, "unit": "Mrays/s", "cores": cores, "kind": "port",
            "sample": f"oracle gather over {len(sel)} seeded pixels x {len(reps)} cluster representatives "
                      f"({pairs} shadow segments + shadings) on {cores} threads, {dt:.1f} s"}


def run_guided(args):
    """Guided path tracing of the 1M-triangle instanced scene (Ward BRDFs):
    1280x720, 16 spp, depth 5, 100k VPLs clustered into 1,000 clusters with
    16x16-pixel slices driving the first-bounce direction sampler.  A step is
    one full-frame render_guided with the guide resident; the guide build
    (LightSlice initialize) is timed once beside it."""
    import renzoku_amd as rz
    from renzoku_amd import scenes, vplgen
    W, H = 1280, 720
    t0 = time.perf_counter()
    sc = scenes.instanced_meshes(W, H, seed=2013, target_tris=args.tris or 1_000_000)
    ctx = rz.Ctx(0)
    scene = rz.Scene(ctx, sc)
    log(f"[bench] scene: {sc.ntri} triangles, BVH depth {scene.info().depth} ({time.perf_counter() - t0:.1f} s)")
    tracer, filler = vplgen.gpu_tracer(scene)
    vpl, vmat, n_emitted = vplgen.generate_vpls(sc, tracer, filler, args.vpls or 100_000, seed=2013, batch=65536)
    vpls = rz.Vpls(ctx, vpl, vmat)
    log(f"[bench] VPLs: {len(vmat)} from {n_emitted:.0f} light samples")
    ls = rz.LightSlice(ctx, scene, vpls, width=W, height=H, tile=16, num_clusters=1000, num_neighbor_slices=16,
                       map_resolution=32, n_emitted=n_emitted, seed=2013)
    t1 = time.perf_counter()
    ls.initialize(sc.camera)
    ctx.synchronize()
    init_ms = (time.perf_counter() - t1) * 1e3
    inf = ls.info()
    log(f"[bench] guide built: {inf.num_slices} slices, {inf.num_clusters} clusters, {init_ms:.0f} ms")
    kw = dict(spp=16, max_bounce=5, direct_lighting=0, guide_mode=1, seed=2013, tmin=1e-3, to_host=False)
    _, rays = ls.render_guided(sc.camera, count_rays=True, **kw)
    ev = HipEvents()
    e0, e1 = ev.create(), ev.create()
    for i in range(args.warmup):
        ls.render_guided(sc.camera, count_rays=False, **kw)
    ctx.synchronize()
    ms = []
    t_start = time.perf_counter()
    for i in range(args.steps):
        ev.record(e0, ctx.stream)
        ls.render_guided(sc.camera, count_rays=False, **kw)
        ev.record(e1, ctx.stream)
        ms.append(ev.elapsed_ms(e0, e1))
        log(f"[bench] step {i + 1}/{args.steps}: {ms[-1]:.2f} ms")
    ctx.synchronize()
    ms_per_step = (time.perf_counter() - t_start) * 1e3 / args.steps
    nrays = rays["closest_rays"] + rays["shadow_rays"]
    # the same frame as one job of the wavefront engine (bit-equal image)
    b = rz.Batch(ctx, scene, dict(engine=1))
    job = [dict(guide=ls.guide(sc.camera), window=(0, 0, W, H), **{k: v for k, v in kw.items() if k != "to_host"})]
    imgs, wrays, _ = b.render(job, count_rays=True)
    ref, _ = ls.render_guided(sc.camera, count_rays=False, **dict(kw, to_host=True))
    same = bool(np.array_equal(imgs[0].view(np.uint32), ref.view(np.uint32)))
    wf_ms = []
    for i in range(args.steps):
        _, _, dev_ms = b.render(job, to_host=False)
        wf_ms.append(dev_ms)
        log(f"[bench] wavefront step {i + 1}/{args.steps}: {dev_ms:.2f} ms")
    wf_ms = float(np.mean(wf_ms))
    out = {"metric": "guided path tracing Mrays/s (closest + shadow)", "value": round(nrays / (ms_per_step / 1e3) / 1e6, 3),
           "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(ms_per_step, 3), "kernel_ms": round(float(np.mean(ms)), 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded instanced-mesh scene, seeded light-traced VPLs)",
           "config": {"workload": "guided path tracing 1M tris 1280x720 16 spp depth 5 100k VPLs 16x16 slices "
                                  "1000 clusters", "triangles": int(sc.ntri), "width": W, "height": H,
                      "spp": 16, "max_bounce": 5, "vpls": int(len(vmat)), "slices": int(inf.num_slices),
                      "clusters": int(inf.num_clusters), "parallelism": "single GPU"},
           "rays": rays, "paths_per_s": round(W * H * 16 / (ms_per_step / 1e3), 1),
           "guide_build_ms": round(init_ms, 1),
           "wavefront": {"ms": round(wf_ms, 3), "mrays_per_s": round(nrays / wf_ms / 1e3, 1),
                         "same_rays": wrays == rays, "bit_equal_image": same, "config": b.config}}
    print(json.dumps(out), flush=True)


def run_refine(args):
    """LightSlice per-slice refinement on the 200k-triangle LightSlice workload
    (512x512, 65,536 VPLs, 32x32-pixel slices, 16 neighbour slices, 300 kept
    clusters refined to 600 per slice, the reference driver's 2x rule).  A step
    is one initialize with refinement; the refinement time is that minus an
    initialize without it (same LightSlice inputs).  The refinement reads the
    reference's single MT stream slice after slice, so it takes one pass per
    slice whose number count differs from 4 per split; the pass count is
    reported.  cpu_baseline: the oracle refining the first slices (the
    sequential computation's own prefix) on one thread."""
    import renzoku_amd as rz
    from renzoku_amd import scenes, vplgen
    import oracle
    W = H = 512
    maxc = 600
    sc = scenes.room(W, H, seed=2013, target_tris=args.tris or 200_000)
    ctx = rz.Ctx(0)
    scene = rz.Scene(ctx, sc)
    tracer, filler = vplgen.gpu_tracer(scene)
    vpl, vmat, n_emitted = vplgen.generate_vpls(sc, tracer, filler, args.vpls or 65_536, seed=2013, batch=65536)
    vpls = rz.Vpls(ctx, vpl, vmat)
    params = dict(width=W, height=H, tile=32, num_clusters=300, num_neighbor_slices=16, map_resolution=32,
                  n_emitted=n_emitted, seed=2013)
    plain = rz.LightSlice(ctx, scene, vpls, **params)
    refined = rz.LightSlice(ctx, scene, vpls, max_slice_clusters=maxc, **params)

    def timed(ls):
        t0 = time.perf_counter()
        ls.initialize(sc.camera)
        ctx.synchronize()
        return (time.perf_counter() - t0) * 1e3
    for i in range(args.warmup):
        timed(plain)
        timed(refined)
    plain_ms, refined_ms = [], []
    for i in range(args.steps):
        plain_ms.append(timed(plain))
        refined_ms.append(timed(refined))
        log(f"[bench] step {i + 1}/{args.steps}: initialize {plain_ms[-1]:.1f} ms, with refinement "
            f"{refined_ms[-1]:.1f} ms")
    inf = refined.info()
    S = inf.num_slices
    refine_ms = float(np.mean(refined_ms)) - float(np.mean(plain_ms))
    cnt = refined.get("slice_count", np.int32, S)
    out = {"metric": "LightSlice per-slice refinement, slices/s", "value": round(S / (refine_ms / 1e3), 3),
           "unit": "slices/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(float(np.mean(refined_ms)), 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded room scene, seeded light-traced VPLs)",
           "config": {"workload": "lightslice refinement room 200k tris 512x512 65536 VPLs 32x32 slices "
                                  "16 neighbours 300 -> 600 clusters", "triangles": int(sc.ntri), "vpls": int(len(vmat)),
                      "slices": S, "kept_clusters": int(inf.num_clusters), "max_slice_clusters": maxc,
                      "parallelism": "single GPU"},
           "refine_ms": round(refine_ms, 1), "initialize_ms": round(float(np.mean(plain_ms)), 2),
           "refine_passes": int(inf.refine_passes), "slice_clusters": [int(cnt.min()), int(cnt.max())]}
    if not args.no_cpu:
        R = refined.get("R", np.float32, S * len(vmat) * 3).reshape(len(vmat), S, 3)
        nb = refined.get("neighbors", np.int32, S * 16).reshape(S, 16)
        k = 24                                   # slices in the sample
        t0 = time.perf_counter()
        oracle.cluster_by_sampling(R, 300, 2013, nthreads=16)
        t1 = time.perf_counter()
        o = oracle.cluster_and_refine(R[:, :, :], 300, maxc, nb, 2013, nthreads=16, slices=k)
        dt = time.perf_counter() - t1 - (t1 - t0)      # the refinement part of the oracle's run
        out["cpu_baseline"] = {"value": round(k / dt, 3), "unit": "slices/s", "cores": 1, "kind": "port",
                               "sample": f"oracle refinement of the first {k} slices (the sequential run's prefix, "
                                         f"one thread; clustering on 16 threads subtracted), {dt:.1f} s"}
        same = all(np.array_equal(o["rep"][s, :o["count"][s]],
                                  refined.get("slice_reps", np.int32, S * maxc).reshape(S, maxc)[s, :o["count"][s]])
                   for s in range(k))
        out["cpu_baseline"]["same_representatives"] = bool(same)
    print(json.dumps(out), flush=True)


def run_vplgen(args):
    """Bounce-1 VPL generation on the device (rzk_generate_vpls) for the
    200k-triangle LightSlice scene: 65,536 VPLs, one closest ray per light
    sample on the reference tree.  A step is one whole generation (light
    sampling, tracing, ordered compaction, copy of the records to the host).
    Roofline: the closest-ray byte model (48 B of ray I/O, 64 B per visited
    binary node, 48 B per tested triangle) over the step time.  cpu_baseline:
    the oracle generating the same VPLs on one thread (the whole workload, well
    under a second of CPU work).  value counts the light samples used; the
    byte model counts every ray traced."""
    import renzoku_amd as rz
    from renzoku_amd import scenes
    import oracle
    W = H = 512
    nv = args.vpls or 65_536
    sc = scenes.room(W, H, seed=2013, target_tris=args.tris or 200_000)
    ctx = rz.Ctx(0)
    scene = rz.Scene(ctx, sc)
    _, _, used, work = rz.generate_vpls(ctx, scene, nv, seed=2013, count_work=True)
    for i in range(args.warmup):
        rz.generate_vpls(ctx, scene, nv, seed=2013)
    ms = []
    for i in range(args.steps):
        t0 = time.perf_counter()
        rz.generate_vpls(ctx, scene, nv, seed=2013)
        ms.append((time.perf_counter() - t0) * 1e3)
        log(f"[bench] step {i + 1}/{args.steps}: {ms[-1]:.2f} ms")
    ms_per_step = float(np.mean(ms))
    rays = work["closest_rays"]           # traced, with the last batch's samples past the last one used
    gbytes = rays * (RAY_IN_BYTES + 16) + work["nodes"] * 64 + work["tris"] * TRI_BYTES
    achieved = gbytes / (ms_per_step / 1e3) / 1e9
    out = {"metric": "bounce-1 VPL generation, light paths/s (closest rays)", "value": round(used / (ms_per_step / 1e3) / 1e6, 3),
           "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic (seeded room scene)",
           "config": {"workload": "vpl generation room 200k tris 65536 bounce-1 VPLs", "triangles": int(sc.ntri),
                      "vpls": nv, "light_samples": int(used), "parallelism": "single GPU"},
           "vpls_per_s": round(nv / (ms_per_step / 1e3), 1), "work": work,
           "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                        "kernel": "rzk_generate_vpls (k_light_vpls, k_vpl_order, k_vpl_scatter, host copy)",
                        "note": "step time includes the ordered compaction and the copy of the records to the host"}}
    if not args.no_cpu:
        k = nv
        osc = oracle.OrcScene(sc)
        t0 = time.perf_counter()
        ov, om, ou = osc.generate_vpls(k, seed=2013)
        dt = time.perf_counter() - t0
        gv, gm, gu = rz.generate_vpls(ctx, scene, k, seed=2013)
        out["cpu_baseline"] = {"value": round(ou / dt / 1e6, 4), "unit": "Mrays/s", "cores": 1, "kind": "port",
                               "sample": f"oracle generation of the first {k} VPLs ({ou} light samples) on one "
                                         f"thread, {dt:.2f} s",
                               "same_vpls": bool(gu == ou and np.array_equal(gv.view(np.uint32), ov.view(np.uint32)))}
    print(json.dumps(out), flush=True)


def scene_diagonal(sc):
    """BoundingBox::diagonal of the scene box (boundingbox.h:109-111), float32"""
    p = np.asarray(sc.positions, np.float32).reshape(-1, 3)
    d = p.max(0) - p.min(0)
    return float(np.sqrt(np.float32(np.float32(d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]), dtype=np.float32))


def run_kdslice(args):
    """LightSlice's slicing of the gather points by the 6-D kd-tree
    (generate_slices, lightslice.cc:252-303) on the 512x512 room: every pixel's
    primary hit with a material mapped by map_vec6, then the kd-tree build with
    slices of at most 256 points.  A step is one build from points resident in
    device memory (rzk_kdslices_build_dev).  Roofline: 36 B per point and depth
    (its 6 coordinates for the node box, its key, its index read and written)
    over the step time.  cpu_baseline: the oracle's build of the same points on
    one thread."""
    import renzoku_amd as rz
    from renzoku_amd import scenes
    import oracle
    W = H = 512
    sc = scenes.room(W, H, seed=2013, target_tris=args.tris or 200_000)
    ctx = rz.Ctx(0)
    scene = rz.Scene(ctx, sc)
    px, py = np.meshgrid(np.arange(W, dtype=np.int32), np.arange(H, dtype=np.int32))
    _, _, recv, m, _ = scene.camera_primary(sc.camera, px.ravel(), py.ravel())
    keep = m >= 0                         # a hit with a material (lightslice.cc:270)
    diag = scene_diagonal(sc)
    pts = rz.map_vec6(ctx, scene, recv[keep], m[keep], diag)
    n = len(pts)
    dptr = C.c_void_p()
    ctx.check(ctx.lib.rzk_device_alloc(ctx.h, pts.nbytes, C.byref(dptr)), "rzk_device_alloc")
    ctx.check(ctx.lib.rzk_copy_h2d(ctx.h, dptr, pts.ctypes.data_as(C.c_void_p), pts.nbytes), "rzk_copy_h2d")
    for i in range(args.warmup):
        rz.KdSlices(ctx, dptr.value, 256, 2013, n=n).close()
    ms = []
    for i in range(args.steps):
        t0 = time.perf_counter()
        kd = rz.KdSlices(ctx, dptr.value, 256, 2013, n=n)
        ms.append((time.perf_counter() - t0) * 1e3)
        log(f"[bench] step {i + 1}/{args.steps}: {ms[-1]:.2f} ms")
        if i + 1 < args.steps:
            kd.close()
    ms_per_step = float(np.mean(ms))
    depths, mm = 0, n
    while mm >= 2:
        depths, mm = depths + 1, (mm + 1) // 2
    gbytes = 36.0 * n * depths
    achieved = gbytes / (ms_per_step / 1e3) / 1e9
    out = {"metric": "LightSlice kd-tree slicing, gather points/s", "value": round(n / (ms_per_step / 1e3) / 1e6, 3),
           "unit": "Mpoints/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic (seeded room scene)",
           "config": {"workload": "kd-tree slicing room 200k tris 512x512 gather points, slices of 256",
                      "points": n, "slices": kd.S, "depths": depths, "parallelism": "single GPU"},
           "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                        "kernel": "rzk_kdslices_build_dev (k_kd_big_*, k_kd_level_lds)",
                        "note": "step time includes the build's own allocations and the host checks of the "
                                "large-node rounds"}}
    if not args.no_cpu:
        t0 = time.perf_counter()
        o = oracle.KdSlices(pts, 256, 2013)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(n / dt / 1e6, 4), "unit": "Mpoints/s", "cores": 1, "kind": "port",
                               "sample": f"oracle build of the same {n} points on one thread, {dt:.2f} s",
                               "same_slices": bool(np.array_equal(kd.get("order"), o.order)
                                                   and np.array_equal(kd.get("slices"), o.slices))}
    kd.close()
    ctx.lib.rzk_device_free(ctx.h, dptr)
    print(json.dumps(out), flush=True)


def c3_scene_and_guides(args, levels):
    """the 1M-triangle instanced scene at 1280x720 with one LightSlice guide per
    VPL count in `levels` (16x16-pixel slices, 1,000 clusters each)"""
    import renzoku_amd as rz
    from renzoku_amd import scenes, vplgen
    W, H = 1280, 720
    t0 = time.perf_counter()
    sc = scenes.instanced_meshes(W, H, seed=2013, target_tris=args.tris or 1_000_000)
    ctx = rz.Ctx(0)
    scene = rz.Scene(ctx, sc)
    log(f"[bench] scene: {sc.ntri} triangles, BVH depth {scene.info().depth} ({time.perf_counter() - t0:.1f} s)")
    tracer, filler = vplgen.gpu_tracer(scene)
    guides = []
    for nv in levels:
        vpl, vmat, n_emitted = vplgen.generate_vpls(sc, tracer, filler, nv, seed=2013, batch=65536)
        vp = rz.Vpls(ctx, vpl, vmat)
        ls = rz.LightSlice(ctx, scene, vp, width=W, height=H, tile=16, num_clusters=min(1000, nv // 4),
                           num_neighbor_slices=16, map_resolution=32, n_emitted=n_emitted, seed=2013)
        t1 = time.perf_counter()
        ls.initialize(sc.camera)
        ctx.synchronize()
        log(f"[bench] guide for {nv} VPLs: {ls.info().num_clusters} clusters, {(time.perf_counter() - t1) * 1e3:.0f} ms")
        guides.append((nv, vp, ls, ls.guide(sc.camera)))
    return sc, ctx, scene, guides


def run_batched(args):
    """256 render jobs of uneven size on the 1M-triangle scene: windows 32..256 px
    a side, 1-64 spp, guides from 4k-100k VPLs; all jobs in one persistent,
    graph-launched kernel with the autotuned launch shape, against the plain
    per-request drop-in (one rzk_lightslice_render_guided per job)."""
    import renzoku_amd as rz
    levels = [4096, 8192, 16384, 32768, 65536, 100_000]
    sc, ctx, scene, guides = c3_scene_and_guides(args, levels)
    rng = np.random.default_rng(2013)
    specs = []
    for i in range(args.jobs):
        w, h = int(rng.integers(32, 257)), int(rng.integers(32, 257))
        nv, vp, ls, g = guides[int(rng.integers(0, len(guides)))]
        specs.append(dict(guide=g, ls=ls, vpls=nv, window=(int(rng.integers(0, 1280 - w + 1)),
                                                          int(rng.integers(0, 720 - h + 1)), w, h),
                          spp=int(rng.integers(1, 65)), max_bounce=5, direct_lighting=0, guide_mode=1,
                          seed=int(rng.integers(0, 2 ** 31))))
    paths = sum(s["window"][2] * s["window"][3] * s["spp"] for s in specs)
    log(f"[bench] {len(specs)} jobs, {paths / 1e6:.1f} M paths")
    b = rz.Batch(ctx, scene)
    tune = [dict(s, spp=max(1, s["spp"] // 4)) for s in specs[:: max(1, len(specs) // 48)]]
    t0 = time.perf_counter()
    best, tried = b.autotune(tune)
    log(f"[bench] autotune: {len(tried)} candidates in {time.perf_counter() - t0:.1f} s -> {best}")
    _, rays, _ = b.render(specs, to_host=False, count_rays=True)
    ms = []
    for i in range(args.steps):
        t0 = time.perf_counter()
        _, _, dev_ms = b.render(specs, to_host=False)
        ms.append((time.perf_counter() - t0) * 1e3)
        log(f"[bench] batched step {i + 1}/{args.steps}: {ms[-1]:.1f} ms (device {dev_ms:.1f} ms)")
    batched_ms = float(np.mean(ms))
    # plain per-request drop-in: one launch per job, back to back
    ctx.synchronize()
    t0 = time.perf_counter()
    for k, s in enumerate(specs):
        s["ls"].render_guided(sc.camera, spp=s["spp"], max_bounce=5, direct_lighting=0, guide_mode=1, seed=s["seed"],
                              window=s["window"], to_host=False, count_rays=False)
        if (k + 1) % 64 == 0:
            ctx.synchronize()
            log(f"[bench] plain: {k + 1}/{len(specs)} jobs, {time.perf_counter() - t0:.1f} s")
    ctx.synchronize()
    plain_ms = (time.perf_counter() - t0) * 1e3
    # equal accuracy: the batched images are the plain ones, bit for bit (checked on a sample)
    imgs, _, _ = b.render(specs, to_host=True)
    checked = 0
    for k in range(0, len(specs), max(1, len(specs) // 8)):
        s = specs[k]
        ref, _ = s["ls"].render_guided(sc.camera, spp=s["spp"], max_bounce=5, direct_lighting=0, guide_mode=1,
                                       seed=s["seed"], window=s["window"], count_rays=False)
        x0, y0, w, h = s["window"]
        assert np.array_equal(imgs[k].view(np.uint32), ref[y0:y0 + h, x0:x0 + w].view(np.uint32)), k
        checked += 1
    nrays = rays["closest_rays"] + rays["shadow_rays"]
    out = {"metric": "batched image-requests/s (guided path tracing jobs)",
           "value": round(len(specs) / (batched_ms / 1e3), 3), "unit": "requests/s", "n_gpus": 1,
           "steps": args.steps, "warmup": 1, "ms_per_step": round(batched_ms, 1), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded instanced-mesh scene, seeded light-traced VPLs, seeded job mix)",
           "config": {"workload": f"batched {len(specs)} guided path tracing jobs on 1M tris, windows 32..256 px, "
                                  "1-64 spp, depth 5, guides from 4k-100k VPLs", "triangles": int(sc.ntri),
                      "jobs": len(specs), "paths": int(paths), "vpl_levels": levels, "parallelism": "single GPU"},
           "plain_requests_per_s": round(len(specs) / (plain_ms / 1e3), 3), "plain_ms": round(plain_ms, 1),
           "speedup_vs_plain": round(plain_ms / batched_ms, 2), "mrays_per_s": round(nrays / batched_ms / 1e3, 1),
           "rays": rays, "autotuned": best, "autotune_trials": [[c, round(t, 2)] for c, t in tried],
           "accuracy": f"bit-equal to the plain render on {checked} sampled jobs"}
    print(json.dumps(out), flush=True)


def random_camera(rng, w, h):
    """a seeded camera pose inside the instanced-mesh room (2000 x 600 x 2000)"""
    import math
    from renzoku_amd import scenes
    eye = (rng.uniform(200, 1800), rng.uniform(120, 480), rng.uniform(100, 900))
    look = (rng.uniform(200, 1800), rng.uniform(0, 300), rng.uniform(1000, 1900))
    return scenes.camera(eye, look, (0, 1, 0), float(rng.uniform(45, 75)), w, h)


def run_served(args):
    """Poisson arrivals of render requests (square images 128..1024 px a side,
    1-32 spp, depth 5, random camera poses) against a 2M-triangle scene with
    200k VPLs whose 1,000 clusters stay resident; the server admits by memory
    budget and batches continuously.  The same requests, arrival times and
    images are then served by the plain per-request drop-in (guide build +
    render_guided, one request at a time, first come first served)."""
    import threading
    import renzoku_amd as rz
    from renzoku_amd import scenes, vplgen
    t0 = time.perf_counter()
    sc = scenes.instanced_meshes(1280, 720, seed=2013, target_tris=args.tris or 2_000_000)
    ctx = rz.Ctx(0)
    scene = rz.Scene(ctx, sc)
    log(f"[bench] scene: {sc.ntri} triangles ({time.perf_counter() - t0:.1f} s)")
    tracer, filler = vplgen.gpu_tracer(scene)
    vpl, vmat, n_emitted = vplgen.generate_vpls(sc, tracer, filler, args.vpls or 200_000, seed=2013, batch=65536)
    vp = rz.Vpls(ctx, vpl, vmat)
    ls = rz.LightSlice(ctx, scene, vp, width=1280, height=720, tile=32, num_clusters=1000, num_neighbor_slices=16,
                       n_emitted=n_emitted, seed=2013)
    ls.initialize(sc.camera)
    clusters = ls.clusters()
    gparams = dict(tile=16, map_resolution=32, map_epsilon=1e-3, clamp_mode=rz._lib.CLAMP_FORM_FACTOR,
                   incident_clamp=0.01, tmin=1e-3, map_gate=0)
    log(f"[bench] {len(vmat)} VPLs, {ls.info().num_clusters} resident clusters")
    rng = np.random.default_rng(args.seed)
    reqs = []
    for i in range(args.requests):
        side = int(rng.integers(128, 1025))
        reqs.append((random_camera(rng, side, side), dict(spp=int(rng.integers(1, 33)), max_bounce=5,
                                                          direct_lighting=0, guide_mode=1,
                                                          seed=int(rng.integers(0, 2 ** 31)))))
    arrivals = np.cumsum(rng.exponential(1.0 / args.rate, len(reqs)))
    paths = sum(int(c[17]) * int(c[18]) * kw["spp"] for c, kw in reqs)
    log(f"[bench] {len(reqs)} requests, {paths / 1e6:.1f} M paths, arrivals over {arrivals[-1]:.2f} s "
        f"at {args.rate}/s")
    # warm the kernels once
    g = rz.Guide.create(ctx, scene, clusters, reqs[0][0], gparams)
    g.render(spp=1, max_bounce=5, window=(0, 0, 64, 64))
    del g
    srv = rz.Server(0, scene, clusters, memory_budget=args.budget_gb << 30, step_iterations=args.step_iterations,
                    guide=gparams)
    srv.wait(srv.submit(reqs[0][0], spp=1, max_bounce=1))
    served = [None] * len(reqs)

    def submitter():
        start = time.perf_counter()
        for i, (cam, kw) in enumerate(reqs):
            dt = arrivals[i] - (time.perf_counter() - start)
            if dt > 0:
                time.sleep(dt)
            served[i] = srv.submit(cam, **kw)

    th = threading.Thread(target=submitter)
    th.start()
    th.join()
    s_imgs, s_lat, s_fin, s_sub = [], [], [], []
    for i, t in enumerate(served):
        img, st = srv.wait(t)
        s_imgs.append(img)
        s_lat.append(st["finished"] - st["submitted"])
        s_fin.append(st["finished"])
        s_sub.append(st["submitted"])
        if (i + 1) % 16 == 0:
            log(f"[bench] served {i + 1}/{len(reqs)}")
    sstats = srv.stats()
    served_span = max(s_fin) - min(s_sub)
    srv.close()
    # plain: first come first served, one request at a time, same arrival times
    p_lat = []
    start = time.perf_counter()
    busy_until = 0.0
    mismatches = 0
    for i, (cam, kw) in enumerate(reqs):
        now = time.perf_counter() - start
        if now < arrivals[i]:
            time.sleep(arrivals[i] - now)
        g = rz.Guide.create(ctx, scene, clusters, cam, gparams)
        img, _ = g.render(**kw)
        del g
        end = time.perf_counter() - start
        p_lat.append(end - arrivals[i])
        busy_until = end
        mismatches += int(not np.array_equal(img.view(np.uint32), s_imgs[i].view(np.uint32)))
        if (i + 1) % 16 == 0:
            log(f"[bench] plain {i + 1}/{len(reqs)}: {end:.1f} s")
    plain_span = busy_until - arrivals[0]
    pct = lambda a, q: round(float(np.percentile(np.asarray(a) * 1e3, q)), 1)
    out = {"metric": "served image-requests/s with p50/p99 latency",
           "value": round(len(reqs) / served_span, 3), "unit": "requests/s", "n_gpus": 1,
           "steps": len(reqs), "warmup": 1, "ms_per_step": round(served_span * 1e3 / len(reqs), 1),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded instanced-mesh scene, seeded VPLs, seeded Poisson request stream)",
           "config": {"workload": f"served guided path tracing, Poisson {args.rate}/s, images 128..1024 px square, "
                                  "1-32 spp, depth 5, random cameras, 2M tris, 200k VPLs, 1000 resident clusters",
                      "triangles": int(sc.ntri), "vpls": int(len(vmat)), "requests": len(reqs),
                      "paths": int(paths), "rate_per_s": args.rate, "memory_budget_gb": args.budget_gb,
                      "step_iterations": args.step_iterations, "parallelism": "single GPU"},
           "latency_ms": {"p50": pct(s_lat, 50), "p99": pct(s_lat, 99), "max": pct(s_lat, 100)},
           "plain_requests_per_s": round(len(reqs) / plain_span, 3),
           "plain_latency_ms": {"p50": pct(p_lat, 50), "p99": pct(p_lat, 99), "max": pct(p_lat, 100)},
           "speedup_vs_plain": round(plain_span / served_span, 2),
           "paths_per_s": round(paths / served_span, 1),
           "server": {k: sstats[k] for k in ("rounds", "items", "peak_bytes", "busy_seconds")},
           "accuracy": f"{len(reqs) - mismatches} of {len(reqs)} images bit-equal to the plain drop-in"}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="lightslice", choices=["lightslice", "guided", "batched", "served", "refine", "vplgen",
                                                                 "kdslice"],
                    help="lightslice: the 200k-triangle LightSlice gather step (the headline line); "
                         "guided: guided path tracing of the 1M-triangle scene; batched: 256 uneven jobs")
    ap.add_argument("--jobs", type=int, default=256, help="batched workload: number of jobs")
    ap.add_argument("--requests", type=int, default=64, help="served workload: number of requests")
    ap.add_argument("--rate", type=float, default=50.0, help="served workload: Poisson arrival rate (1/s)")
    ap.add_argument("--budget-gb", type=int, default=16, help="served workload: memory budget (GiB)")
    ap.add_argument("--step-iterations", type=int, default=2,
                    help="served workload: engine iterations between looks at finished requests")
    ap.add_argument("--seed", type=int, default=2013, help="served workload: request stream seed")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--tris", type=int, default=0, help="triangle count (default: the workload's)")
    ap.add_argument("--kd", type=int, default=0,
                    help="lightslice workload with the reference's kd-tree slices of at most this many points "
                         "(0: 32x32 pixel tiles)")
    ap.add_argument("--vpls", type=int, default=0, help="VPL count (default: the workload's)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="after the run, copy MIB MiB device-to-device once (a known-size transfer for "
                         "scaling FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU sharding is a "
              f"design, not a measured configuration.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "guided":
        return run_guided(args)
    if args.workload == "batched":
        return run_batched(args)
    if args.workload == "served":
        return run_served(args)
    if args.workload == "refine":
        return run_refine(args)
    if args.workload == "vplgen":
        return run_vplgen(args)
    if args.workload == "kdslice":
        return run_kdslice(args)

    import renzoku_amd as rz
    from renzoku_amd import scenes, vplgen

    W = H = 512
    params = dict(width=W, height=H, tile=32, num_clusters=300, num_neighbor_slices=16, map_resolution=32,
                  map_epsilon=1e-3, clamp_mode=rz._lib.CLAMP_FORM_FACTOR, clamp=0.01, incident_clamp=0.01,
                  tmin=1e-3, seed=2013, map_gate=0, kd_max_slice_size=args.kd)
    t0 = time.perf_counter()
    sc = scenes.room(W, H, seed=2013, target_tris=args.tris or 200_000)
    log(f"[bench] scene: {sc.ntri} triangles ({time.perf_counter() - t0:.1f} s)")
    ctx = rz.Ctx(0)
    scene = rz.Scene(ctx, sc)
    inf = scene.info()
    log(f"[bench] BVH: {inf.num_interior_nodes} interior nodes, depth {inf.depth}")
    tracer, filler = vplgen.gpu_tracer(scene)
    vpl, vmat, n_emitted = vplgen.generate_vpls(sc, tracer, filler, args.vpls or 65_536, seed=2013, batch=65536)
    params["n_emitted"] = n_emitted
    vpls = rz.Vpls(ctx, vpl, vmat)
    log(f"[bench] VPLs: {len(vmat)} from {n_emitted:.0f} light samples")
    ls = rz.LightSlice(ctx, scene, vpls, **params)

    ev = HipEvents()
    stream = ctx.stream
    e0, e1 = ev.create(), ev.create()
    for i in range(args.warmup):
        ls.initialize(sc.camera)
        ls.gather(to_host=False)
        ctx.synchronize()
        log(f"[bench] warmup {i + 1}/{args.warmup} done")
    work = ls.count_work()
    lsinf = ls.info()
    log(f"[bench] work per step: {work}")

    ctx.synchronize()
    gather_ms = []
    t_start = time.perf_counter()
    for i in range(args.steps):
        ls.initialize(sc.camera)
        ev.record(e0, stream)
        ls.gather(to_host=False)
        ev.record(e1, stream)
        gather_ms.append(ev.elapsed_ms(e0, e1))
    ctx.synchronize()
    total_s = time.perf_counter() - t_start
    ms_per_step = total_s * 1e3 / args.steps
    g_ms = float(np.mean(gather_ms))
    log(f"[bench] {args.steps} steps: {ms_per_step:.2f} ms/step, gather kernel {g_ms:.3f} ms")

    shadow_rays = work["reduced_pairs"] + work["gather_pairs"]
    value = shadow_rays / (ms_per_step / 1e3) / 1e6
    gbytes = (work["gather_pairs"] * (RAY_IN_BYTES + SHADOW_OUT_BYTES) + work["gather_nodes"] * NODE4_BYTES +
              work["gather_tris"] * TRI_BYTES)
    achieved = gbytes / (g_ms / 1e3) / 1e9
    out = {
        "metric": "shadow Mrays/s + VPL-gather Mpairs/s; image req/s + p50/p99 latency vs plain",
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded room scene, seeded light-traced VPLs)",
        "config": {"workload": ("lightslice-gather room 200k tris 512x512 65536 VPLs 32x32 slices 300 clusters"
                                if not args.kd else
                                f"lightslice-gather room 200k tris 512x512 65536 VPLs kd-tree slices of {args.kd} "
                                "points 300 clusters"),
                   "triangles": int(sc.ntri), "width": W, "height": H, "vpls": int(len(vmat)),
                   "slices": int(lsinf.num_slices), "clusters": int(lsinf.num_clusters),
                   "parallelism": "single GPU"},
        "gather_mpairs_per_s": round(work["gather_pairs"] / (g_ms / 1e3) / 1e6, 3),
        "step_mpairs_per_s": round(shadow_rays / (ms_per_step / 1e3) / 1e6, 3),
        "gather_kernel_ms": round(g_ms, 4),
        "work": work,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": GATHER_HBM_BYTES,
                     "kernel": "k_slice_visibility + k_slice_shade (the gather)",
                     "bytes_per_launch": int(gbytes),
                     "note": "achieved counts every lane's node visits (128 B per 4-wide node, 48 B per triangle "
                             "test, 36 B of segment I/O); a wave fetches each node once and L2/MALL serve it, so "
                             "measured HBM traffic per launch (traffic) is far below; the visibility kernel is "
                             "vector-issue bound (VALUBusy reads 108.6%, saturated, rocprofv3 --pmc)"},
    }
    if not args.no_cpu:
        reps = ls.get("reps", np.int32, lsinf.num_clusters)
        rep_w = ls.get("rep_w", np.float32, lsinf.num_clusters)
        log("[bench] CPU baseline ...")
        out["cpu_baseline"] = cpu_baseline(sc, ls, vpl, vmat, reps, rep_w, params)
        log(f"[bench] CPU baseline: {out['cpu_baseline']}")
    print(json.dumps(out), flush=True)
    if args.copy_calibration > 0:
        nb = args.copy_calibration << 20
        src, dst = C.c_void_p(), C.c_void_p()
        lib = ctx.lib
        ctx.check(lib.rzk_device_alloc(ctx.h, nb, C.byref(src)), "alloc")
        ctx.check(lib.rzk_device_alloc(ctx.h, nb, C.byref(dst)), "alloc")
        hip = C.CDLL("libamdhip64.so")
        hip.hipMemcpyDtoD.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        assert hip.hipMemcpyDtoD(dst, src, nb) == 0
        ctx.synchronize()
        log(f"[bench] copy calibration: {args.copy_calibration} MiB device-to-device")
        lib.rzk_device_free(ctx.h, src)
        lib.rzk_device_free(ctx.h, dst)


if __name__ == "__main__":
    main()
