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
peak with the algorithmic bytes of SURVEY's traversal model.  cpu_baseline is
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
           "guide_build_ms": round(init_ms, 1)}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="lightslice", choices=["lightslice", "guided"],
                    help="lightslice: the 200k-triangle LightSlice gather step (the headline line); "
                         "guided: guided path tracing of the 1M-triangle scene")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--tris", type=int, default=0, help="triangle count (default: the workload's)")
    ap.add_argument("--vpls", type=int, default=0, help="VPL count (default: the workload's)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="after the run, copy MIB MiB device-to-device once (a known-size transfer for "
                         "scaling FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU sharding is a design "
              f"described in DESIGN.md, not a measured configuration.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "guided":
        return run_guided(args)

    import renzoku_amd as rz
    from renzoku_amd import scenes, vplgen

    W = H = 512
    params = dict(width=W, height=H, tile=32, num_clusters=300, num_neighbor_slices=16, map_resolution=32,
                  map_epsilon=1e-3, clamp_mode=rz._lib.CLAMP_FORM_FACTOR, clamp=0.01, incident_clamp=0.01,
                  tmin=1e-3, seed=2013, map_gate=0)
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
        "config": {"workload": "lightslice-gather room 200k tris 512x512 65536 VPLs 32x32 slices 300 clusters",
                   "triangles": int(sc.ntri), "width": W, "height": H, "vpls": int(len(vmat)),
                   "slices": int(lsinf.num_slices), "clusters": int(lsinf.num_clusters),
                   "parallelism": "single GPU"},
        "gather_mpairs_per_s": round(work["gather_pairs"] / (g_ms / 1e3) / 1e6, 3),
        "step_mpairs_per_s": round(shadow_rays / (ms_per_step / 1e3) / 1e6, 3),
        "gather_kernel_ms": round(g_ms, 4),
        "work": work,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_slice_gather",
                     "bytes_per_launch": int(gbytes)},
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
