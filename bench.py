"""Benchmark of the hot path on one MI355X.

Default run (no --workload): the headline LightSlice step, then compact legs
for the other configurations, then live rocprofv3 counter passes; one JSON line.

Headline (C2): seeded room of ~200k triangles (tessellated boxes and UV
spheres, half Lambertian, half Phong), 512x512 pixels, 65,536 bounce-1 VPLs,
32x32-pixel slices (256 rows), 300 cluster centres.  A step is the whole hot
path with the scene BVH and the VPL set resident in HBM: primary closest hits
for every pixel, the reduced matrix (256 x 65,536 shadow segments + incident
radiance), alphas, centre sampling, assignment, representatives, per-slice
cluster pdfs and luminance maps, and the final per-pixel gather over the
cluster representatives (one shadow segment per (pixel, representative) pair).
value = shadow Mrays/s over the step (reduced-matrix + gather segments / step
time).

Roofline: the gather (k_slice_visibility + k_slice_shade) is timed with HIP
events on the context stream and is bound by vector issue, so its roofline is
SQ_INSTS_VALU x 64 lanes per launch (a live rocprofv3 counter pass of the same
build and inputs) / launch time against 256 CUs x 64 lanes x 2.4 GHz; its HBM
traffic is FETCH_SIZE + WRITE_SIZE of the same launch, each scaled by what the
counter reads for a 512 MiB device copy.  cpu_baseline: the whole step on the
oracle (a C restatement of the reference) on every usable host core.

Legs: C1 plain VPL render (Cornell box, 256x256, 1,024 VPLs), C3 guided path
tracing (1M triangles, 1280x720, 16 spp, depth 5) with an HBM roofline over the
traversal byte model, C4 batched jobs and C5 served requests against the plain
per-request drop-in.
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
# MI355X fp32 vector peak as published: 157.3 TFLOP/s, i.e. 78.65 T FMA lane-operations per second
# (256 CUs x 4 SIMDs x 32 lanes x 2.4 GHz).  The probes below measure what instruction classes reach:
# with every operand in a vector register about 0.92 of it (multiply + add), with a scalar-register
# operand (a wave-uniform value, as the wave-coherent walks' node planes) about half of it.
VALU_LANE_OPS_PEAK = 157.3e12 / 2
# counter passes of the live PMC leg (one rocprofv3 run each; at most 8 SQ / 4 TCC counters per pass)
PMC_PASSES = [["SQ_INSTS_VALU", "SQ_WAVES"], ["VALUUtilization", "VALUBusy"], ["FETCH_SIZE"], ["WRITE_SIZE"]]
# kernels whose per-launch counters the PMC child reports (substring of the kernel name)
# (substring of the kernel name, which dispatch): the PMC child runs the C2 step first (its reduced matrix
# is the first one), then the C1 gather, then C3 (whose guide build has a reduced matrix of its own), then
# the calibration copy (the last copy dispatch)
PMC_KERNELS = {"slice_visibility": ("k_slice_visibility", "first"), "slice_shade": ("k_slice_shade", "first"),
               "reduced_matrix": ("k_reduced_matrix", "first"), "visibility": ("k_visibility", "first"),
               "gather_shade": ("k_gather_shade", "first"), "guided_pt": ("k_guided_pt", "first"),
               "copy": ("copyBuffer", "last"), "probe_fma": ("k_valu_probe(", "last"),
               "probe_slab": ("k_valu_probe_slab(", "last"), "probe_muladd_vv": ("k_valu_probe_vv(", "last"),
               "probe_slab_vv": ("k_valu_probe_slab_vv(", "last"), "probe_fma_vv": ("k_valu_probe_fma_vv(", "last")}
# vector-issue probes (renzoku_amd/libprobe.so kind): FMA and the slab-test mix with scalar-register
# operands, then multiply + add, the slab mix and FMA with vector-register operands only
PROBES = {"probe_fma": 0, "probe_slab": 1, "probe_muladd_vv": 4, "probe_slab_vv": 5, "probe_fma_vv": 6}
# the vector-issue probe launch (renzoku_amd/libprobe.so): 4,096 blocks of 256 lanes (16 waves per SIMD),
# 16 independent FMA chains per lane, about 8 ms
PROBE_BLOCKS, PROBE_ITERS = 4096, 8192
PMC_COPY_MIB = 512
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


def host_cpu():
    """(model name, usable cores): the CPU affinity set, capped by a cgroup CPU quota if there is one"""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            quota, period = f.read().split()[:2]
        if quota != "max":
            cores = max(1, min(cores, int(int(quota) // int(period))))
    except (OSError, ValueError):
        pass
    return model, cores


def _pmc_parse(outdir):
    """{kernel key: {counter: value}} for the last dispatch of each PMC_KERNELS key, from the
    counter-collection CSVs rocprofv3 wrote under outdir (values of one dispatch summed)"""
    import csv
    import glob
    rows = []
    for path in glob.glob(os.path.join(outdir, "**", "*counter_collection*.csv"), recursive=True):
        with open(path, newline="") as f:
            rows.extend(csv.DictReader(f))
    if not rows:
        return {}
    keys = {k.lower(): k for k in rows[0].keys()}
    kn, cn, cv = keys.get("kernel_name"), keys.get("counter_name"), keys.get("counter_value")
    di = keys.get("dispatch_id") or keys.get("correlation_id")
    ts, te = keys.get("start_timestamp"), keys.get("end_timestamp")
    disp = {}
    for r in rows:
        d = disp.setdefault(int(r[di]), {"name": r[kn], "c": {}})
        d["c"][r[cn]] = d["c"].get(r[cn], 0.0) + float(r[cv])
        if ts and te and r.get(ts) and r.get(te):
            d["ns"] = float(r[te]) - float(r[ts])
    out = {}
    for key, (sub, which) in PMC_KERNELS.items():
        ids = [i for i, d in disp.items() if sub in d["name"]]
        if ids:
            d = disp[min(ids) if which == "first" else max(ids)]
            out[key] = dict(d["c"])
            if "ns" in d:
                out[key]["_ns"] = d["ns"]
    return out


def pmc_collect(child_args, deadline):
    """Runs this file as a child under `rocprofv3 --pmc`, one counter pass at a time (PMC_PASSES), each
    child launching every measured kernel once and then copying PMC_COPY_MIB MiB device to device (the
    known-size transfer that scales FETCH_SIZE / WRITE_SIZE).  Returns ({kernel key: {counter: value}},
    note).  A pass that overruns its time limit is killed with its process group."""
    import shutil
    import signal
    import subprocess
    import tempfile
    exe = shutil.which("rocprofv3")
    if not exe:
        return {}, "rocprofv3 not found"
    merged, notes = {}, []
    for counters in PMC_PASSES:
        left = deadline - time.perf_counter()
        if left < 60:
            notes.append(f"pass {'+'.join(counters)} skipped (time budget)")
            continue
        out = tempfile.mkdtemp(prefix="rzk_pmc_", dir=os.environ.get("TMPDIR", "/tmp"))
        cmd = [exe, "--pmc", *counters, "--output-format", "csv", "-d", out, "-o", "pmc", "--",
               sys.executable, "-u", os.path.abspath(__file__), "--pmc-child", *child_args]
        log(f"[bench] PMC pass {'+'.join(counters)} ...")
        t0 = time.perf_counter()
        proc = subprocess.Popen(cmd, stdout=subprocess.DEVNULL, start_new_session=True, cwd=out)
        try:
            rc = proc.wait(timeout=min(180.0, left))
        except subprocess.TimeoutExpired:
            os.killpg(proc.pid, signal.SIGKILL)
            proc.wait()
            notes.append(f"pass {'+'.join(counters)} killed at its time limit")
            shutil.rmtree(out, ignore_errors=True)
            break
        got = _pmc_parse(out)
        shutil.rmtree(out, ignore_errors=True)
        if rc != 0 and not got:
            notes.append(f"pass {'+'.join(counters)} exited {rc}")
            continue
        for k, c in got.items():
            merged.setdefault(k, {}).update(c)
        log(f"[bench] PMC pass {'+'.join(counters)}: {len(got)} kernels in {time.perf_counter() - t0:.1f} s")
    return merged, "; ".join(notes)


def pmc_traffic(pmc, keys):
    """HBM bytes of one launch of the kernels `keys`: FETCH_SIZE and WRITE_SIZE (KiB per dispatch), each
    scaled by what it reads for the PMC_COPY_MIB MiB copy; None without both passes"""
    cp = pmc.get("copy", {})
    if "FETCH_SIZE" not in cp or "WRITE_SIZE" not in cp or cp["FETCH_SIZE"] <= 0 or cp["WRITE_SIZE"] <= 0:
        return None, None
    kib = PMC_COPY_MIB * 1024.0
    ff, fw = kib / cp["FETCH_SIZE"], kib / cp["WRITE_SIZE"]
    total = 0.0
    for k in keys:
        c = pmc.get(k, {})
        if "FETCH_SIZE" not in c or "WRITE_SIZE" not in c:
            return None, None
        total += (c["FETCH_SIZE"] * ff + c["WRITE_SIZE"] * fw) * 1024.0
    return total, {"fetch_scale": round(ff, 4), "write_scale": round(fw, 4)}


def probe_ms():
    """{probe key: launch time} of the vector-issue probes (libprobe.so: independent FMA chains, and the
    slab-test mix of subtract, multiply, min, max, compare, select); {} without the library"""
    path = os.path.join(ROOT, "renzoku_amd", "libprobe.so")
    if not os.path.exists(path):
        return {}
    lib = C.CDLL(path)
    lib.rzk_probe_valu.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float)]
    out = {}
    for key, kind in PROBES.items():
        ms = C.c_float()
        best = None
        for _ in range(3):
            if lib.rzk_probe_valu(kind, PROBE_BLOCKS, PROBE_ITERS, C.byref(ms)) != 0:
                break
            best = ms.value if best is None else min(best, ms.value)
        if best:
            out[key] = best
    return out


def valu_peak(pmc, probe_launch_ms):
    """(VALU peak in lane-ops/s, how it was obtained, {probe: measured lane-ops/s}): the published fp32
    vector peak, with each probe's SQ_INSTS_VALU x 64 over its launch time when the probes ran in the
    counter pass"""
    rates, desc = {}, []
    for key in PROBES:
        insts, ms = pmc.get(key, {}).get("SQ_INSTS_VALU"), probe_launch_ms.get(key)
        if insts and ms:
            rates[key] = insts * 64.0 / (ms / 1e3)
            desc.append(f"{key}: {int(insts)} wave instructions in {ms:.4f} ms = "
                        f"{rates[key] / 1e12:.2f} T ({rates[key] / VALU_LANE_OPS_PEAK:.3f} of the peak)")
    src = "published: MI355X fp32 vector 157.3 TFLOP/s = 78.65 T FMA lane-ops/s"
    if desc:
        src += "; measured probes: " + "; ".join(desc)
    return VALU_LANE_OPS_PEAK, src, rates


def valu_roofline(pmc, keys, launch_ms, kernel_name, byte_model=None, peak=None):
    """roofline of a vector-issue-bound launch: SQ_INSTS_VALU x 64 lanes per launch / launch time against
    the sustained issue peak (valu_peak); measured HBM traffic beside it"""
    peak_rate, peak_src, probe_rates = peak if peak is not None else (VALU_LANE_OPS_PEAK, "published", {})
    insts = [pmc.get(k, {}).get("SQ_INSTS_VALU") for k in keys]
    traffic, scales = pmc_traffic(pmc, keys)
    r = {"bound": "valu", "unit": "T VALU lane-ops/s", "peak": round(peak_rate / 1e12, 3),
         "achieved": None, "frac": None, "traffic": None if traffic is None else int(traffic),
         "kernel": kernel_name, "launch_ms": round(launch_ms, 4), "peak_source": peak_src}
    if all(v is not None for v in insts):
        ops = sum(insts) * 64.0
        r["achieved"] = round(ops / (launch_ms / 1e3) / 1e12, 3)
        r["frac"] = round(ops / (launch_ms / 1e3) / peak_rate, 4)
        r["sq_insts_valu"] = {k: int(pmc[k]["SQ_INSTS_VALU"]) for k in keys}
        if "probe_slab" in probe_rates:
            # the issue rate of the class these kernels run (box tests whose node planes are
            # wave-uniform scalar-register operands), as measured by the slab-mix probe
            r["scalar_operand_issue"] = round(probe_rates["probe_slab"] / 1e12, 3)
            r["frac_of_scalar_operand_issue"] = round(ops / (launch_ms / 1e3) / probe_rates["probe_slab"], 4)
    for k in keys:
        for c in ("VALUUtilization", "VALUBusy"):
            if c in pmc.get(k, {}):
                r.setdefault(c.lower(), {})[k] = round(pmc[k][c], 2)
    if traffic is not None:
        r["hbm_gbs"] = round(traffic / (launch_ms / 1e3) / 1e9, 2)
        r["hbm_frac"] = round(traffic / (launch_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)
        r["traffic_scales"] = scales
    if byte_model is not None:
        r["byte_model"] = byte_model
    return r


def hbm_roofline(pmc, keys, launch_ms, algo_bytes, kernel_name, note, peak=None):
    """roofline of a traversal launch against HBM peak: algorithmic bytes (the traversal byte model over
    each lane's own counted work) per launch / launch time; measured traffic beside it"""
    traffic, scales = pmc_traffic(pmc, keys)
    achieved = algo_bytes / (launch_ms / 1e3) / 1e9
    r = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None if traffic is None else int(traffic),
         "kernel": kernel_name, "launch_ms": round(launch_ms, 4), "bytes_per_launch": int(algo_bytes), "note": note}
    if traffic is not None:
        r["measured_gbs"] = round(traffic / (launch_ms / 1e3) / 1e9, 2)
        r["traffic_scales"] = scales
    insts = [pmc.get(k, {}).get("SQ_INSTS_VALU") for k in keys]
    if all(v is not None for v in insts):
        r["valu_frac"] = round(sum(insts) * 64.0 / (launch_ms / 1e3) / (peak[0] if peak else VALU_LANE_OPS_PEAK), 4)
        r["valu_peak"] = round((peak[0] if peak else VALU_LANE_OPS_PEAK) / 1e12, 3)
    for k in keys:
        if "VALUUtilization" in pmc.get(k, {}):
            r.setdefault("valuutilization", {})[k] = round(pmc[k]["VALUUtilization"], 2)
    return r


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
    vpl, vmat, n_emitted = vplgen.device_vpls(scene, args.vpls or 100_000, seed=2013)
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


def run_build(args):
    """Scene creation with the trees built on the host (rzk_scene_create: Morton binary tree and binned-SAH
    4-wide tree) against the device build (rzk_scene_create_ex, RZK_SCENE_BUILD_DEVICE: the same binary
    tree, the 4-wide tree collapsed from it) on the 1M- and 2M-triangle instanced scenes; then the guided
    path-tracing frame of the 1M scene (1280x720, 16 spp, depth 5, the same 100k VPLs) on either 4-wide tree."""
    import renzoku_amd as rz
    from renzoku_amd import scenes, vplgen
    ctx = rz.Ctx(0)
    warm = scenes.cornell_box(64, 64)
    rz.Scene(ctx, warm, build="device")
    res = {}
    for ntri in (1_000_000, 2_000_000):
        sc = scenes.instanced_meshes(1280, 720, seed=2013, target_tris=ntri)
        t0 = time.perf_counter()
        hs = rz.Scene(ctx, sc)
        ctx.synchronize()
        host_ms = (time.perf_counter() - t0) * 1e3
        dev_ms = []
        for i in range(max(args.steps, 1)):
            t0 = time.perf_counter()
            ds = rz.Scene(ctx, sc, build="device")
            ctx.synchronize()
            dev_ms.append((time.perf_counter() - t0) * 1e3)
            if i + 1 < args.steps:
                del ds
        hi, di = hs.info(), ds.info()
        same = all(np.array_equal(a.view(np.uint32), b.view(np.uint32))
                   for a, b in zip(hs.export_trees()[:2], ds.export_trees()[:2]))
        res[int(sc.ntri)] = dict(host_ms=round(host_ms, 1), device_ms=round(float(np.median(dev_ms)), 1),
                                 speedup=round(host_ms / float(np.median(dev_ms)), 2),
                                 binary_tree_bit_equal=bool(same), depth=int(hi.depth),
                                 device_bytes_host_build=int(hi.device_bytes),
                                 device_bytes_device_build=int(di.device_bytes))
        log(f"[bench] {sc.ntri} triangles: host build {host_ms:.0f} ms, device build {np.median(dev_ms):.1f} ms "
            f"(binary tree bit-equal: {same})")
        if ntri == 1_000_000:
            sc1, hs1, ds1 = sc, hs, ds
        else:
            del hs, ds
    W, H = 1280, 720
    vpl, vmat, ne = vplgen.device_vpls(hs1, 100_000, seed=2013)
    vpls = rz.Vpls(ctx, vpl, vmat)
    kw = dict(spp=16, max_bounce=5, direct_lighting=0, guide_mode=1, seed=2013, tmin=1e-3)
    frames, imgs = {}, {}
    ev = HipEvents()
    e0, e1 = ev.create(), ev.create()
    for name, scene in (("host", hs1), ("device", ds1)):
        ls = rz.LightSlice(ctx, scene, vpls, width=W, height=H, tile=16, num_clusters=1000, num_neighbor_slices=16,
                           map_resolution=32, n_emitted=ne, seed=2013)
        ls.initialize(sc1.camera)
        imgs[name], _ = ls.render_guided(sc1.camera, count_rays=False, to_host=True, **kw)
        ms = []
        for i in range(args.steps):
            ev.record(e0, ctx.stream)
            ls.render_guided(sc1.camera, count_rays=False, to_host=False, **kw)
            ev.record(e1, ctx.stream)
            ms.append(ev.elapsed_ms(e0, e1))
        frames[name] = round(float(np.mean(ms)), 2)
        log(f"[bench] guided frame on the {name}-built 4-wide tree: {frames[name]:.2f} ms")
        del ls
    a, b = imgs["host"], imgs["device"]
    same_px = float(np.mean(np.all(a.view(np.uint32) == b.view(np.uint32), axis=-1)))
    out = {"metric": "scene build ms (host vs device)", "value": res[max(res)]["device_ms"], "unit": "ms",
           "n_gpus": 1, "steps": args.steps, "warmup": 1, "higher_is_better": False, "dtype": "f32",
           "data": "synthetic (seeded instanced-mesh scenes)",
           "config": {"workload": "scene creation, 1M and 2M triangles; guided frame 1280x720 16 spp depth 5",
                      "parallelism": "single GPU"},
           "builds": res, "guided_frame_ms": frames, "guided_image_equal_pixels": round(same_px, 6)}
    print(json.dumps(out), flush=True)


def run_refine(args):
    """LightSlice per-slice refinement on the 200k-triangle LightSlice workload
    (512x512, 65,536 VPLs, 32x32-pixel slices, 16 neighbour slices, 300 kept
    clusters refined to 600 per slice, the reference driver's 2x rule).  A step
    is one initialize with refinement; the refinement time is that minus an
    initialize without it (same LightSlice inputs).  The refinement reads the
    reference's single MT stream slice after slice, so a slice's offset is
    known only once the slices before it ran at theirs; passes run slices at
    guessed offsets and keep the chain that guessed right; the pass count is
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
    vpl, vmat, n_emitted = vplgen.device_vpls(scene, args.vpls or 65_536, seed=2013)
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


def guides_for(ctx, sc, scene, levels):
    """one LightSlice guide per VPL count in `levels` on a 1280x720 scene (16x16-pixel slices, up to
    1,000 clusters each)"""
    import renzoku_amd as rz
    from renzoku_amd import vplgen
    guides = []
    for nv in levels:
        vpl, vmat, n_emitted = vplgen.device_vpls(scene, nv, seed=2013)
        vp = rz.Vpls(ctx, vpl, vmat)
        ls = rz.LightSlice(ctx, scene, vp, width=1280, height=720, tile=16, num_clusters=min(1000, nv // 4),
                           num_neighbor_slices=16, map_resolution=32, n_emitted=n_emitted, seed=2013)
        t1 = time.perf_counter()
        ls.initialize(sc.camera)
        ctx.synchronize()
        log(f"[bench] guide for {nv} VPLs: {ls.info().num_clusters} clusters, {(time.perf_counter() - t1) * 1e3:.0f} ms")
        guides.append((nv, vp, ls, ls.guide(sc.camera)))
    return guides


BATCH_LEVELS = [4096, 8192, 16384, 32768, 65536, 100_000]


def measure_batched(args, sc, ctx, scene, guides, njobs, tune_max, steps):
    """njobs render jobs of uneven size on a 1280x720 scene: windows 32..256 px a side, 1-64 spp, guides
    from the given VPL levels; all jobs in one persistent, graph-launched kernel with the autotuned launch
    shape, against the plain per-request drop-in (one rzk_lightslice_render_guided per job)."""
    import renzoku_amd as rz
    rng = np.random.default_rng(2013)
    specs = []
    for i in range(njobs):
        w, h = int(rng.integers(32, 257)), int(rng.integers(32, 257))
        nv, vp, ls, g = guides[int(rng.integers(0, len(guides)))]
        specs.append(dict(guide=g, ls=ls, vpls=nv, window=(int(rng.integers(0, 1280 - w + 1)),
                                                          int(rng.integers(0, 720 - h + 1)), w, h),
                          spp=int(rng.integers(1, 65)), max_bounce=5, direct_lighting=0, guide_mode=1,
                          seed=int(rng.integers(0, 2 ** 31))))
    paths = sum(s["window"][2] * s["window"][3] * s["spp"] for s in specs)
    log(f"[bench] {len(specs)} jobs, {paths / 1e6:.1f} M paths")
    b = rz.Batch(ctx, scene)
    best, tried = None, []
    if tune_max > 0:
        tune = [dict(s, spp=max(1, s["spp"] // 4)) for s in specs[:: max(1, len(specs) // 48)]]
        t0 = time.perf_counter()
        best, tried = b.autotune(tune, max_tried=tune_max)
        log(f"[bench] autotune: {len(tried)} candidates in {time.perf_counter() - t0:.1f} s -> {best}")
    _, rays, _ = b.render(specs, to_host=False, count_rays=True)
    ms = []
    for i in range(steps):
        t0 = time.perf_counter()
        _, _, dev_ms = b.render(specs, to_host=False)
        ms.append((time.perf_counter() - t0) * 1e3)
        log(f"[bench] batched step {i + 1}/{steps}: {ms[-1]:.1f} ms (device {dev_ms:.1f} ms)")
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
    return {"metric": "batched image-requests/s (guided path tracing jobs)",
            "value": round(len(specs) / (batched_ms / 1e3), 3), "unit": "requests/s", "n_gpus": 1,
            "steps": steps, "warmup": 1, "ms_per_step": round(batched_ms, 1), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (seeded instanced-mesh scene, seeded light-traced VPLs, seeded job mix)",
            "config": {"workload": f"batched {len(specs)} guided path tracing jobs on {sc.ntri} tris, windows "
                                   "32..256 px, 1-64 spp, depth 5, guides from 4k-100k VPLs", "triangles": int(sc.ntri),
                       "jobs": len(specs), "paths": int(paths), "vpl_levels": [g[0] for g in guides],
                       "parallelism": "single GPU"},
            "plain_requests_per_s": round(len(specs) / (plain_ms / 1e3), 3), "plain_ms": round(plain_ms, 1),
            "speedup_vs_plain": round(plain_ms / batched_ms, 2), "mrays_per_s": round(nrays / batched_ms / 1e3, 1),
            "rays": rays, "autotuned": best, "autotune_trials": [[c, round(t, 2)] for c, t in tried],
            "accuracy": f"bit-equal to the plain render on {checked} sampled jobs", "_paths": int(paths)}


def run_batched(args):
    """256 render jobs of uneven size on the 1M-triangle scene (see measure_batched)"""
    import renzoku_amd as rz
    from renzoku_amd import scenes
    sc = scenes.instanced_meshes(1280, 720, seed=2013, target_tris=args.tris or 1_000_000)
    ctx = rz.Ctx(0)
    scene = rz.Scene(ctx, sc)
    guides = guides_for(ctx, sc, scene, BATCH_LEVELS)
    out = measure_batched(args, sc, ctx, scene, guides, args.jobs, 64, args.steps)
    out.pop("_paths")
    print(json.dumps(out), flush=True)


def random_camera(rng, w, h):
    """a seeded camera pose inside the instanced-mesh room (2000 x 600 x 2000)"""
    import math
    from renzoku_amd import scenes
    eye = (rng.uniform(200, 1800), rng.uniform(120, 480), rng.uniform(100, 900))
    look = (rng.uniform(200, 1800), rng.uniform(0, 300), rng.uniform(1000, 1900))
    return scenes.camera(eye, look, (0, 1, 0), float(rng.uniform(45, 75)), w, h)


def setup_c5(args, ctx):
    """the 2M-triangle served scene with 200k VPLs whose 1,000 clusters stay resident"""
    import renzoku_amd as rz
    from renzoku_amd import scenes, vplgen
    t0 = time.perf_counter()
    sc = scenes.instanced_meshes(1280, 720, seed=2013, target_tris=args.tris5 or 2_000_000)
    t1 = time.perf_counter()
    scene = rz.Scene(ctx, sc)
    host_ms = (time.perf_counter() - t1) * 1e3
    log(f"[bench] C5 scene: {sc.ntri} triangles ({time.perf_counter() - t0:.1f} s)")
    # the same scene with its trees built on the device (second build timed: the first loads the kernels)
    dev_ms = []
    for _ in range(2):
        t1 = time.perf_counter()
        ds = rz.Scene(ctx, sc, build="device")
        ctx.synchronize()
        dev_ms.append((time.perf_counter() - t1) * 1e3)
    same = all(np.array_equal(a.view(np.uint32), b.view(np.uint32))
               for a, b in zip(scene.export_trees()[:2], ds.export_trees()[:2]))
    del ds
    build = {"host_ms": round(host_ms, 1), "device_ms": round(dev_ms[-1], 1),
             "speedup": round(host_ms / dev_ms[-1], 1), "binary_tree_bit_equal": bool(same)}
    log(f"[bench] C5 scene build: host {host_ms:.0f} ms, device {dev_ms[-1]:.1f} ms, bit-equal {same}")
    vpl, vmat, n_emitted = vplgen.device_vpls(scene, args.vpls5 or 200_000, seed=2013)
    vp = rz.Vpls(ctx, vpl, vmat)
    ls = rz.LightSlice(ctx, scene, vp, width=1280, height=720, tile=32, num_clusters=1000, num_neighbor_slices=16,
                       n_emitted=n_emitted, seed=2013)
    ls.initialize(sc.camera)
    clusters = ls.clusters()
    log(f"[bench] C5: {len(vmat)} VPLs, {ls.info().num_clusters} resident clusters")
    return dict(sc=sc, scene=scene, vp=vp, ls=ls, clusters=clusters, nvpl=len(vmat), build=build)


def measure_served(args, ctx, c5, nreq, rate, seed):
    """Poisson arrivals (rate per second) of nreq render requests (square images 128..1024 px a side, 1-32
    spp, depth 5, random camera poses) against the resident scene, VPLs and clusters; the server admits by
    memory budget and batches continuously.  The same requests, arrival times and images are then served
    by the plain per-request drop-in (guide build + render_guided, one request at a time, first come
    first served)."""
    import threading
    import renzoku_amd as rz
    sc, scene, clusters = c5["sc"], c5["scene"], c5["clusters"]
    gparams = dict(tile=16, map_resolution=32, map_epsilon=1e-3, clamp_mode=rz._lib.CLAMP_FORM_FACTOR,
                   incident_clamp=0.01, tmin=1e-3, map_gate=0)
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(nreq):
        side = int(rng.integers(128, 1025))
        reqs.append((random_camera(rng, side, side), dict(spp=int(rng.integers(1, 33)), max_bounce=5,
                                                          direct_lighting=0, guide_mode=1,
                                                          seed=int(rng.integers(0, 2 ** 31)))))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, len(reqs)))
    paths = sum(int(c[17]) * int(c[18]) * kw["spp"] for c, kw in reqs)
    log(f"[bench] {len(reqs)} requests, {paths / 1e6:.1f} M paths, arrivals over {arrivals[-1]:.2f} s at {rate}/s")
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
    pct = lambda a, q: round(float(np.percentile(np.asarray(a) * 1e3, q)), 1)  # noqa: E731
    return {"metric": "served image-requests/s with p50/p99 latency",
            "value": round(len(reqs) / served_span, 3), "unit": "requests/s", "n_gpus": 1,
            "steps": len(reqs), "warmup": 1, "ms_per_step": round(served_span * 1e3 / len(reqs), 1),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (seeded instanced-mesh scene, seeded VPLs, seeded Poisson request stream)",
            "config": {"workload": f"served guided path tracing, Poisson {rate}/s, images 128..1024 px square, "
                                   f"1-32 spp, depth 5, random cameras, {sc.ntri} tris, {c5['nvpl']} VPLs, 1000 "
                                   "resident clusters",
                       "triangles": int(sc.ntri), "vpls": int(c5["nvpl"]), "requests": len(reqs),
                       "paths": int(paths), "rate_per_s": rate, "memory_budget_gb": args.budget_gb,
                       "step_iterations": args.step_iterations, "parallelism": "single GPU"},
            "latency_ms": {"p50": pct(s_lat, 50), "p99": pct(s_lat, 99), "max": pct(s_lat, 100)},
            "plain_requests_per_s": round(len(reqs) / plain_span, 3),
            "plain_latency_ms": {"p50": pct(p_lat, 50), "p99": pct(p_lat, 99), "max": pct(p_lat, 100)},
            "speedup_vs_plain": round(plain_span / served_span, 2),
            "paths_per_s": round(paths / served_span, 1),
            "server": {k: sstats[k] for k in ("rounds", "items", "peak_bytes", "busy_seconds")},
            "accuracy": f"{len(reqs) - mismatches} of {len(reqs)} images bit-equal to the plain drop-in",
            "_paths": int(paths)}


def run_served(args):
    """the served workload at each rate of --rates (see measure_served), one JSON line per rate"""
    import renzoku_amd as rz
    ctx = rz.Ctx(0)
    c5 = setup_c5(args, ctx)
    for rate in [float(r) for r in args.rates.split(",")]:
        out = measure_served(args, ctx, c5, args.requests, rate, args.seed)
        out.pop("_paths")
        print(json.dumps(out), flush=True)


def device_copy(ctx, mib):
    """one MIB MiB device-to-device copy (a known-size transfer for scaling FETCH_SIZE / WRITE_SIZE)"""
    nb = int(mib) << 20
    src, dst = C.c_void_p(), C.c_void_p()
    lib = ctx.lib
    ctx.check(lib.rzk_device_alloc(ctx.h, nb, C.byref(src)), "rzk_device_alloc")
    ctx.check(lib.rzk_device_alloc(ctx.h, nb, C.byref(dst)), "rzk_device_alloc")
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpyDtoD.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    assert hip.hipMemcpyDtoD(dst, src, nb) == 0
    ctx.synchronize()
    lib.rzk_device_free(ctx.h, src)
    lib.rzk_device_free(ctx.h, dst)


# --------------------------------------------------------------------------- C2 (the headline)
C2_DESC = "lightslice-gather room 200k tris 512x512 65536 VPLs 32x32 slices 300 clusters"


def setup_c2(args):
    import renzoku_amd as rz
    from renzoku_amd import scenes, vplgen
    W = H = 512
    params = dict(width=W, height=H, tile=32, num_clusters=300, num_neighbor_slices=16, map_resolution=32,
                  map_epsilon=1e-3, clamp_mode=rz._lib.CLAMP_FORM_FACTOR, clamp=0.01, incident_clamp=0.01,
                  tmin=1e-3, seed=2013, map_gate=0, kd_max_slice_size=args.kd)
    t0 = time.perf_counter()
    sc = scenes.room(W, H, seed=2013, target_tris=args.tris or 200_000)
    ctx = rz.Ctx(0)
    scene = rz.Scene(ctx, sc)
    log(f"[bench] C2 scene: {sc.ntri} triangles, {scene.info().num_interior_nodes} interior nodes "
        f"({time.perf_counter() - t0:.1f} s)")
    vpl, vmat, n_emitted = vplgen.device_vpls(scene, args.vpls or 65_536, seed=2013)
    params["n_emitted"] = n_emitted
    vpls = rz.Vpls(ctx, vpl, vmat)
    log(f"[bench] C2 VPLs: {len(vmat)} from {n_emitted:.0f} light samples")
    ls = rz.LightSlice(ctx, scene, vpls, **params)
    return dict(sc=sc, ctx=ctx, scene=scene, vpl=vpl, vmat=vmat, vpls=vpls, ls=ls, params=params)


def cpu_baseline_c2(c2, lsinf, cores, model):
    """the whole LightSlice step on the oracle (a C restatement of the reference), all usable host cores:
    primary hits of every pixel, the reduced matrix of the slice rows, alphas + centre sampling +
    assignment + representatives, and the final gather over the representatives; plus the gather on one
    thread over a seeded sample of pixels"""
    import oracle
    from renzoku_amd import scenes
    sc, ls, vpl, vmat, prm = c2["sc"], c2["ls"], c2["vpl"], c2["vmat"], c2["params"]
    W, H = prm["width"], prm["height"]
    osc = oracle.OrcScene(sc)
    parts = {}
    t_all = time.perf_counter()
    t0 = time.perf_counter()
    px, py = scenes.pixel_grid(W, H)
    rays = oracle.camera_rays(sc.camera, px, py)
    t, b, g, tri = osc.trace_closest(rays)
    recv, m, _ = osc.fill_receivers(rays, t, b, g, tri)
    parts["primary_s"] = time.perf_counter() - t0
    row_pix = ls.get("row_pix", np.int32, lsinf.num_slices)
    rows_recv, rows_mat = recv[row_pix], m[row_pix]
    t0 = time.perf_counter()
    R, _ = osc.reduced_matrix(rows_recv, rows_mat, vpl, vmat, prm["clamp_mode"], prm["incident_clamp"], prm["tmin"],
                              nthreads=cores)
    parts["reduced_matrix_s"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    cl = oracle.cluster_by_sampling(R, prm["num_clusters"], prm["seed"], nthreads=cores)
    parts["clustering_s"] = time.perf_counter() - t0
    reps, rep_w = cl["reps"], cl["rep_w"]
    vr, vm = vpl[reps].copy(), vmat[reps].copy()
    t0 = time.perf_counter()
    osc.gather_clusters(recv, m, vr, vm, rep_w, prm["clamp_mode"], prm["clamp"], prm["tmin"], prm["n_emitted"],
                        nthreads=cores)
    parts["gather_s"] = time.perf_counter() - t0
    step_s = time.perf_counter() - t_all
    lit = int((m >= 0).sum())
    rows_ok = int((rows_mat >= 0).sum())
    segs = rows_ok * len(vmat) + lit * len(reps)
    # one thread: the gather over a seeded pixel sample, about 5 s
    rng = np.random.default_rng(1)
    ok = np.nonzero(m >= 0)[0]
    n1 = 256
    while True:
        sel = np.sort(rng.choice(ok, min(n1, len(ok)), replace=False))
        t0 = time.perf_counter()
        osc.gather_clusters(recv[sel], m[sel], vr, vm, rep_w, prm["clamp_mode"], prm["clamp"], prm["tmin"],
                            prm["n_emitted"], nthreads=1)
        dt1 = time.perf_counter() - t0
        if dt1 >= 2.5 or n1 >= len(ok):
            break
        n1 = int(min(len(ok), n1 * max(2.0, 5.0 / max(dt1, 1e-3))))
    return {"value": round(segs / step_s / 1e6, 3), "unit": "Mrays/s", "cores": cores, "kind": "port",
            "cpu": model,
            "sample": f"the whole step on the oracle: {W * H} primary hits, reduced matrix {rows_ok} rows x "
                      f"{len(vmat)} VPLs, clustering into {prm['num_clusters']}, gather {lit} pixels x {len(reps)} "
                      f"representatives ({segs} shadow segments) on {cores} threads, {step_s:.1f} s",
            "step_s": round(step_s, 3), "parts_s": {k: round(v, 3) for k, v in parts.items()},
            "one_thread": {"value": round(len(sel) * len(reps) / dt1 / 1e6, 3), "unit": "Mrays/s",
                           "sample": f"oracle gather of {len(sel)} seeded pixels x {len(reps)} representatives on "
                                     f"one thread, {dt1:.1f} s"}}


def measure_c2(args, c2):
    import renzoku_amd as rz  # noqa: F401
    ctx, ls, sc = c2["ctx"], c2["ls"], c2["sc"]
    ev = HipEvents()
    stream = ctx.stream
    e0, e1 = ev.create(), ev.create()
    for i in range(args.warmup):
        ls.initialize(sc.camera)
        ls.gather(to_host=False)
        ctx.synchronize()
        log(f"[bench] warmup {i + 1}/{args.warmup} done")
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
    # the step's work, counted after the timed steps (the LightSlice is initialized by then)
    work = ls.count_work()
    lsinf = ls.info()
    log(f"[bench] work per step: {work}")
    shadow_rays = work["reduced_pairs"] + work["gather_pairs"]
    value = shadow_rays / (ms_per_step / 1e3) / 1e6
    out = {
        "metric": "shadow Mrays/s + VPL-gather Mpairs/s; image req/s + p50/p99 latency vs plain",
        "value": round(value, 3), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded room scene, seeded light-traced VPLs)",
        "config": {"workload": (C2_DESC if not args.kd else
                                f"lightslice-gather room 200k tris 512x512 65536 VPLs kd-tree slices of {args.kd} "
                                "points 300 clusters"),
                   "triangles": int(sc.ntri), "width": 512, "height": 512, "vpls": int(len(c2["vmat"])),
                   "slices": int(lsinf.num_slices), "clusters": int(lsinf.num_clusters),
                   "parallelism": "single GPU"},
        "gather_mpairs_per_s": round(work["gather_pairs"] / (g_ms / 1e3) / 1e6, 3),
        "step_mpairs_per_s": round(shadow_rays / (ms_per_step / 1e3) / 1e6, 3),
        "gather_kernel_ms": round(g_ms, 4), "work": work}
    # the traversal byte model over each lane's own work (36 B segment I/O, 128 B per 4-wide node reached,
    # 48 B per triangle tested); node fetches are shared by the 64 lanes of a wave (scalar loads) and served
    # by L2 / MALL, so this is logical traffic, far above what reaches HBM
    gbytes = (work["gather_pairs"] * (RAY_IN_BYTES + SHADOW_OUT_BYTES) + work["gather_nodes"] * NODE4_BYTES +
              work["gather_tris"] * TRI_BYTES)
    out["_byte_model"] = {"bytes_per_launch": int(gbytes), "gbs": round(gbytes / (g_ms / 1e3) / 1e9, 1),
                          "nodes_per_pair": round(work["gather_nodes"] / max(1, work["gather_pairs"]), 3),
                          "tris_per_pair": round(work["gather_tris"] / max(1, work["gather_pairs"]), 3)}
    return out, lsinf, g_ms


# --------------------------------------------------------------------------- C1 (plain VPL render)
def setup_c1(ctx):
    import renzoku_amd as rz
    from renzoku_amd import scenes, vplgen
    sc = scenes.cornell_box(256, 256)
    scene = rz.Scene(ctx, sc)
    vpl, vmat, ne = vplgen.device_vpls(scene, 1024, seed=2013)
    vpls = rz.Vpls(ctx, vpl, vmat)
    px, py = scenes.pixel_grid(256, 256)
    _, _, recv, m, _ = scene.camera_primary(sc.camera, px, py)
    dev = []
    for a in (np.ascontiguousarray(recv, np.float32), np.ascontiguousarray(m, np.int32),
              np.zeros((len(m), 3), np.float32)):
        p = C.c_void_p()
        ctx.check(ctx.lib.rzk_device_alloc(ctx.h, a.nbytes, C.byref(p)), "rzk_device_alloc")
        ctx.check(ctx.lib.rzk_copy_h2d(ctx.h, p, a.ctypes.data_as(C.c_void_p), a.nbytes), "rzk_copy_h2d")
        dev.append(p)
    return dict(sc=sc, scene=scene, vpl=vpl, vmat=vmat, ne=ne, vpls=vpls, recv=recv, m=m, dev=dev)


def c1_gather_dev(ctx, c1):
    """the gather of every pixel's receiver over the 1,024 VPLs, device resident (no allocation)"""
    d_recv, d_mat, d_out = c1["dev"]
    ctx.check(ctx.lib.rzk_gather_dev_noalloc(ctx.h, c1["scene"].h, c1["vpls"].h, d_recv, d_mat, len(c1["m"]), 1,
                                             0.01, 1e-3, c1["ne"], d_out), "rzk_gather_dev_noalloc")


def measure_c1(args, ctx, pmc_hook, cores, model):
    """Plain VPL render of the seeded Cornell box: 256x256 gather points at pixel centres, 1,024 bounce-1
    VPLs, form-factor clamp 0.01.  A step is one rzk_render_vpl (primary hits + the fused gather, image
    left on the device); the gather kernels alone (k_visibility + k_gather_shade) are timed with HIP
    events around the device-resident no-allocation gather of the same receivers."""
    c1 = setup_c1(ctx)
    sc, cam = c1["sc"], np.ascontiguousarray(c1["sc"].camera, np.float32)
    render = lambda: ctx.check(ctx.lib.rzk_render_vpl(ctx.h, c1["scene"].h, c1["vpls"].h,  # noqa: E731
                                                      cam.ctypes.data_as(C.c_void_p), 1, 0.01, 1e-3, c1["ne"], None),
                               "rzk_render_vpl")
    for _ in range(2):
        render()
        c1_gather_dev(ctx, c1)
    ctx.synchronize()
    steps = max(3, args.steps)
    t0 = time.perf_counter()
    for _ in range(steps):
        render()
    ctx.synchronize()
    render_ms = (time.perf_counter() - t0) * 1e3 / steps
    ev = HipEvents()
    e0, e1 = ev.create(), ev.create()
    g = []
    for _ in range(steps):
        ev.record(e0, ctx.stream)
        c1_gather_dev(ctx, c1)
        ev.record(e1, ctx.stream)
        g.append(ev.elapsed_ms(e0, e1))
    g_ms = float(np.mean(g))
    lit = int((c1["m"] >= 0).sum())
    pairs = lit * len(c1["vmat"])
    leg = {"workload": "plain VPL render Cornell box 34 tris + light 256x256 1024 VPLs G clamp 0.01",
           "value": round(pairs / (render_ms / 1e3) / 1e6, 3), "unit": "Mpairs/s (= Mrays/s, one shadow segment "
           "per pair)", "ms_per_step": round(render_ms, 4), "gather_kernel_ms": round(g_ms, 4),
           "gather_mpairs_per_s": round(pairs / (g_ms / 1e3) / 1e6, 3), "pairs": pairs}
    pmc_hook("c1", leg, lambda pmc, peak: valu_roofline(pmc, ["visibility", "gather_shade"], g_ms,
                                                        "k_visibility + k_gather_shade (the C1 gather)", peak=peak))
    if not args.no_cpu:
        import oracle
        osc = oracle.OrcScene(sc)
        t0 = time.perf_counter()
        osc.gather(c1["recv"], c1["m"], c1["vpl"], c1["vmat"], 1, 0.01, 1e-3, c1["ne"], nthreads=cores)
        dt = time.perf_counter() - t0
        sel = np.nonzero(c1["m"] >= 0)[0][::16]
        t1 = time.perf_counter()
        osc.gather(c1["recv"][sel], c1["m"][sel], c1["vpl"], c1["vmat"], 1, 0.01, 1e-3, c1["ne"], nthreads=1)
        dt1 = time.perf_counter() - t1
        leg["cpu_baseline"] = {"value": round(pairs / dt / 1e6, 3), "unit": "Mpairs/s", "cores": cores,
                               "kind": "port", "cpu": model,
                               "sample": f"the whole image on the oracle: {lit} pixels x {len(c1['vmat'])} VPLs "
                                         f"on {cores} threads, {dt:.2f} s",
                               "one_thread": {"value": round(len(sel) * len(c1["vmat"]) / dt1 / 1e6, 3),
                                              "unit": "Mpairs/s",
                                              "sample": f"{len(sel)} pixels x {len(c1['vmat'])} VPLs, {dt1:.2f} s"}}
    for p in c1["dev"]:
        ctx.lib.rzk_device_free(ctx.h, p)
    return leg


# --------------------------------------------------------------------------- C3 (guided path tracing)
def setup_c3(args, ctx):
    import renzoku_amd as rz
    from renzoku_amd import scenes, vplgen
    W, H = 1280, 720
    t0 = time.perf_counter()
    sc = scenes.instanced_meshes(W, H, seed=2013, target_tris=args.tris3 or 1_000_000)
    scene = rz.Scene(ctx, sc)
    log(f"[bench] C3 scene: {sc.ntri} triangles ({time.perf_counter() - t0:.1f} s)")
    vpl, vmat, ne = vplgen.device_vpls(scene, 100_000, seed=2013)
    vpls = rz.Vpls(ctx, vpl, vmat)
    ls = rz.LightSlice(ctx, scene, vpls, width=W, height=H, tile=16, num_clusters=1000, num_neighbor_slices=16,
                       map_resolution=32, n_emitted=ne, seed=2013)
    t1 = time.perf_counter()
    ls.initialize(sc.camera)
    ctx.synchronize()
    init_ms = (time.perf_counter() - t1) * 1e3
    return dict(sc=sc, scene=scene, vpls=vpls, ls=ls, init_ms=init_ms)


C3_KW = dict(spp=16, max_bounce=5, direct_lighting=0, guide_mode=1, seed=2013, tmin=1e-3)


def measure_c3(args, ctx, c3, pmc_hook, cores, model):
    """Guided path tracing of the seeded 1M-triangle instanced scene (Ward BRDFs), 1280x720, 16 spp, depth 5,
    guided by 100k VPLs in 1,000 clusters per 16x16-pixel slice.  A step is one frame of the plain drop-in
    (rzk_lightslice_render_guided: one k_guided_pt launch), timed with HIP events."""
    sc, ls = c3["sc"], c3["ls"]
    W, H = 1280, 720
    ls.render_guided(sc.camera, count_rays=False, to_host=False, **C3_KW)
    _, rays = ls.render_guided(sc.camera, count_rays=True, to_host=False, **C3_KW)
    pw = ls.get("pt_work", np.int64, 4)
    ev = HipEvents()
    e0, e1 = ev.create(), ev.create()
    ms = []
    for i in range(max(2, min(args.steps, 4))):
        ev.record(e0, ctx.stream)
        ls.render_guided(sc.camera, count_rays=False, to_host=False, **C3_KW)
        ev.record(e1, ctx.stream)
        ms.append(ev.elapsed_ms(e0, e1))
        log(f"[bench] C3 frame {i + 1}: {ms[-1]:.1f} ms")
    k_ms = float(np.mean(ms))
    nrays = rays["closest_rays"] + rays["shadow_rays"]
    paths = W * H * C3_KW["spp"]
    algo = (rays["closest_rays"] * (RAY_IN_BYTES + 16) + rays["shadow_rays"] * (RAY_IN_BYTES + SHADOW_OUT_BYTES) +
            int(pw[2]) * NODE4_BYTES + int(pw[3]) * TRI_BYTES)
    leg = {"workload": "guided path tracing 1M tris 1280x720 16 spp depth 5 100k VPLs 16x16 slices 1000 clusters",
           "value": round(nrays / (k_ms / 1e3) / 1e6, 3), "unit": "Mrays/s (closest + shadow)",
           "ms_per_step": round(k_ms, 3), "paths_per_s": round(paths / (k_ms / 1e3), 1),
           "guide_build_ms": round(c3["init_ms"], 1), "rays": rays,
           "work": {"nodes4": int(pw[2]), "tris": int(pw[3]),
                    "nodes_per_ray": round(int(pw[2]) / max(1, nrays), 2),
                    "tris_per_ray": round(int(pw[3]) / max(1, nrays), 2)}}
    pmc_hook("c3", leg, lambda pmc, peak: hbm_roofline(
        pmc, ["guided_pt"], k_ms, algo, "k_guided_pt (one lane per pixel, all bounces)",
        "48 B per closest ray + 36 B per shadow ray of I/O, 128 B per 4-wide node and 48 B per triangle "
        "each lane's own traversal reaches (counted by the kernel)", peak))
    if not args.no_cpu:
        import oracle
        osc = oracle.OrcScene(sc)
        S = ls.info().num_slices
        res = 32
        guide = dict(res=res, maps=ls.get("maps", np.float32, S * res * res),
                     row_cdf=ls.get("row_cdf", np.float32, S * res),
                     col_cdf=ls.get("col_cdf", np.float32, S * res * res),
                     totals=ls.get("map_totals", np.float32, S))
        side = 16
        while True:
            win = (W // 2 - side // 2, H // 2 - side // 2, side, side)
            t0 = time.perf_counter()
            oracle.guided_pt(osc, sc.camera, C3_KW["spp"], C3_KW["max_bounce"], 0, 1, C3_KW["seed"], C3_KW["tmin"],
                             16, window=win, nthreads=cores, **guide)
            dt = time.perf_counter() - t0
            if dt >= 4.0 or side >= 256:
                break
            side = min(256, int(side * max(1.5, min(4.0, (8.0 / max(dt, 1e-3)) ** 0.5))))
        cpu_pps = side * side * C3_KW["spp"] / dt
        leg["cpu_baseline"] = {"value": round(cpu_pps / 1e6, 4), "unit": "Mpaths/s", "cores": cores,
                               "kind": "port", "cpu": model,
                               "sample": f"oracle guided path tracing of a centred {side}x{side} window x "
                                         f"{C3_KW['spp']} spp, depth 5, on {cores} threads, {dt:.1f} s",
                               "gpu_mpaths_per_s": round(paths / (k_ms / 1e3) / 1e6, 3)}
        c3["cpu_paths_per_s"] = cpu_pps
    return leg


# --------------------------------------------------------------------------- the default run
def pmc_child(args):
    """the kernels the PMC passes count, each launched once (C2 step, C1 gather, C3 frame), then the
    known-size copy"""
    legs = [x for x in args.legs.split(",") if x]
    c2 = setup_c2(args)
    ctx, ls = c2["ctx"], c2["ls"]
    ls.initialize(c2["sc"].camera)
    ls.gather(to_host=False)
    ctx.synchronize()
    if "c1" in legs:
        c1 = setup_c1(ctx)
        c1_gather_dev(ctx, c1)
        ctx.synchronize()
    if "c3" in legs:
        c3 = setup_c3(args, ctx)
        c3["ls"].render_guided(c3["sc"].camera, count_rays=False, to_host=False, **C3_KW)
        ctx.synchronize()
    device_copy(ctx, PMC_COPY_MIB)
    probe_ms()
    log("[bench] PMC child done")


# rough seconds each leg needs on the box (setup + timing + its CPU baseline), for the time budget
LEG_SECONDS = {"c1": 20, "c3": 60, "c4": 90, "c5": 90}


def run_default(args):
    """The headline C2 LightSlice step, then compact C1 / C3 / C4 / C5 legs, then the live PMC passes that
    give every roofline its measured counters; one JSON line."""
    import traceback
    t_begin = time.perf_counter()
    deadline = t_begin + args.time_budget
    model, cores = host_cpu()
    wanted = [x for x in args.legs.split(",") if x]
    c2 = setup_c2(args)
    ctx = c2["ctx"]
    out, lsinf, g_ms = measure_c2(args, c2)
    byte_model = out.pop("_byte_model")
    if not args.no_cpu:
        log("[bench] C2 CPU baseline (the whole step on the oracle) ...")
        out["cpu_baseline"] = cpu_baseline_c2(c2, lsinf, cores, model)
        log(f"[bench] C2 CPU baseline: {out['cpu_baseline']['value']} Mrays/s on {cores} threads")
    pending = []

    def pmc_hook(name, leg, fn):
        pending.append((leg, fn))

    legs, c3 = {}, None
    for name in wanted:
        left = deadline - time.perf_counter()
        if left < LEG_SECONDS.get(name, 60) + (0 if args.no_pmc else 150):
            legs[name] = {"skipped": f"time budget ({left:.0f} s left)"}
            continue
        log(f"[bench] leg {name} ...")
        try:
            if name == "c1":
                legs[name] = measure_c1(args, ctx, pmc_hook, cores, model)
            elif name == "c3":
                c3 = setup_c3(args, ctx)
                legs[name] = measure_c3(args, ctx, c3, pmc_hook, cores, model)
            elif name == "c4":
                if c3 is None:
                    c3 = setup_c3(args, ctx)
                guides = guides_for(ctx, c3["sc"], c3["scene"], BATCH_LEVELS)
                r = measure_batched(args, c3["sc"], ctx, c3["scene"], guides, args.jobs, 12, 2)
                leg = {k: r[k] for k in ("value", "unit", "ms_per_step", "plain_requests_per_s", "speedup_vs_plain",
                                         "mrays_per_s", "autotuned", "accuracy")}
                leg["workload"] = r["config"]["workload"]
                if "cpu_paths_per_s" in c3:
                    leg["cpu_baseline"] = {
                        "value": round(args.jobs / (r["_paths"] / c3["cpu_paths_per_s"]), 4), "unit": "requests/s",
                        "cores": cores, "kind": "port", "cpu": model,
                        "sample": f"the job mix's {r['_paths']} paths at the oracle's guided path tracing rate "
                                  "measured in the C3 leg on the same scene"}
                legs[name] = leg
            elif name == "c5":
                c5 = setup_c5(args, ctx)
                r = measure_served(args, ctx, c5, args.requests, args.rate, args.seed)
                leg = {k: r[k] for k in ("value", "unit", "latency_ms", "plain_requests_per_s", "plain_latency_ms",
                                         "speedup_vs_plain", "paths_per_s", "accuracy")}
                leg["workload"] = r["config"]["workload"]
                leg["scene_build"] = c5["build"]
                if c3 is not None and "cpu_paths_per_s" in c3:
                    leg["cpu_baseline"] = {
                        "value": round(args.requests / (r["_paths"] / c3["cpu_paths_per_s"]), 4),
                        "unit": "requests/s", "cores": cores, "kind": "port", "cpu": model,
                        "sample": f"the request stream's {r['_paths']} paths at the oracle's guided path tracing "
                                  "rate measured in the C3 leg (the 1M-triangle half of this scene)"}
                legs[name] = leg
            log(f"[bench] leg {name}: {legs[name].get('value')} {legs[name].get('unit')}")
        except Exception as e:  # a failed leg is reported, the headline line still prints
            traceback.print_exc()
            legs[name] = {"error": f"{type(e).__name__}: {e}"}
    p_ms = probe_ms()
    pmc, note = {}, "PMC passes not run (--no-pmc)"
    if not args.no_pmc:
        child = ["--legs", ",".join(x for x in ("c1", "c3") if x in legs and "error" not in legs[x]
                                    and "skipped" not in legs[x])]
        for flag, v in (("--tris", args.tris), ("--vpls", args.vpls), ("--kd", args.kd), ("--tris3", args.tris3)):
            if v:
                child += [flag, str(v)]
        pmc, note = pmc_collect(child, deadline)
    peak = valu_peak(pmc, p_ms)
    out["roofline"] = valu_roofline(pmc, ["slice_visibility", "slice_shade"], g_ms,
                                    "k_slice_visibility + k_slice_shade (the gather)", byte_model, peak)
    if note:
        out["roofline"]["pmc_note"] = note
    rm = pmc.get("reduced_matrix", {})
    if "_ns" in rm:
        out["roofline_reduced_matrix"] = valu_roofline(pmc, ["reduced_matrix"], rm["_ns"] / 1e6,
                                                       "k_reduced_matrix (launch time from the counter pass)",
                                                       peak=peak)
    for leg, fn in pending:
        leg["roofline"] = fn(pmc, peak)
    out["legs"] = legs
    out["bench_seconds"] = round(time.perf_counter() - t_begin, 1)
    print(json.dumps(out), flush=True)

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="lightslice", choices=["lightslice", "guided", "batched", "served", "refine", "vplgen",
                                                                 "kdslice", "build"],
                    help="lightslice: the 200k-triangle LightSlice gather step (the headline line); "
                         "guided: guided path tracing of the 1M-triangle scene; batched: 256 uneven jobs")
    ap.add_argument("--jobs", type=int, default=256, help="batched workload: number of jobs")
    ap.add_argument("--requests", type=int, default=32, help="served workload / C5 leg: number of requests")
    ap.add_argument("--rate", type=float, default=200.0, help="C5 leg: Poisson arrival rate (1/s)")
    ap.add_argument("--budget-gb", type=int, default=16, help="served workload: memory budget (GiB)")
    ap.add_argument("--step-iterations", type=int, default=8,
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
    ap.add_argument("--tris3", type=int, default=0, help="C3 / C4 scene triangles (default 1M)")
    ap.add_argument("--tris5", type=int, default=0, help="C5 scene triangles (default 2M)")
    ap.add_argument("--vpls5", type=int, default=0, help="C5 VPL count (default 200k)")
    ap.add_argument("--rates", default="50", help="served workload: comma-separated Poisson rates (1/s)")
    ap.add_argument("--legs", default="c1,c3,c4,c5", help="default run: the legs beside the C2 headline")
    ap.add_argument("--no-pmc", action="store_true", help="default run: skip the live rocprofv3 counter passes")
    ap.add_argument("--time-budget", type=float, default=540.0,
                    help="default run: seconds after which legs and counter passes are skipped")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.pmc_child:
        return pmc_child(args)
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
    if args.workload == "build":
        return run_build(args)

    return run_default(args)




if __name__ == "__main__":
    main()
