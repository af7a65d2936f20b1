// batch.hip — the batched form of guided path tracing: a job table of render
// windows (each with its own guide, spp, depth and seed), one persistent
// kernel over a work queue of (job, 64-pixel block, sample block) items, an
// in-order per-pixel reduction, the three launches captured once as a HIP
// graph, and a coordinate-descent autotuner over the launch shape.
//
// The host side replaces the reference's per-bundle dispatch
// (IntegratorScheduler, integrator_scheduler.cc:185-229, one 64x64 bundle per
// worker) with one launch for any number of windows of any size.

#include <algorithm>
#include <cstring>
#include <memory>

#include "pathtrace.h"

using namespace rzk;

namespace {

template <typename T>
struct DevArray {
    T* p = nullptr;
    size_t n = 0;
    void reserve(size_t count) {
        if (count <= n && p) return;
        if (p) RZK_HIP(hipFree(p));
        p = nullptr;
        n = 0;
        RZK_HIP(hipMalloc((void**)&p, sizeof(T) * std::max<size_t>(count, 1)));
        n = std::max<size_t>(count, 1);
    }
    ~DevArray() { if (p) (void)hipFree(p); }
};

template <typename F>
int batch_guard(rzk_ctx* ctx, F&& f) {
    try {
        f();
        if (ctx) ctx->last_error.clear();
        return RZK_OK;
    } catch (const RzkError& e) {
        if (ctx) ctx->last_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        if (ctx) ctx->last_error = "host allocation failed";
        return RZK_E_OOM;
    } catch (const std::exception& e) {
        if (ctx) ctx->last_error = e.what();
        return RZK_E_INVALID;
    }
}

bool valid_config(const rzk_batch_config& c) {
    return (c.block == 64 || c.block == 128 || c.block == 192 || c.block == 256) && c.blocks_per_cu >= 1 &&
           c.blocks_per_cu <= 64 && c.items_per_dequeue >= 1 && c.items_per_dequeue <= 64 &&
           c.samples_per_item >= 1 && c.samples_per_item <= 1024 && (c.engine == 0 || c.engine == 1) &&
           c.path_slots >= 4096 && c.path_slots <= (1 << 26);
}

// whether two configurations capture the same graph: the same launch shape
// and the same persistent kernel (one item per dequeue has its own)
bool same_launch(const rzk_batch_config& a, const rzk_batch_config& b) {
    return a.block == b.block && a.blocks_per_cu == b.blocks_per_cu && a.engine == b.engine &&
           a.path_slots == b.path_slots && (a.items_per_dequeue == 1) == (b.items_per_dequeue == 1);
}

}  // namespace

struct rzk_batch {
    rzk_ctx* ctx = nullptr;
    const rzk_scene* scene = nullptr;
    rzk_batch_config cfg{256, 8, 1, 1, 1, 1 << 20};
    WfEngine* wf = nullptr;
    DevArray<JobGpu> jobs;
    DevArray<float> samples, out;
    DevArray<BatchHeader> hdr;
    DevArray<unsigned long long> counters;   // {queue head, closest rays, shadow rays}
    DevArray<int> spill;
    PinnedArray<JobGpu> h_jobs;
    PinnedArray<BatchHeader> h_hdr;
    PinnedArray<unsigned long long> h_counters;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    SceneView sv{};

    int grid() const { return std::max(1, ctx->num_cus) * cfg.blocks_per_cu; }

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        if (wf) wf_destroy(wf);
        wf = nullptr;
    }

    // capture reset + persistent kernel + reduction for the current launch shape
    void capture() {
        drop_graph();
        sv = scene_view(scene);
        int64_t threads = (int64_t)grid() * cfg.block;
        int64_t extra = 2 * (int64_t)sv.stack4 - kPtLdsStack;   // closest_hit4 pairs
        if (extra > 0) {
            spill.reserve((size_t)(extra * threads));
            sv.spill = spill.p;
            sv.spill_pitch = threads;
        } else {
            sv.spill = nullptr;
            sv.spill_pitch = 0;
        }
        hipStream_t st = ctx->stream;
        RZK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        launch_batch(st, sv, hdr.p, grid(), cfg.block, cfg.items_per_dequeue, std::max(1, ctx->num_cus) * 8);
        hipError_t le = hipGetLastError();
        hipGraph_t g = nullptr;
        hipError_t ee = hipStreamEndCapture(st, &g);
        RZK_HIP(le);
        RZK_HIP(ee);
        graph = g;
        RZK_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }

    ~rzk_batch() {
        drop_graph();
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
    }
};

PtParams job_params(const rzk_guide& g, const rzk_render_job& j) {
    PtParams pp;
    pp.width = g.W; pp.height = g.H; pp.spp = j.spp; pp.max_bounce = j.max_bounce;
    pp.direct = j.direct_lighting ? 1 : 0; pp.guide_mode = j.guide_mode; pp.tile = g.tile;
    pp.res = g.res; pp.tiles_x = g.tiles_x; pp.tiles_y = g.tiles_y;
    pp.x0 = j.x0; pp.y0 = j.y0; pp.w = j.w; pp.h = j.h;
    pp.seed = j.seed; pp.tmin = j.tmin;
    return pp;
}

void check_job(const rzk_guide* g, const rzk_render_job& j) {
    RZK_CHECK(g != nullptr, RZK_E_INVALID, "job without a guide");
    RZK_CHECK(j.spp > 0 && j.max_bounce >= 0 && (j.guide_mode == 0 || j.guide_mode == 1), RZK_E_INVALID,
              "bad path tracing parameters in a job");
    RZK_CHECK(j.w > 0 && j.h > 0 && j.x0 >= 0 && j.y0 >= 0 && j.x0 + j.w <= g->W && j.y0 + j.h <= g->H,
              RZK_E_INVALID, "job window outside its guide's image");
    RZK_CHECK(g->tile > 0 && g->tiles_x > 0 && g->tiles_y > 0, RZK_E_INVALID, "guide without slices");
}

static void batch_render(rzk_batch* b, const rzk_render_job* jobs, int32_t n, float* out_rgb, int64_t* out_rays,
                         float* out_ms) {
    rzk_ctx* ctx = b->ctx;
    RZK_CHECK(jobs && n > 0, RZK_E_INVALID, "empty batch");
    for (int32_t i = 0; i < n; ++i) {
        check_job(jobs[i].guide, jobs[i]);
        RZK_CHECK(jobs[i].guide->scene == b->scene, RZK_E_INVALID, "job guide belongs to another scene");
    }
    // previous launch must be done before the pinned staging is rewritten
    RZK_HIP(hipStreamSynchronize(ctx->stream));
    b->h_jobs.reserve((size_t)n);
    int64_t items = 0, pixels = 0, samples = 0;
    const bool wave = b->cfg.engine == 1;
    const int S = wave ? 1 : b->cfg.samples_per_item;
    int max_bounce = 0;
    for (int32_t i = 0; i < n; ++i) {
        const rzk_render_job& j = jobs[i];
        JobGpu& J = b->h_jobs.p[i];
        std::memset(&J, 0, sizeof(J));
        const rzk_guide& g = *j.guide;
        std::memcpy(J.cam.v, g.cam, sizeof(J.cam.v));
        J.pp = job_params(g, j);
        J.gv = g.gv;
        int64_t npx = (int64_t)j.w * j.h;
        J.pix_blocks = (int)((npx + 63) / 64);
        J.samples_per_item = std::min(S, j.spp);
        J.samp_blocks = (j.spp + J.samples_per_item - 1) / J.samples_per_item;
        int strata = 1;
        while ((strata + 1) * (strata + 1) <= j.spp) ++strata;
        J.strata = strata * strata == j.spp ? strata : 0;
        J.item_first = items;
        J.pix_first = pixels;
        J.sample_off = samples;
        max_bounce = std::max(max_bounce, j.max_bounce);
        // persistent engine: work items; wavefront engine: paths (sample-major)
        items += wave ? npx * j.spp : (int64_t)J.pix_blocks * J.samp_blocks;
        pixels += npx;
        samples += npx * j.spp * 3;
    }
    b->jobs.reserve((size_t)n);
    b->samples.reserve((size_t)samples);
    b->out.reserve((size_t)pixels * 3);
    if (wave) {
        if (!b->wf) b->wf = wf_create(ctx, b->scene, b->cfg.path_slots, b->cfg.block, b->cfg.blocks_per_cu);
    } else if (!b->exec) {
        b->capture();
    }
    BatchHeader& h = *b->h_hdr.p;
    h.jobs = b->jobs.p;
    h.n_jobs = n;
    h.n_items = items;
    h.n_pixels = pixels;
    h.samples = b->samples.p;
    h.out = b->out.p;
    h.counter = b->counters.p;
    h.work = out_rays ? b->counters.p + 1 : nullptr;
    h.items_per_dequeue = b->cfg.items_per_dequeue;
    h.pad = 0;
    h.job_lo = 0;
    h.job_mask = -1;
    h.item_lo = 0;
    hipStream_t st = ctx->stream;
    RZK_HIP(hipMemcpyAsync(b->jobs.p, b->h_jobs.p, sizeof(JobGpu) * n, hipMemcpyHostToDevice, st));
    RZK_HIP(hipMemcpyAsync(b->hdr.p, b->h_hdr.p, sizeof(BatchHeader), hipMemcpyHostToDevice, st));
    RZK_HIP(hipEventRecord(b->ev0, st));
    if (wave) {
        if (out_rays) RZK_HIP(hipMemsetAsync(b->counters.p + 1, 0, 2 * sizeof(unsigned long long), st));
        wf_run(b->wf, b->jobs.p, n, items, max_bounce, b->samples.p, out_rays ? b->counters.p + 1 : nullptr, nullptr);
        launch_batch_reduce(st, b->hdr.p, std::max(1, ctx->num_cus) * 8);
        RZK_HIP(hipGetLastError());
    } else {
        RZK_HIP(hipGraphLaunch(b->exec, st));
    }
    RZK_HIP(hipEventRecord(b->ev1, st));
    if (out_rgb)
        RZK_HIP(hipMemcpyAsync(out_rgb, b->out.p, sizeof(float) * 3 * pixels, hipMemcpyDeviceToHost, st));
    if (out_rays)
        RZK_HIP(hipMemcpyAsync(b->h_counters.p, b->counters.p, sizeof(unsigned long long) * 3,
                               hipMemcpyDeviceToHost, st));
    RZK_HIP(hipStreamSynchronize(st));
    if (out_rays) {
        out_rays[0] = (int64_t)b->h_counters.p[1];
        out_rays[1] = (int64_t)b->h_counters.p[2];
    }
    if (out_ms) RZK_HIP(hipEventElapsedTime(out_ms, b->ev0, b->ev1));
}

extern "C" {

int rzk_batch_create(rzk_ctx* ctx, const rzk_scene* scene, const rzk_batch_config* cfg, rzk_batch** out) {
    if (!ctx || !scene || !out) return RZK_E_INVALID;
    *out = nullptr;
    return batch_guard(ctx, [&] {
        RZK_HIP(hipSetDevice(ctx->device));
        std::unique_ptr<rzk_batch> b(new rzk_batch());
        b->ctx = ctx;
        b->scene = scene;
        if (cfg) {
            RZK_CHECK(valid_config(*cfg), RZK_E_INVALID, "bad batch configuration");
            b->cfg = *cfg;
        }
        b->hdr.reserve(1);
        b->counters.reserve(3);
        b->h_hdr.reserve(1);
        b->h_counters.reserve(3);
        RZK_HIP(hipEventCreate(&b->ev0));
        RZK_HIP(hipEventCreate(&b->ev1));
        ctx_retain(ctx);
        *out = b.release();
    });
}

int rzk_batch_destroy(rzk_batch* b) {
    if (!b) return RZK_OK;
    (void)hipStreamSynchronize(b->ctx->stream);
    rzk_ctx* ctx = b->ctx;
    delete b;
    ctx_release(ctx);
    return RZK_OK;
}

int rzk_batch_set_config(rzk_batch* b, const rzk_batch_config* cfg) {
    if (!b || !cfg) return RZK_E_INVALID;
    return batch_guard(b->ctx, [&] {
        RZK_CHECK(valid_config(*cfg), RZK_E_INVALID, "bad batch configuration");
        RZK_HIP(hipStreamSynchronize(b->ctx->stream));
        bool shape = !same_launch(*cfg, b->cfg);
        b->cfg = *cfg;
        if (shape) b->drop_graph();
    });
}

int rzk_batch_get_config(const rzk_batch* b, rzk_batch_config* out) {
    if (!b || !out) return RZK_E_INVALID;
    *out = b->cfg;
    return RZK_OK;
}

int rzk_batch_render(rzk_batch* b, const rzk_render_job* jobs, int32_t n, float* out_rgb, int64_t* out_rays,
                     float* out_ms) {
    if (!b) return RZK_E_INVALID;
    return batch_guard(b->ctx, [&] { batch_render(b, jobs, n, out_rgb, out_rays, out_ms); });
}

// timed renders per autotuner candidate (after one warm-up render)
static constexpr int kTuneRuns = 3;

int rzk_batch_autotune(rzk_batch* b, const rzk_render_job* jobs, int32_t n, rzk_batch_config* best,
                       rzk_batch_trial* tried, int32_t max_tried, int32_t* n_tried) {
    if (!b) return RZK_E_INVALID;
    return batch_guard(b->ctx, [&] {
        RZK_CHECK(jobs && n > 0, RZK_E_INVALID, "empty batch");
        int32_t count = 0;
        auto time_cfg = [&](const rzk_batch_config& c) {
            bool shape = !same_launch(c, b->cfg);
            b->cfg = c;
            if (shape) b->drop_graph();
            batch_render(b, jobs, n, nullptr, nullptr, nullptr);   // instantiate + warm
            float ms = 0.f;
            for (int r = 0; r < kTuneRuns; ++r) {                  // the fastest of a few runs
                float t = 0.f;
                batch_render(b, jobs, n, nullptr, nullptr, &t);
                ms = r == 0 ? t : std::min(ms, t);
            }
            if (tried && count < max_tried) tried[count] = rzk_batch_trial{c, ms};
            count++;
            return ms;
        };
        rzk_batch_config cur = b->cfg;
        float cur_ms = time_cfg(cur);
        // one field at a time; the value a stage starts from is already timed
        // (every candidate differs from all earlier ones)
        auto stage = [&](int rzk_batch_config::*field, std::initializer_list<int> values) {
            const int start = cur.*field;
            for (int v : values) {
                if (v == start) continue;
                rzk_batch_config c = cur;
                c.*field = v;
                float ms = time_cfg(c);
                if (ms < cur_ms) { cur_ms = ms; cur = c; }
            }
        };
        stage(&rzk_batch_config::engine, {0, 1});
        if (cur.engine == 1) {
            stage(&rzk_batch_config::block, {128, 256});
            stage(&rzk_batch_config::blocks_per_cu, {4, 8, 16});
            stage(&rzk_batch_config::path_slots, {1 << 18, 1 << 19, 1 << 20, 1 << 21});
        } else {
            stage(&rzk_batch_config::samples_per_item, {1, 2, 4, 8, 16});
            stage(&rzk_batch_config::block, {64, 128, 256});
            stage(&rzk_batch_config::blocks_per_cu, {2, 4, 8, 16});
            stage(&rzk_batch_config::items_per_dequeue, {1, 2, 4, 8});
        }
        bool shape = !same_launch(cur, b->cfg);
        b->cfg = cur;
        if (shape) b->drop_graph();
        if (best) *best = cur;
        if (n_tried) *n_tried = count;
    });
}

}  // extern "C"
