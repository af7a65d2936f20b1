// serve.hip — the served form: a request loop over a resident scene, VPL
// clusters and memory budget, feeding one wavefront path-tracing engine.
//
// Requests (camera, spp, depth, seed) are queued by any thread.  A worker
// thread admits them oldest first while the budget allows (guide tables +
// per-sample radiance + image), builds each admitted request's guide from the
// resident clusters, and appends the request as a job to the engine's job
// ring.  Two engines (launch.engine):
//  * 0, the persistent path kernel: a job's work items (64-pixel blocks x
//    samples) are numbered after the previous job's; each engine step runs
//    the next range of items, up to step_iterations items per wave of the
//    grid, across whichever jobs they belong to, so requests that arrive
//    while others are on the GPU start at the next step.  A job is finished
//    when the step holding its last item has finished.
//  * 1, the wavefront engine: free path slots take the next paths of
//    whichever jobs are open (continuous batching at path granularity); a
//    job is finished when its finished-path count, read after each step,
//    reaches its paths.
// A finished job's pixels are summed in sample order, the image is copied out
// asynchronously, and its memory returns to the budget.  A request's image is
// therefore the same, bit for bit, alone or in any mix, and equal to the plain
// per-request render.
//
// Replaces the reference's per-frame loop (Scene::start -> IntegratorScheduler,
// scene.cc:441, integrator_scheduler.cc:143-229), which renders one image at
// a time over a thread pool.

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "pathtrace.h"

using namespace rzk;

namespace {

constexpr int64_t kJobRing = 1 << 14;      // jobs open at once (a power of two)
constexpr int kMaxStepIterations = 32;    // work items per wave (or wavefront iterations) per step

// first-fit allocator over one device block: the memory budget
struct Arena {
    char* base = nullptr;
    size_t size = 0, used = 0, peak = 0;
    std::map<size_t, size_t> free_;   // offset -> bytes, coalesced

    void init(size_t bytes) {
        size = (bytes + 255) & ~(size_t)255;
        RZK_HIP(hipMalloc((void**)&base, size));
        free_.clear();
        free_[0] = size;
    }
    bool alloc(size_t bytes, size_t& off) {
        bytes = (bytes + 255) & ~(size_t)255;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            off = it->first;
            size_t rest = it->second - bytes;
            free_.erase(it);
            if (rest) free_[off + bytes] = rest;
            used += bytes;
            peak = std::max(peak, used);
            return true;
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = (bytes + 255) & ~(size_t)255;
        used -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    ~Arena() { if (base) (void)hipFree(base); }
};

struct Request {
    int64_t ticket = 0;
    rzk_request r{};
    int W = 0, H = 0;
    double t_submit = 0, t_admit = 0, t_finish = 0;
    int steps = 0;                   // engine steps the request was open for
    GuideLayout L;
    size_t bytes_samples = 0, bytes_img = 0;
    size_t off_tab = 0, off_smp = 0, off_out = 0;
    rzk_guide guide;
    int64_t job = -1, paths = 0;
    int64_t item_end = 0;            // persistent engine: one past the job's last work item
    hipEvent_t done_ev = nullptr;
    float* host_img = nullptr;       // pinned, filled by the final copy
    size_t host_bytes = 0;
    int state = 0;                   // 0 queued, 1 open, 2 copying, 3 finished, 4 failed
    int64_t held_bytes() const { return (int64_t)(L.bytes_tab + bytes_samples + bytes_img); }
};

size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

// device bytes a request takes at admission (the guide-build temporaries are
// returned right after the build is queued)
size_t request_bytes(const GuideLayout& L, size_t samples, size_t img) {
    return round256(L.bytes_tab) + round256(L.bytes_tmp) + round256(samples) + round256(img);
}

}  // namespace

struct rzk_server {
    rzk_ctx* ctx = nullptr;                // owned: the server's stream
    const rzk_scene* scene = nullptr;
    const rzk_clusters* cl = nullptr;
    rzk_server_params prm{};
    rzk_guide_params gp{};
    Arena arena;
    WfEngine* wf = nullptr;
    // persistent engine: launch shape, traversal spill, step headers, queue head, item counts
    bool persistent = false;
    int grid = 0, block = 0;
    int64_t step_items = 0;
    SceneView sv{};
    int* d_spill = nullptr;
    BatchHeader* d_hdr = nullptr;          // one per step slot
    BatchHeader* h_hdr = nullptr;          // pinned staging, one per step slot
    unsigned long long* d_head = nullptr;
    int64_t total_items = 0, next_item = 0;
    int64_t step_end[2] = {0, 0};
    JobGpu* d_jobs = nullptr;              // job ring
    JobGpu* h_jobs = nullptr;              // pinned staging, one entry per ring slot
    unsigned int* d_done = nullptr;        // finished paths per ring slot
    unsigned int* h_done = nullptr;        // two pinned snapshots
    hipEvent_t ev_step[2] = {nullptr, nullptr}, ev_t0[2] = {nullptr, nullptr};
    bool step_busy[2] = {false, false};
    int64_t step_jobs[2] = {0, 0};         // jobs admitted before each snapshot was queued
    int step_slot = 0;
    int64_t n_jobs = 0, total_paths = 0;
    // request bookkeeping
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> incoming;
    std::unordered_map<int64_t, std::unique_ptr<Request>> reqs;
    std::deque<Request*> waiting, open, copying;
    std::vector<std::pair<float*, size_t>> pinned_pool;
    int64_t next_ticket = 1;
    std::atomic<bool> stop{false};
    std::string error;
    std::thread worker;
    std::chrono::steady_clock::time_point t0;
    rzk_server_stats stats{};

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    // a pooled pinned buffer of at least `bytes` (its size in `bytes`), or null
    float* pinned_get(size_t& bytes) {
        std::lock_guard<std::mutex> lk(mu);
        size_t best = (size_t)-1;
        int bi = -1;
        for (size_t i = 0; i < pinned_pool.size(); ++i)
            if (pinned_pool[i].second >= bytes && pinned_pool[i].second < best) { best = pinned_pool[i].second; bi = (int)i; }
        if (bi >= 0) {
            float* p = pinned_pool[bi].first;
            bytes = pinned_pool[bi].second;
            pinned_pool.erase(pinned_pool.begin() + bi);
            return p;
        }
        return nullptr;
    }

    int64_t oldest_open_job() const { return open.empty() ? n_jobs : open.front()->job; }

    void publish() {
        wf_publish(wf, WfJobs{d_jobs, oldest_open_job(), n_jobs, kJobRing, total_paths, (float*)arena.base, nullptr,
                              d_done});
    }

    // admit queued requests oldest first while the budget and the job ring allow
    void admit() {
        while (!waiting.empty() && n_jobs - oldest_open_job() < kJobRing) {
            Request* q = waiting.front();
            size_t o_tab, o_tmp, o_smp, o_out;
            if (!arena.alloc(q->L.bytes_tab, o_tab)) break;
            if (!arena.alloc(q->L.bytes_tmp, o_tmp)) { arena.release(o_tab, q->L.bytes_tab); break; }
            if (!arena.alloc(q->bytes_samples, o_smp)) {
                arena.release(o_tab, q->L.bytes_tab);
                arena.release(o_tmp, q->L.bytes_tmp);
                break;
            }
            if (!arena.alloc(q->bytes_img, o_out)) {
                arena.release(o_tab, q->L.bytes_tab);
                arena.release(o_tmp, q->L.bytes_tmp);
                arena.release(o_smp, q->bytes_samples);
                break;
            }
            q->off_tab = o_tab; q->off_smp = o_smp; q->off_out = o_out;
            hipStream_t st = ctx->stream;
            GuideBuffers b = guide_buffers(q->L, arena.base + o_tmp, arena.base + o_tab);
            enqueue_guide_build(ctx, scene, *cl, q->r.camera, gp, q->L, b);
            // temporaries are only read by the kernels just queued: later users are stream-ordered after them
            arena.release(o_tmp, q->L.bytes_tmp);
            guide_set(q->guide, scene, q->r.camera, gp, q->L, b);
            q->guide.ctx = ctx;
            // the job: paths numbered from total_paths, samples in its arena region
            int64_t j = n_jobs++;
            JobGpu& J = h_jobs[j & (kJobRing - 1)];
            std::memset(&J, 0, sizeof(J));
            std::memcpy(J.cam.v, q->r.camera, sizeof(J.cam.v));
            rzk_render_job rj{&q->guide, 0, 0, q->W, q->H, q->r.spp, q->r.max_bounce, q->r.direct_lighting,
                              q->r.guide_mode, q->r.seed, q->r.tmin};
            J.pp = job_params(q->guide, rj);
            J.gv = q->guide.gv;
            J.sample_off = (int64_t)(o_smp / sizeof(float));
            int strata = 1;
            while ((strata + 1) * (strata + 1) <= q->r.spp) ++strata;
            J.strata = strata * strata == q->r.spp ? strata : 0;
            const int64_t npx = (int64_t)q->W * q->H;
            J.pix_blocks = (int)((npx + 63) / 64);
            J.samples_per_item = 1;
            J.samp_blocks = q->r.spp;
            if (persistent) {
                J.item_first = total_items;
                total_items += (int64_t)J.pix_blocks * J.samp_blocks;
                q->item_end = total_items;
            } else {
                J.item_first = total_paths;
            }
            RZK_HIP(hipMemcpyAsync(d_jobs + (j & (kJobRing - 1)), &J, sizeof(JobGpu), hipMemcpyHostToDevice, st));
            RZK_HIP(hipMemsetAsync(d_done + (j & (kJobRing - 1)), 0, sizeof(unsigned int), st));
            q->job = j;
            q->paths = (int64_t)q->W * q->H * q->r.spp;
            total_paths += q->paths;
            q->t_admit = now();
            q->state = 1;
            waiting.pop_front();
            open.push_back(q);
        }
        std::lock_guard<std::mutex> lk(mu);
        stats.peak_bytes = std::max<int64_t>(stats.peak_bytes, (int64_t)arena.peak);
    }

    void retire_step(int k) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ev_t0[k], ev_step[k]) == hipSuccess) {
            std::lock_guard<std::mutex> lk(mu);
            stats.busy_seconds += ms * 1e-3;
        }
        step_busy[k] = false;
    }

    // requests whose paths are all finished (as of snapshot `k`): reduce, copy out
    void finish_done(int k) {
        const unsigned int* done = h_done + k * kJobRing;
        hipStream_t st = ctx->stream;
        int64_t finished_paths = 0;
        for (auto it = open.begin(); it != open.end();) {
            Request* q = *it;
            // a job admitted after the snapshot was queued has no count in it yet; with the
            // persistent engine, every item before step k's end is done
            const bool finished = persistent ? q->item_end <= step_end[k]
                                             : q->job < step_jobs[k] &&
                                                   (int64_t)done[q->job & (kJobRing - 1)] >= q->paths;
            if (!finished) {
                q->steps++;
                ++it;
                continue;
            }
            q->steps++;
            launch_job_reduce(st, d_jobs + (q->job & (kJobRing - 1)), (const float*)arena.base,
                              (float*)(arena.base + q->off_out));
            RZK_HIP(hipGetLastError());
            q->host_bytes = q->bytes_img;
            q->host_img = pinned_get(q->host_bytes);
            if (!q->host_img) RZK_HIP(hipHostMalloc((void**)&q->host_img, q->host_bytes, hipHostMallocDefault));
            RZK_HIP(hipMemcpyAsync(q->host_img, arena.base + q->off_out, q->bytes_img, hipMemcpyDeviceToHost, st));
            RZK_HIP(hipEventRecord(q->done_ev, st));
            q->state = 2;
            finished_paths += q->paths;
            copying.push_back(q);
            it = open.erase(it);
        }
        std::lock_guard<std::mutex> lk(mu);
        stats.items += finished_paths;
    }

    void retire_copies() {
        while (!copying.empty()) {
            Request* q = copying.front();
            hipError_t e = hipEventQuery(q->done_ev);
            if (e == hipErrorNotReady) break;
            RZK_HIP(e);
            copying.pop_front();
            arena.release(q->off_tab, q->L.bytes_tab);
            arena.release(q->off_smp, q->bytes_samples);
            arena.release(q->off_out, q->bytes_img);
            std::lock_guard<std::mutex> lk(mu);
            q->t_finish = now();
            q->state = 3;
            stats.finished++;
            cv_done.notify_all();
        }
    }

    void loop() {
        const int iters = std::max(1, std::min(kMaxStepIterations, prm.step_iterations));
        while (true) {
            {
                std::unique_lock<std::mutex> lk(mu);
                if (stop) break;
                while (!incoming.empty()) { waiting.push_back(incoming.front()); incoming.pop_front(); }
                bool idle = waiting.empty() && open.empty() && copying.empty() && !step_busy[0] && !step_busy[1];
                if (idle) {
                    cv_work.wait_for(lk, std::chrono::milliseconds(50));
                    continue;
                }
            }
            retire_copies();
            admit();
            if (open.empty()) {
                // nothing on the engine: drain outstanding steps, then copies
                bool waited = false;
                for (int k = 0; k < 2; ++k)
                    if (step_busy[k]) { RZK_HIP(hipEventSynchronize(ev_step[k])); retire_step(k); waited = true; }
                if (!waited) {
                    if (!copying.empty()) RZK_HIP(hipEventSynchronize(copying.front()->done_ev));
                    else std::this_thread::sleep_for(std::chrono::microseconds(200));
                }
                continue;
            }
            if (!persistent) publish();
            // one engine step: a few iterations (or a range of work items), then a
            // snapshot of the per-job finished counts
            int k = step_slot;
            step_slot ^= 1;
            if (step_busy[k]) {
                RZK_HIP(hipEventSynchronize(ev_step[k]));
                retire_step(k);
                finish_done(k);
            }
            hipStream_t st = ctx->stream;
            bool launched = true;
            if (persistent) {
                // the next range of work items
                launched = next_item < total_items;
                if (launched) {
                    const int64_t a = next_item, e = std::min(total_items, a + step_items);
                    BatchHeader& h = h_hdr[k];
                    std::memset(&h, 0, sizeof(h));
                    h.jobs = d_jobs;
                    h.n_jobs = n_jobs;
                    h.n_items = e;
                    h.samples = (float*)arena.base;
                    h.counter = d_head;
                    h.items_per_dequeue = std::max(1, prm.launch.items_per_dequeue);
                    h.job_lo = oldest_open_job();
                    h.job_mask = kJobRing - 1;
                    h.item_lo = a;
                    RZK_HIP(hipMemcpyAsync(d_hdr + k, &h, sizeof(BatchHeader), hipMemcpyHostToDevice, st));
                    RZK_HIP(hipEventRecord(ev_t0[k], st));
                    launch_batch_items(st, sv, d_hdr + k, grid, block, h.items_per_dequeue);
                    RZK_HIP(hipGetLastError());
                    RZK_HIP(hipEventRecord(ev_step[k], st));
                    step_end[k] = e;
                    next_item = e;
                }
            } else {
                RZK_HIP(hipEventRecord(ev_t0[k], st));
                wf_iterate(wf, iters);
                RZK_HIP(hipMemcpyAsync(h_done + k * kJobRing, d_done, sizeof(unsigned int) * kJobRing,
                                       hipMemcpyDeviceToHost, st));
                RZK_HIP(hipEventRecord(ev_step[k], st));
            }
            if (launched) {
                step_busy[k] = true;
                step_jobs[k] = n_jobs;
                std::lock_guard<std::mutex> lk(mu);
                stats.rounds++;
            }
            // act on the previous step's snapshot while this one runs
            int prev = k ^ 1;
            if (step_busy[prev]) {
                RZK_HIP(hipEventSynchronize(ev_step[prev]));
                retire_step(prev);
                finish_done(prev);
            }
        }
    }

    void run() {
        try {
            RZK_HIP(hipSetDevice(ctx->device));
            loop();
        } catch (const RzkError& e) {
            fail(e.msg);
        } catch (const std::exception& e) {
            fail(e.what());
        }
    }

    void fail(const std::string& msg) {
        std::lock_guard<std::mutex> lk(mu);
        error = msg.empty() ? "server worker failed" : msg;
        for (auto& kv : reqs)
            if (kv.second->state < 3) kv.second->state = 4;
        cv_done.notify_all();
    }

    ~rzk_server() {
        stop = true;
        cv_work.notify_all();
        if (worker.joinable()) worker.join();
        if (ctx) (void)hipStreamSynchronize(ctx->stream);
        if (wf) wf_destroy(wf);
        if (d_spill) (void)hipFree(d_spill);
        if (d_hdr) (void)hipFree(d_hdr);
        if (d_head) (void)hipFree(d_head);
        if (h_hdr) (void)hipHostFree(h_hdr);
        for (int k = 0; k < 2; ++k) {
            if (ev_step[k]) (void)hipEventDestroy(ev_step[k]);
            if (ev_t0[k]) (void)hipEventDestroy(ev_t0[k]);
        }
        for (auto& kv : reqs) {
            if (kv.second->done_ev) (void)hipEventDestroy(kv.second->done_ev);
            if (kv.second->host_img) (void)hipHostFree(kv.second->host_img);
        }
        for (auto& p : pinned_pool) (void)hipHostFree(p.first);
        if (d_jobs) (void)hipFree(d_jobs);
        if (d_done) (void)hipFree(d_done);
        if (h_jobs) (void)hipHostFree(h_jobs);
        if (h_done) (void)hipHostFree(h_done);
        if (ctx) rzk_ctx_destroy(ctx);
    }
};

namespace {

template <typename F>
int serve_guard(rzk_server* s, F&& f) {
    try {
        f();
        return RZK_OK;
    } catch (const RzkError& e) {
        if (s) { std::lock_guard<std::mutex> lk(s->mu); s->error = e.msg; }
        return e.code;
    } catch (const std::bad_alloc&) {
        if (s) { std::lock_guard<std::mutex> lk(s->mu); s->error = "host allocation failed"; }
        return RZK_E_OOM;
    } catch (const std::exception& e) {
        if (s) { std::lock_guard<std::mutex> lk(s->mu); s->error = e.what(); }
        return RZK_E_INVALID;
    }
}

}  // namespace

extern "C" {

int rzk_server_create(int device, const rzk_scene* scene, const rzk_clusters* cl, const rzk_server_params* prm,
                      rzk_server** out) {
    if (!scene || !cl || !prm || !out) return RZK_E_INVALID;
    *out = nullptr;
    std::unique_ptr<rzk_server> s(new rzk_server());
    int rc = rzk_ctx_create(device, &s->ctx);
    if (rc != RZK_OK) return rc;
    rc = serve_guard(s.get(), [&] {
        RZK_CHECK(cl->scene == scene, RZK_E_INVALID, "clusters belong to another scene");
        RZK_CHECK(prm->memory_budget > 0, RZK_E_INVALID, "bad server parameters");
        const rzk_batch_config& c = prm->launch;
        RZK_CHECK((c.block == 64 || c.block == 128 || c.block == 192 || c.block == 256) && c.blocks_per_cu >= 1 &&
                      c.blocks_per_cu <= 64 && c.path_slots >= 4096 && c.path_slots <= (1 << 26) &&
                      (c.engine == 0 || c.engine == 1) && c.items_per_dequeue >= 1 && c.items_per_dequeue <= 64,
                  RZK_E_INVALID, "bad launch shape");
        s->scene = scene;
        s->cl = cl;
        s->prm = *prm;
        s->gp = prm->use_cluster_params ? cl->gp : prm->guide;
        RZK_CHECK(s->gp.tile > 0 && s->gp.map_resolution > 0 && s->gp.map_resolution <= 64, RZK_E_INVALID,
                  "bad guide parameters");
        s->t0 = std::chrono::steady_clock::now();
        s->arena.init((size_t)prm->memory_budget);
        RZK_HIP(hipMalloc((void**)&s->d_jobs, sizeof(JobGpu) * kJobRing));
        RZK_HIP(hipMalloc((void**)&s->d_done, sizeof(unsigned int) * kJobRing));
        RZK_HIP(hipMemset(s->d_done, 0, sizeof(unsigned int) * kJobRing));
        RZK_HIP(hipHostMalloc((void**)&s->h_jobs, sizeof(JobGpu) * kJobRing, hipHostMallocDefault));
        RZK_HIP(hipHostMalloc((void**)&s->h_done, 2 * sizeof(unsigned int) * kJobRing, hipHostMallocDefault));
        for (int k = 0; k < 2; ++k) {
            RZK_HIP(hipEventCreate(&s->ev_step[k]));
            RZK_HIP(hipEventCreate(&s->ev_t0[k]));
        }
        s->persistent = c.engine == 0;
        if (s->persistent) {
            s->block = c.block;
            s->grid = std::max(1, s->ctx->num_cus) * c.blocks_per_cu;
            const int64_t threads = (int64_t)s->grid * s->block;
            s->step_items = (int64_t)std::max(1, std::min(kMaxStepIterations, prm->step_iterations)) * (threads / 64);
            s->sv = scene_view(scene);
            const int64_t extra = 2 * (int64_t)s->sv.stack4 - kPtLdsStack;   // closest_hit4 pairs beyond LDS
            if (extra > 0) {
                RZK_HIP(hipMalloc((void**)&s->d_spill, sizeof(int) * (size_t)(extra * threads)));
                s->sv.spill = s->d_spill;
                s->sv.spill_pitch = threads;
            } else {
                s->sv.spill = nullptr;
                s->sv.spill_pitch = 0;
            }
            RZK_HIP(hipMalloc((void**)&s->d_hdr, 2 * sizeof(BatchHeader)));
            RZK_HIP(hipMalloc((void**)&s->d_head, sizeof(unsigned long long)));
            RZK_HIP(hipHostMalloc((void**)&s->h_hdr, 2 * sizeof(BatchHeader), hipHostMallocDefault));
        } else {
            s->wf = wf_create(s->ctx, scene, c.path_slots, c.block, c.blocks_per_cu);
            wf_reset(s->wf);
            s->publish();
        }
        RZK_HIP(hipStreamSynchronize(s->ctx->stream));
        rzk_server* raw = s.get();
        s->worker = std::thread([raw] { raw->run(); });
    });
    if (rc != RZK_OK) return rc;
    *out = s.release();
    return RZK_OK;
}

int rzk_server_destroy(rzk_server* s) {
    delete s;
    return RZK_OK;
}

int rzk_server_submit(rzk_server* s, const rzk_request* r, int64_t* ticket) {
    if (!s || !r || !ticket) return RZK_E_INVALID;
    return serve_guard(s, [&] {
        std::unique_ptr<Request> q(new Request());
        q->r = *r;
        q->W = (int)r->camera[17];
        q->H = (int)r->camera[18];
        RZK_CHECK(q->W > 0 && q->H > 0 && r->spp > 0 && r->max_bounce >= 0 &&
                      (r->guide_mode == 0 || r->guide_mode == 1),
                  RZK_E_INVALID, "bad request");
        q->L = guide_layout(q->W, q->H, s->gp.tile, s->gp.map_resolution, s->cl->C);
        q->bytes_img = sizeof(float) * 3 * (size_t)q->W * q->H;
        q->bytes_samples = q->bytes_img * (size_t)r->spp;
        RZK_CHECK(request_bytes(q->L, q->bytes_samples, q->bytes_img) <= s->arena.size, RZK_E_OOM,
                  "request needs more device memory than the whole budget");
        RZK_HIP(hipEventCreateWithFlags(&q->done_ev, hipEventDisableTiming));
        std::lock_guard<std::mutex> lk(s->mu);
        RZK_CHECK(s->error.empty(), RZK_E_HIP, s->error);
        q->ticket = s->next_ticket++;
        q->t_submit = s->now();
        *ticket = q->ticket;
        s->incoming.push_back(q.get());
        s->reqs[q->ticket] = std::move(q);
        s->stats.submitted++;
        s->cv_work.notify_one();
    });
}

int rzk_server_wait(rzk_server* s, int64_t ticket, double timeout_s, float* out_rgb, rzk_request_stats* st) {
    if (!s) return RZK_E_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return RZK_E_INVALID;
    Request* q = it->second.get();
    auto ready = [&] { return q->state >= 3; };
    if (timeout_s < 0) s->cv_done.wait(lk, ready);
    else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) return RZK_E_STATE;
    if (q->state == 4) return RZK_E_HIP;
    if (out_rgb) std::memcpy(out_rgb, q->host_img, q->bytes_img);
    if (st) {
        st->submitted = q->t_submit;
        st->admitted = q->t_admit;
        st->finished = q->t_finish;
        st->rounds = q->steps;
        st->pad = 0;
        st->device_bytes = q->held_bytes();
    }
    s->pinned_pool.push_back({q->host_img, q->host_bytes});
    q->host_img = nullptr;
    (void)hipEventDestroy(q->done_ev);
    q->done_ev = nullptr;
    s->reqs.erase(it);
    return RZK_OK;
}

int rzk_server_get_stats(rzk_server* s, rzk_server_stats* out) {
    if (!s || !out) return RZK_E_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    out->now = s->now();
    return RZK_OK;
}

const char* rzk_server_last_error(rzk_server* s) {
    if (!s) return "null server";
    std::lock_guard<std::mutex> lk(s->mu);
    return s->error.c_str();
}

}  // extern "C"
