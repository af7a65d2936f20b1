// serve.hip — the served form: a request loop over a resident scene, VPL
// clusters and memory budget.  Requests (camera, spp, depth, seed) are queued
// by any thread; a worker thread admits them oldest first while the budget
// allows (guide tables + running sums + image), builds each admitted
// request's guide from the resident clusters, and runs rounds: each round is
// one graph launch of a persistent kernel over up to round_items work items
// drawn from all admitted requests, oldest first.  Requests that arrive while
// a round runs join the next one; finished images are copied out
// asynchronously and their memory returns to the budget.
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
    int rounds = 0;
    // admitted state
    GuideLayout L;
    size_t off_tab = 0, off_tmp = 0, off_acc = 0, off_out = 0;
    size_t bytes_img = 0;
    rzk_guide guide;
    int64_t pix_blocks = 0, samp_blocks = 0, total_items = 0, next_item = 0;
    int S = 1, strata = 0;
    hipEvent_t done_ev = nullptr;
    float* host_img = nullptr;       // pinned, filled by the final copy
    size_t host_bytes = 0;
    int state = 0;                   // 0 queued, 1 active, 2 copying, 3 finished, 4 failed
    int64_t held_bytes() const { return (int64_t)(L.bytes_tab + 2 * bytes_img); }
};

size_t request_bytes(const GuideLayout& L, size_t img) { return L.bytes_tab + L.bytes_tmp + 2 * img + 4 * 256; }

}  // namespace

struct rzk_server {
    rzk_ctx* ctx = nullptr;                // owned: the server's stream
    const rzk_scene* scene = nullptr;
    const rzk_clusters* cl = nullptr;
    rzk_server_params prm{};
    rzk_guide_params gp{};
    Arena arena;
    SceneView sv{};
    int* spill = nullptr;
    // two round slots: device job table + header, pinned staging, graph
    ServeJobGpu* d_jobs[2] = {nullptr, nullptr};
    ServeHeader* d_hdr[2] = {nullptr, nullptr};
    ServeJobGpu* h_jobs[2] = {nullptr, nullptr};
    ServeHeader* h_hdr[2] = {nullptr, nullptr};
    size_t job_cap = 0;
    unsigned long long* d_counter = nullptr;
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipEvent_t ev_start[2] = {nullptr, nullptr}, ev_end[2] = {nullptr, nullptr};
    bool slot_busy[2] = {false, false};
    // request bookkeeping
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> incoming;
    std::unordered_map<int64_t, std::unique_ptr<Request>> reqs;
    std::deque<Request*> waiting, active, copying;
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
    int grid() const { return std::max(1, ctx->num_cus) * prm.launch.blocks_per_cu; }

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

    void capture() {
        for (int k = 0; k < 2; ++k) {
            hipStream_t st = ctx->stream;
            RZK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            launch_serve_round(st, sv, d_hdr[k], grid(), prm.launch.block);
            hipError_t le = hipGetLastError();
            hipGraph_t g = nullptr;
            hipError_t ee = hipStreamEndCapture(st, &g);
            RZK_HIP(le);
            RZK_HIP(ee);
            graph[k] = g;
            RZK_HIP(hipGraphInstantiate(&exec[k], g, nullptr, nullptr, 0));
        }
    }

    void ensure_job_cap(size_t n) {
        if (n <= job_cap) return;
        size_t cap = std::max(n, job_cap * 2);
        for (int k = 0; k < 2; ++k) {
            if (slot_busy[k]) { RZK_HIP(hipEventSynchronize(ev_end[k])); retire_slot(k); }
            if (d_jobs[k]) RZK_HIP(hipFree(d_jobs[k]));
            if (h_jobs[k]) RZK_HIP(hipHostFree(h_jobs[k]));
            RZK_HIP(hipMalloc((void**)&d_jobs[k], sizeof(ServeJobGpu) * cap));
            RZK_HIP(hipHostMalloc((void**)&h_jobs[k], sizeof(ServeJobGpu) * cap, hipHostMallocDefault));
        }
        job_cap = cap;
    }

    void retire_slot(int k) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ev_start[k], ev_end[k]) == hipSuccess) {
            std::lock_guard<std::mutex> lk(mu);
            stats.busy_seconds += ms * 1e-3;
        }
        slot_busy[k] = false;
    }

    // admit queued requests oldest first while the budget allows
    void admit() {
        while (!waiting.empty()) {
            Request* q = waiting.front();
            size_t need = request_bytes(q->L, q->bytes_img);
            size_t o_tab, o_tmp, o_acc, o_out;
            if (!arena.alloc(q->L.bytes_tab, o_tab)) break;
            if (!arena.alloc(q->L.bytes_tmp, o_tmp)) { arena.release(o_tab, q->L.bytes_tab); break; }
            if (!arena.alloc(q->bytes_img, o_acc)) {
                arena.release(o_tab, q->L.bytes_tab);
                arena.release(o_tmp, q->L.bytes_tmp);
                break;
            }
            if (!arena.alloc(q->bytes_img, o_out)) {
                arena.release(o_tab, q->L.bytes_tab);
                arena.release(o_tmp, q->L.bytes_tmp);
                arena.release(o_acc, q->bytes_img);
                break;
            }
            (void)need;
            q->off_tab = o_tab; q->off_tmp = o_tmp; q->off_acc = o_acc; q->off_out = o_out;
            GuideBuffers b = guide_buffers(q->L, arena.base + o_tmp, arena.base + o_tab);
            enqueue_guide_build(ctx, scene, *cl, q->r.camera, gp, q->L, b);
            // temporaries are only read by the kernels just queued: later users are stream-ordered after them
            arena.release(o_tmp, q->L.bytes_tmp);
            guide_set(q->guide, scene, q->r.camera, gp, q->L, b);
            q->guide.ctx = ctx;
            q->t_admit = now();
            q->state = 1;
            waiting.pop_front();
            active.push_back(q);
        }
        std::lock_guard<std::mutex> lk(mu);
        stats.peak_bytes = std::max<int64_t>(stats.peak_bytes, (int64_t)arena.peak);
    }

    // fill one slot's job table from the active requests, oldest first
    int64_t plan_round(int k, std::vector<Request*>& finished_now) {
        int64_t budget = prm.round_items, total = 0;
        size_t n = 0;
        std::vector<std::pair<Request*, int64_t>> shares;
        for (Request* q : active) {
            if (budget <= 0) break;
            int64_t left = q->total_items - q->next_item;
            int64_t take = std::min(std::min(left, q->pix_blocks), budget);
            if (take <= 0) continue;
            shares.push_back({q, take});
            budget -= take;
        }
        ensure_job_cap(shares.size());
        for (auto& sh : shares) {
            Request* q = sh.first;
            ServeJobGpu& J = h_jobs[k][n++];
            std::memset(&J, 0, sizeof(J));
            std::memcpy(J.cam.v, q->r.camera, sizeof(J.cam.v));
            rzk_render_job j{&q->guide, 0, 0, q->W, q->H, q->r.spp, q->r.max_bounce, q->r.direct_lighting,
                             q->r.guide_mode, q->r.seed, q->r.tmin};
            J.pp = job_params(q->guide, j);
            J.gv = q->guide.gv;
            J.item_first = total;
            J.item_begin = q->next_item;
            J.pix_blocks = (int)q->pix_blocks;
            J.samples_per_item = q->S;
            J.strata = q->strata;
            J.acc = (float*)(arena.base + q->off_acc);
            J.out = (float*)(arena.base + q->off_out);
            total += sh.second;
            q->next_item += sh.second;
            q->rounds++;
            if (q->next_item == q->total_items) finished_now.push_back(q);
        }
        ServeHeader& h = *h_hdr[k];
        h.jobs = d_jobs[k];
        h.n_jobs = (int64_t)n;
        h.n_items = total;
        h.counter = d_counter;
        h.items_per_dequeue = prm.launch.items_per_dequeue;
        h.pad = 0;
        return total;
    }

    void loop() {
        int slot = 0;
        while (true) {
            {
                std::unique_lock<std::mutex> lk(mu);
                if (stop) break;
                while (!incoming.empty()) { waiting.push_back(incoming.front()); incoming.pop_front(); }
                bool idle = waiting.empty() && active.empty() && copying.empty() && !slot_busy[0] && !slot_busy[1];
                if (idle) {
                    cv_work.wait_for(lk, std::chrono::milliseconds(50));
                    continue;
                }
            }
            // retire finished copies
            while (!copying.empty()) {
                Request* q = copying.front();
                hipError_t e = hipEventQuery(q->done_ev);
                if (e == hipErrorNotReady) break;
                RZK_HIP(e);
                copying.pop_front();
                arena.release(q->off_tab, q->L.bytes_tab);
                arena.release(q->off_acc, q->bytes_img);
                arena.release(q->off_out, q->bytes_img);
                std::lock_guard<std::mutex> lk(mu);
                q->t_finish = now();
                q->state = 3;
                stats.finished++;
                cv_done.notify_all();
            }
            admit();
            // the slot about to be refilled must have finished its previous round
            if (slot_busy[slot]) {
                RZK_HIP(hipEventSynchronize(ev_end[slot]));
                retire_slot(slot);
            }
            std::vector<Request*> fin;
            int64_t items = active.empty() ? 0 : plan_round(slot, fin);
            if (items == 0) {
                // nothing to launch: wait for an outstanding copy or round
                if (!copying.empty()) RZK_HIP(hipEventSynchronize(copying.front()->done_ev));
                else if (slot_busy[slot ^ 1]) { RZK_HIP(hipEventSynchronize(ev_end[slot ^ 1])); retire_slot(slot ^ 1); }
                else std::this_thread::sleep_for(std::chrono::microseconds(200));
                continue;
            }
            hipStream_t st = ctx->stream;
            RZK_HIP(hipMemcpyAsync(d_jobs[slot], h_jobs[slot], sizeof(ServeJobGpu) * h_hdr[slot]->n_jobs,
                                   hipMemcpyHostToDevice, st));
            RZK_HIP(hipMemcpyAsync(d_hdr[slot], h_hdr[slot], sizeof(ServeHeader), hipMemcpyHostToDevice, st));
            RZK_HIP(hipEventRecord(ev_start[slot], st));
            RZK_HIP(hipGraphLaunch(exec[slot], st));
            RZK_HIP(hipEventRecord(ev_end[slot], st));
            slot_busy[slot] = true;
            {
                std::lock_guard<std::mutex> lk(mu);
                stats.rounds++;
                stats.items += items;
            }
            for (Request* q : fin) {
                active.erase(std::find(active.begin(), active.end(), q));
                q->host_bytes = q->bytes_img;
                q->host_img = pinned_get(q->host_bytes);
                if (!q->host_img) RZK_HIP(hipHostMalloc((void**)&q->host_img, q->host_bytes, hipHostMallocDefault));
                RZK_HIP(hipMemcpyAsync(q->host_img, arena.base + q->off_out, q->bytes_img, hipMemcpyDeviceToHost, st));
                RZK_HIP(hipEventRecord(q->done_ev, st));
                q->state = 2;
                copying.push_back(q);
            }
            slot ^= 1;
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
        for (int k = 0; k < 2; ++k) {
            if (exec[k]) (void)hipGraphExecDestroy(exec[k]);
            if (graph[k]) (void)hipGraphDestroy(graph[k]);
            if (ev_start[k]) (void)hipEventDestroy(ev_start[k]);
            if (ev_end[k]) (void)hipEventDestroy(ev_end[k]);
            if (d_jobs[k]) (void)hipFree(d_jobs[k]);
            if (d_hdr[k]) (void)hipFree(d_hdr[k]);
            if (h_jobs[k]) (void)hipHostFree(h_jobs[k]);
            if (h_hdr[k]) (void)hipHostFree(h_hdr[k]);
        }
        for (auto& kv : reqs) {
            if (kv.second->done_ev) (void)hipEventDestroy(kv.second->done_ev);
            if (kv.second->host_img) (void)hipHostFree(kv.second->host_img);
        }
        for (auto& p : pinned_pool) (void)hipHostFree(p.first);
        if (d_counter) (void)hipFree(d_counter);
        if (spill) (void)hipFree(spill);
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
        RZK_CHECK(prm->memory_budget > 0 && prm->round_items > 0, RZK_E_INVALID, "bad server parameters");
        const rzk_batch_config& c = prm->launch;
        RZK_CHECK((c.block == 64 || c.block == 128 || c.block == 192 || c.block == 256) && c.blocks_per_cu >= 1 &&
                      c.blocks_per_cu <= 64 && c.items_per_dequeue >= 1 && c.samples_per_item >= 1,
                  RZK_E_INVALID, "bad launch shape");
        s->scene = scene;
        s->cl = cl;
        s->prm = *prm;
        s->gp = prm->use_cluster_params ? cl->gp : prm->guide;
        RZK_CHECK(s->gp.tile > 0 && s->gp.map_resolution > 0 && s->gp.map_resolution <= 64, RZK_E_INVALID,
                  "bad guide parameters");
        s->t0 = std::chrono::steady_clock::now();
        s->arena.init((size_t)prm->memory_budget);
        s->sv = scene_view(scene);
        int64_t threads = (int64_t)s->grid() * c.block;
        int64_t extra = 2 * (int64_t)s->sv.stack4 - kLdsStack;   // closest_hit4 pairs
        if (extra > 0) {
            RZK_HIP(hipMalloc((void**)&s->spill, sizeof(int) * (size_t)(extra * threads)));
            s->sv.spill = s->spill;
            s->sv.spill_pitch = threads;
        } else {
            s->sv.spill = nullptr;
            s->sv.spill_pitch = 0;
        }
        RZK_HIP(hipMalloc((void**)&s->d_counter, sizeof(unsigned long long)));
        for (int k = 0; k < 2; ++k) {
            RZK_HIP(hipMalloc((void**)&s->d_hdr[k], sizeof(ServeHeader)));
            RZK_HIP(hipHostMalloc((void**)&s->h_hdr[k], sizeof(ServeHeader), hipHostMallocDefault));
            RZK_HIP(hipEventCreate(&s->ev_start[k]));
            RZK_HIP(hipEventCreate(&s->ev_end[k]));
        }
        s->ensure_job_cap(64);
        s->capture();
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
        RZK_CHECK(request_bytes(q->L, q->bytes_img) <= s->arena.size, RZK_E_OOM,
                  "request needs more device memory than the whole budget");
        int64_t npx = (int64_t)q->W * q->H;
        q->pix_blocks = (npx + 63) / 64;
        q->S = std::min(s->prm.launch.samples_per_item, r->spp);
        q->samp_blocks = (r->spp + q->S - 1) / q->S;
        q->total_items = q->pix_blocks * q->samp_blocks;
        int strata = 1;
        while ((strata + 1) * (strata + 1) <= r->spp) ++strata;
        q->strata = strata * strata == r->spp ? strata : 0;
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
        st->rounds = q->rounds;
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
