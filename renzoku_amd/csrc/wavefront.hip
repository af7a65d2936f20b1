// wavefront.hip — host side of the wavefront path tracing engine: slot state,
// queues and the two captured iteration graphs (one per queue parity), and
// the run loop that replays them until every path of the job table is done.
//
// The device side (pathtrace.hip, k_wf_*) splits trace_wellington
// (vpl.cc:934-1041) at its traversals, so the engine produces the same
// per-sample radiance as the one-lane-per-pixel kernel while its kernels stay
// small: traversal kernels carry only the ray and a short stack, shading
// kernels no stack at all.

#include <algorithm>
#include <cstring>
#include <memory>

#include "pathtrace.h"

using namespace rzk;

struct WfEngine {
    rzk_ctx* ctx = nullptr;
    const rzk_scene* scene = nullptr;
    SceneView sv{};
    int64_t P = 0;
    int block = 256, grid = 0;
    std::vector<void*> bufs;
    WfView v{};
    WfJobs* d_jh = nullptr;
    WfJobs* h_jh = nullptr;                 // 2 pinned staging entries, reused after their copy completed
    hipEvent_t jh_ev[2] = {nullptr, nullptr};
    int jh_next = 0;
    int parity = 0;                         // queue parity of the next iteration
    unsigned long long* h_ctrl = nullptr;   // 2 x {handed out, active}
    int* h_fc = nullptr;                    // initial free counts {P, 0}
    int* all_slots = nullptr;               // 0 .. P-1
    hipGraph_t g[2] = {nullptr, nullptr};
    hipGraphExec_t ex[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};

    template <typename T>
    T* alloc(size_t count) {
        void* p = nullptr;
        RZK_HIP(hipMalloc(&p, sizeof(T) * std::max<size_t>(count, 1)));
        bufs.push_back(p);
        return static_cast<T*>(p);
    }

    ~WfEngine() {
        if (ctx) (void)hipStreamSynchronize(ctx->stream);
        for (int k = 0; k < 2; ++k) {
            if (ex[k]) (void)hipGraphExecDestroy(ex[k]);
            if (g[k]) (void)hipGraphDestroy(g[k]);
            if (ev[k]) (void)hipEventDestroy(ev[k]);
            if (jh_ev[k]) (void)hipEventDestroy(jh_ev[k]);
        }
        for (void* p : bufs) (void)hipFree(p);
        if (h_jh) (void)hipHostFree(h_jh);
        if (h_ctrl) (void)hipHostFree(h_ctrl);
        if (h_fc) (void)hipHostFree(h_fc);
    }
};

WfEngine* wf_create(rzk_ctx* ctx, const rzk_scene* scene, int64_t slots, int block, int blocks_per_cu) {
    std::unique_ptr<WfEngine> e(new WfEngine());
    e->ctx = ctx;
    e->scene = scene;
    e->P = std::max<int64_t>(slots, 64);
    e->block = block;
    e->grid = std::max(1, ctx->num_cus) * blocks_per_cu;
    const int64_t P = e->P;
    WfView& v = e->v;
    v.P = P;
    v.job = e->alloc<int>(P);
    v.pix = e->alloc<int>(P);
    v.smp = e->alloc<int>(P);
    v.bounce = e->alloc<int>(P);
    v.flags = e->alloc<int>(P);
    v.ro = e->alloc<float>(3 * P);
    v.rd = e->alloc<float>(3 * P);
    v.T = e->alloc<float>(3 * P);
    v.Tc = e->alloc<float>(3 * P);
    v.L = e->alloc<float>(3 * P);
    v.hit = e->alloc<float4>(P);
    v.np = e->alloc<float>(3 * P);
    v.nq = e->alloc<float>(3 * P);
    v.lo1 = e->alloc<float>(3 * P);
    v.nwi = e->alloc<float>(3 * P);
    v.nd = e->alloc<float>(3 * P);
    v.nbr = e->alloc<float>(3 * P);
    v.ndpdf = e->alloc<float>(P);
    v.ncos = e->alloc<float>(P);
    v.nk = e->alloc<int>(P);
    v.nhit = e->alloc<float4>(P);
    v.blocked = e->alloc<int>(P);
    for (int par = 0; par < 2; ++par)
        for (int q = 0; q < 4; ++q) v.q[par][q] = e->alloc<int>(P);
    v.counts = e->alloc<int>(8);
    v.freeq[0] = e->alloc<int>(P);
    v.freeq[1] = e->alloc<int>(P);
    v.free_counts = e->alloc<int>(2);
    v.ctrl = e->alloc<unsigned long long>(2);
    e->d_jh = e->alloc<WfJobs>(1);
    RZK_HIP(hipHostMalloc((void**)&e->h_jh, 2 * sizeof(WfJobs), hipHostMallocDefault));
    for (int k = 0; k < 2; ++k) RZK_HIP(hipEventCreateWithFlags(&e->jh_ev[k], hipEventDisableTiming));
    RZK_HIP(hipHostMalloc((void**)&e->h_ctrl, 4 * sizeof(unsigned long long), hipHostMallocDefault));
    RZK_HIP(hipHostMalloc((void**)&e->h_fc, 2 * sizeof(int), hipHostMallocDefault));
    e->h_fc[0] = (int)P;
    e->h_fc[1] = 0;
    e->all_slots = e->alloc<int>(P);
    {
        std::vector<int> iota((size_t)P);
        for (int64_t i = 0; i < P; ++i) iota[i] = (int)i;
        RZK_HIP(hipMemcpy(e->all_slots, iota.data(), sizeof(int) * P, hipMemcpyHostToDevice));
    }
    RZK_HIP(hipMemset(v.flags, 0, sizeof(int) * P));
    RZK_HIP(hipMemset(v.counts, 0, sizeof(int) * 8));
    RZK_HIP(hipMemset(v.ctrl, 0, sizeof(unsigned long long) * 2));
    // traversal stacks: 16 entries per lane in LDS, the rest of 2 * stack4 in HBM
    e->sv = scene_view(scene);
    int64_t threads = (int64_t)e->grid * block;
    int64_t extra = 2 * (int64_t)e->sv.stack4 - kLdsStack;
    if (extra > 0) {
        e->sv.spill = e->alloc<int>((size_t)(extra * threads));
        e->sv.spill_pitch = threads;
    } else {
        e->sv.spill = nullptr;
        e->sv.spill_pitch = 0;
    }
    hipStream_t st = ctx->stream;
    for (int par = 0; par < 2; ++par) {
        RZK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        launch_wf_iteration(st, e->sv, v, e->d_jh, par, e->grid, block);
        hipError_t le = hipGetLastError();
        hipGraph_t gr = nullptr;
        hipError_t ee = hipStreamEndCapture(st, &gr);
        RZK_HIP(le);
        RZK_HIP(ee);
        e->g[par] = gr;
        RZK_HIP(hipGraphInstantiate(&e->ex[par], gr, nullptr, nullptr, 0));
        RZK_HIP(hipEventCreateWithFlags(&e->ev[par], hipEventDisableTiming));
    }
    return e.release();
}

void wf_destroy(WfEngine* e) { delete e; }

int64_t wf_slots(const WfEngine* e) { return e->P; }

void wf_reset(WfEngine* e) {
    hipStream_t st = e->ctx->stream;
    RZK_HIP(hipMemsetAsync(e->v.flags, 0, sizeof(int) * e->P, st));
    RZK_HIP(hipMemsetAsync(e->v.counts, 0, sizeof(int) * 8, st));
    RZK_HIP(hipMemsetAsync(e->v.ctrl, 0, sizeof(unsigned long long) * 2, st));
    // every slot starts free, queued for the first iteration (parity 0)
    RZK_HIP(hipMemcpyAsync(e->v.freeq[0], e->all_slots, sizeof(int) * e->P, hipMemcpyDeviceToDevice, st));
    RZK_HIP(hipMemcpyAsync(e->v.free_counts, e->h_fc, 2 * sizeof(int), hipMemcpyHostToDevice, st));
    e->parity = 0;
}

void wf_publish(WfEngine* e, const WfJobs& jobs) {
    RZK_CHECK(jobs.job_cap > 0 && (jobs.job_cap & (jobs.job_cap - 1)) == 0, RZK_E_INVALID,
              "job ring capacity must be a power of two");
    hipStream_t st = e->ctx->stream;
    int k = e->jh_next;
    e->jh_next ^= 1;
    RZK_HIP(hipEventSynchronize(e->jh_ev[k]));     // the staging entry's previous copy is done
    e->h_jh[k] = jobs;
    RZK_HIP(hipMemcpyAsync(e->d_jh, e->h_jh + k, sizeof(WfJobs), hipMemcpyHostToDevice, st));
    RZK_HIP(hipEventRecord(e->jh_ev[k], st));
}

void wf_iterate(WfEngine* e, int n) {
    for (int i = 0; i < n; ++i) {
        RZK_HIP(hipGraphLaunch(e->ex[e->parity], e->ctx->stream));
        e->parity ^= 1;
    }
}

// Every path of the job table, to completion.  d_jobs[j].item_first is the
// job's first path index; paths of a job are sample-major.  Returns the
// number of iterations.
int64_t wf_run(WfEngine* e, const JobGpu* d_jobs, int64_t n_jobs, int64_t total_paths, int max_bounce,
               float* d_samples, unsigned long long* d_work, unsigned int* d_job_done) {
    if (total_paths <= 0 || n_jobs <= 0) return 0;
    // a path finishes at most max_bounce + 2 iterations after it is generated,
    // and each such window frees every slot at least once
    const int64_t bound = (total_paths / e->P + 2) * (int64_t)(std::max(max_bounce, 0) + 3) + 16;
    hipStream_t st = e->ctx->stream;
    int64_t cap = 1;
    while (cap < n_jobs) cap <<= 1;
    wf_reset(e);
    wf_publish(e, WfJobs{d_jobs, 0, n_jobs, cap, total_paths, d_samples, d_work, d_job_done});
    int64_t it = 0;
    for (;; ++it) {
        int k = (int)(it & 1);
        wf_iterate(e, 1);
        RZK_HIP(hipMemcpyAsync(e->h_ctrl + 2 * k, e->v.ctrl, 2 * sizeof(unsigned long long),
                               hipMemcpyDeviceToHost, st));
        RZK_HIP(hipEventRecord(e->ev[k], st));
        if (it == 0) continue;
        // counters after the previous iteration (one iteration stays queued)
        int prev = k ^ 1;
        RZK_HIP(hipEventSynchronize(e->ev[prev]));
        unsigned long long handed = e->h_ctrl[2 * prev], active = e->h_ctrl[2 * prev + 1];
        if (handed >= (unsigned long long)total_paths && active == 0) break;
        if (it > bound) {
            (void)hipStreamSynchronize(st);
            throw RzkError{RZK_E_STATE, "wavefront engine did not drain its paths"};
        }
    }
    RZK_HIP(hipStreamSynchronize(st));
    return it + 1;
}
