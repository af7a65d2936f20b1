// pathtrace.h — parameters of the guided path tracer (pathtrace.hip), the
// guide handle shared by the plain, batched and served forms, and the batch
// engine's job table (batch.hip).
#pragma once

#include <vector>

#include "rzk_internal.h"

namespace rzk {

// Traversal stack entries per lane kept in LDS by the path kernels (deeper
// ones spill to HBM); with the LDS state columns it sets how many waves fit
// on a compute unit.
#ifndef RZK_PT_LDS_STACK
#define RZK_PT_LDS_STACK 12
#endif
constexpr int kPtLdsStack = RZK_PT_LDS_STACK;

struct PtParams {
    int width, height, spp, max_bounce, direct, guide_mode, tile, res, tiles_x, tiles_y;
    int x0, y0, w, h;              // rendered window of the image
    uint32_t seed;
    float tmin;
    const int* pix_slice = nullptr;   // kd-tree slices: each pixel's slice (-1: none), else tiles
};

struct GuideView {
    const float* maps;             // [slice][res][res]
    const float* row_cdf;          // [slice][res]
    const float* col_cdf;          // [slice][res][res]
    const float* totals;           // [slice]
};

struct CamView { float v[RZK_CAM_FLOATS]; };

// One render job of a batch as the persistent kernel reads it.  Work items of
// a job are (64-pixel block, block of samples_per_item samples), numbered
// item_first + pixel_block * samp_blocks + sample_block.
struct JobGpu {
    CamView cam;
    PtParams pp;
    GuideView gv;
    int64_t item_first;            // first work item of the job
    int64_t pix_first;             // first output pixel of the job (jobs concatenated)
    int64_t sample_off;            // first float of the job's per-sample radiance [spp][pixel][3]
    int pix_blocks, samp_blocks, samples_per_item, strata;
};

// Device-resident description of the batch in flight; every kernel of the
// captured graph reads sizes and buffers from here, so one graph serves any
// batch that fits the buffers.
struct BatchHeader {
    const JobGpu* jobs;            // job j at jobs[j & job_mask] (a ring when job_mask < n_jobs)
    int64_t n_jobs, n_items, n_pixels;
    float* samples;                // per-sample radiance
    float* out;                    // per-pixel mean, jobs concatenated, [h][w][3] each
    unsigned long long* counter;   // work-queue head
    unsigned long long* work;      // {closest rays, shadow rays} or null
    int items_per_dequeue, pad;
    int64_t job_lo;                // first job whose items may be in the launch
    int64_t job_mask;
    int64_t item_lo;               // the launch runs items item_lo .. n_items - 1
};

// Wavefront engine: path state in structure-of-arrays slots ([3][P] planar
// vectors), four compacted queues per iteration parity.
struct WfView {
    int64_t P;
    int *job, *pix, *smp, *bounce, *flags;
    float *ro, *rd, *T, *Tc, *L;        // [3][P]
    float4* hit;                        // current ray: t, beta, gamma, tri bits
    float *np, *nq, *lo1, *nwi, *nd, *nbr;   // next-event data, [3][P]
    float *ndpdf, *ncos;
    int* nk;
    float4* nhit;                       // next-event BSDF ray hit
    int* blocked;
    int* q[2][4];                       // [parity][extend, nee ray, shadow, nee resolve]
    int* counts;                        // [parity][4]
    int* freeq[2];                      // slots freed in an iteration, taken by the next one
    int* free_counts;                   // [parity]
    unsigned long long* ctrl;           // {paths handed out, active slots}
};

// The job table in flight (read by every kernel of the captured iteration).
// Jobs are numbered 0, 1, 2, ... for the life of the table; job j lives in
// ring entry j & (job_cap - 1), and paths are numbered across jobs in order
// (item_first = the job's first path, paths sample-major within a job).  New
// jobs may be appended between iterations; only jobs [job_lo, n_jobs) can
// still hand out paths.
struct WfJobs {
    const JobGpu* jobs;                 // ring of job_cap entries (a power of two)
    int64_t job_lo, n_jobs, job_cap;
    int64_t total_paths;
    float* samples;                     // base of every job's per-sample radiance ([spp][pixel][3] at sample_off)
    unsigned long long* work;           // {closest rays, shadow rays} or null
    unsigned int* job_done;             // finished paths per ring entry, or null
};

}  // namespace rzk

// A guide: the per-slice guiding distributions of one camera, borrowed from a
// LightSlice or owned (built from resident clusters).
struct rzk_guide {
    rzk_ctx* ctx = nullptr;
    const rzk_scene* scene = nullptr;
    float cam[RZK_CAM_FLOATS];
    int W = 0, H = 0, tile = 0, tiles_x = 0, tiles_y = 0, res = 0;
    rzk::GuideView gv{};
    std::vector<void*> owned;      // device buffers freed with the guide
    int64_t device_bytes = 0;
};

// Resident cluster set: the kept representatives as a compact VPL set (their
// own Morton order), weights, identity indices, and the guide parameters of
// the LightSlice they came from.
struct rzk_clusters {
    rzk_ctx* ctx = nullptr;
    const rzk_scene* scene = nullptr;
    rzk_vpls* reps = nullptr;      // compact representative VPLs
    float* d_rep_w = nullptr;
    int* d_ident = nullptr;        // 0..C-1
    int C = 0;
    float n_emitted = 1.f;
    rzk_guide_params gp{};
    std::vector<int> h_reps;       // indices into the original VPL set
    std::vector<float> h_rep_w;
};

// Device buffers of one guide build.  Temporaries are dead once the build's
// kernels have run; the tables are the guide.
struct GuideBuffers {
    int* px = nullptr; int* py = nullptr;
    float* rays = nullptr; float4* hits = nullptr;
    float* recv = nullptr; int* mat = nullptr; int* light = nullptr;
    float* R = nullptr; float* V = nullptr; float* slice_wi = nullptr; int* upper = nullptr; float* row_pdf = nullptr;
    float* maps = nullptr; float* col_cdf = nullptr; float* row_cdf = nullptr; float* totals = nullptr;
};

struct GuideLayout {
    int S = 0, tx = 0, ty = 0;
    // byte offsets of every buffer within two regions (temporaries, tables), 256-B aligned
    size_t off_tmp[12], off_tab[4];
    size_t bytes_tmp = 0, bytes_tab = 0;
};

GuideLayout guide_layout(int W, int H, int tile, int res, int C);
GuideBuffers guide_buffers(const GuideLayout& L, char* tmp, char* tab);
// enqueue the guide build on ctx's stream (no synchronisation)
void enqueue_guide_build(rzk_ctx* ctx, const rzk_scene* scene, const rzk_clusters& cl, const float* cam,
                         const rzk_guide_params& gp, const GuideLayout& L, const GuideBuffers& b);
// fill a guide handle's camera, sizes and views over the tables of `b`
void guide_set(rzk_guide& g, const rzk_scene* scene, const float* cam, const rzk_guide_params& gp,
               const GuideLayout& L, const GuideBuffers& b);

void launch_guided_pt_rays(rzk_ctx* ctx, const SceneView& sv, const rzk::PtParams& pp, const rzk::GuideView& gv,
                           const float* d_rays, const int* d_pix, const int* d_smp, int64_t n, float* d_out);
void launch_guided_pt(rzk_ctx* ctx, const SceneView& sv, const float* cam, const rzk::PtParams& pp,
                      const rzk::GuideView& gv, float* d_out, unsigned long long* d_work);
// the three kernels of a batch on `stream` (captured into a graph by the caller)
// (items_per_dequeue: the header's; it picks the persistent kernel)
void launch_batch(hipStream_t stream, const SceneView& sv, const rzk::BatchHeader* d_hdr, int grid, int block,
                  int items_per_dequeue, int reduce_grid);
// the queue reset and the persistent kernel alone (items item_lo .. n_items - 1)
void launch_batch_items(hipStream_t stream, const SceneView& sv, const rzk::BatchHeader* d_hdr, int grid, int block,
                        int items_per_dequeue);
// one iteration of the wavefront engine (queues of parity `in` consumed,
// parity in ^ 1 produced): extend + next-event rays, shadow rays, next-event
// resolution, bounce shading, path regeneration
void launch_wf_iteration(hipStream_t stream, const SceneView& sv, const rzk::WfView& v, const rzk::WfJobs* d_jh,
                         int in, int grid, int block);
// the wavefront engine (wavefront.hip): `slots` paths in flight, kernels of
// `block` lanes on compute units x blocks_per_cu workgroups
struct WfEngine;
WfEngine* wf_create(rzk_ctx* ctx, const rzk_scene* scene, int64_t slots, int block, int blocks_per_cu);
void wf_destroy(WfEngine* e);
int64_t wf_slots(const WfEngine* e);
// run every path of a device job table (item_first = first path index) to
// completion on ctx's stream; max_bounce bounds the iterations; returns them
int64_t wf_run(WfEngine* e, const rzk::JobGpu* d_jobs, int64_t n_jobs, int64_t total_paths, int max_bounce,
               float* d_samples, unsigned long long* d_work, unsigned int* d_job_done);
// incremental use (a request loop): reset the slots, publish a job table
// (queued on the stream), run iterations, read {paths handed out, active
// slots} as of the last completed wf_counters_async
void wf_reset(WfEngine* e);
void wf_publish(WfEngine* e, const rzk::WfJobs& jobs);
void wf_iterate(WfEngine* e, int n);
// each job's finished paths: k_job_reduce sums a job's samples per pixel in
// sample order into out ([h][w][3] means)
void launch_job_reduce(hipStream_t stream, const rzk::JobGpu* d_job, const float* samples, float* out);
// per-pixel means of finished per-sample radiance (k_batch_reduce)
void launch_batch_reduce(hipStream_t stream, const rzk::BatchHeader* d_hdr, int grid);
// PtParams of a job over a guide
rzk::PtParams job_params(const rzk_guide& g, const rzk_render_job& j);
void check_job(const rzk_guide* g, const rzk_render_job& j);
