// pathtrace.h — parameters of the guided path tracer (pathtrace.hip), the
// guide handle shared by the plain, batched and served forms, and the batch
// engine's job table (batch.hip).
#pragma once

#include <vector>

#include "rzk_internal.h"

namespace rzk {

struct PtParams {
    int width, height, spp, max_bounce, direct, guide_mode, tile, res, tiles_x, tiles_y;
    int x0, y0, w, h;              // rendered window of the image
    uint32_t seed;
    float tmin;
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
    int64_t sample_off;            // first float of the job's per-sample radiance [pixel][spp][3]
    int pix_blocks, samp_blocks, samples_per_item, strata;
};

// Device-resident description of the batch in flight; every kernel of the
// captured graph reads sizes and buffers from here, so one graph serves any
// batch that fits the buffers.
struct BatchHeader {
    const JobGpu* jobs;
    int64_t n_jobs, n_items, n_pixels;
    float* samples;                // per-sample radiance
    float* out;                    // per-pixel mean, jobs concatenated, [h][w][3] each
    unsigned long long* counter;   // work-queue head
    unsigned long long* work;      // {closest rays, shadow rays} or null
    int items_per_dequeue, pad;
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

void launch_guided_pt(rzk_ctx* ctx, const SceneView& sv, const float* cam, const rzk::PtParams& pp,
                      const rzk::GuideView& gv, float* d_out, unsigned long long* d_work);
// the three kernels of a batch on `stream` (captured into a graph by the caller)
void launch_batch(hipStream_t stream, const SceneView& sv, const rzk::BatchHeader* d_hdr, int grid, int block,
                  int reduce_grid);
// PtParams of a job over a guide
rzk::PtParams job_params(const rzk_guide& g, const rzk_render_job& j);
void check_job(const rzk_guide* g, const rzk_render_job& j);
