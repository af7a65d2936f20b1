// pathtrace.h — parameters of the guided path tracer (pathtrace.hip).
#pragma once

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

}  // namespace rzk

void launch_guided_pt(rzk_ctx* ctx, const SceneView& sv, const float* cam, const rzk::PtParams& pp,
                      const rzk::GuideView& gv, float* d_out, unsigned long long* d_work);
