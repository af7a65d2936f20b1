// gpu_guided_integrator.h — the body of MonteCarloIntegrator::integrate
// (monte_carlo_integrator.cc:45-70) for VirtualPointLight's guided path
// tracing (vpl.cc:907-1041): one bundle (a pixel window of the image,
// sensor.cc:52-79) rendered in one call of the engine and written to the
// frame's pixels.  The guide is a LightSlice's (rzk_lightslice_guide) or one
// built from resident clusters (rzk_guide_create).
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/rzk.h"

namespace Renzoku {

struct Bundle {
    int x = 0, y = 0, width = 0, height = 0;         // pixel window
};

class GpuGuidedPathTracer {
public:
    GpuGuidedPathTracer(rzk_ctx* ctx, const rzk_guide* guide, int width, int height, int spp, int max_bounce,
                        bool direct_lighting, uint32_t seed)
        : ctx_(ctx), guide_(guide), width_(width), height_(height), image_((size_t)width * height * 3) {
        params_.spp = spp;
        params_.max_bounce = max_bounce;
        params_.direct_lighting = direct_lighting ? 1 : 0;
        params_.guide_mode = 1;
        params_.seed = seed;
        params_.tmin = 1e-3f;                        // Scene::get_tmin default (common.h:48)
    }

    // frame: [height][width][3]; only the bundle's pixels change
    void integrate(const Bundle& b, std::vector<float>& frame) {
        rzk_pt_params p = params_;
        p.x0 = b.x; p.y0 = b.y; p.w = b.width; p.h = b.height;
        int rc = rzk_render_guided(ctx_, guide_, &p, image_.data(), nullptr);
        if (rc != RZK_OK) throw std::runtime_error(std::string("rzk_render_guided: ") + rzk_ctx_last_error(ctx_));
        frame.resize(image_.size());
        for (int y = b.y; y < b.y + b.height; ++y)
            for (int x = b.x; x < b.x + b.width; ++x)
                for (int c = 0; c < 3; ++c) {
                    size_t i = ((size_t)y * width_ + x) * 3 + c;
                    frame[i] = image_[i];
                }
    }

private:
    rzk_ctx* ctx_;
    const rzk_guide* guide_;
    int width_, height_;
    rzk_pt_params params_{};
    std::vector<float> image_;
};

}  // namespace Renzoku
