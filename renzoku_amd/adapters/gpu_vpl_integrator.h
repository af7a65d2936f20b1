// gpu_vpl_integrator.h — GpuVplIntegrator: Renzoku's VirtualPointLight
// integrator seen through MonteCarloIntegrator (monte_carlo_integrator.h:28-54,
// vpl.h:24-33), backed by the MI355X engine.  integrate(bundle, frame)
// evaluates every primary ray of the bundle in one engine call, then runs
// MonteCarloIntegrator::integrate unchanged, whose loop keeps the reference's
// 2x2 box-filter accumulation into target / weight and the frame's current
// image (monte_carlo_integrator.cc:45-70); radiance(PrimaryRay&) is answered
// from that batch.  Two estimates, as VirtualPointLight::radiance(PrimaryRay&)
// selects them (vpl.cc:443-461):
//  * GATHER: at the ray's closest hit, the sum of gather_point over the VPL set
//    divided by the emitted light paths (vpl.cc:661-688, :488-489), through
//    rzk_trace_closest + rzk_fill_receivers + rzk_gather;
//  * GUIDED_PATH_TRACE: LightSlice-guided path tracing (vpl.cc:934-1041) of the
//    ray itself, through rzk_render_guided_rays over a guide (rzk_lightslice_guide
//    or rzk_guide_create).  A ray's random numbers are keyed by (seed, pixel,
//    sample, bounce); its sample number is frames_done x samples-per-pixel of the
//    bundle + its rank among its pixel's rays in the bundle, so every frame draws
//    new numbers.
#pragma once

#include <algorithm>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/rzk.h"
#include "aggregate_gpu.h"
#include "renzoku_iface.h"

namespace Renzoku {

class GpuVplIntegrator : public MonteCarloIntegrator {
public:
    enum Estimate { GATHER = 0, GUIDED_PATH_TRACE = 1 };

    // GATHER over a VPL set resident in the engine
    GpuVplIntegrator(const AggregateGpu* agg, const rzk_vpls* vpls, Float num_emitted, int clamp_mode,
                     Float clamp_threshold, Float tmin = 1e-3f)
        : estimate_(GATHER), ctx_(agg->context()), agg_(agg), vpls_(vpls), num_emitted_(num_emitted),
          clamp_mode_(clamp_mode), clamp_(clamp_threshold), tmin_(tmin) {}

    // GUIDED_PATH_TRACE with a guide's per-slice distributions
    GpuVplIntegrator(rzk_ctx* ctx, const rzk_guide* guide, int max_bounce, bool direct_lighting, uint32_t seed,
                     Float tmin = 1e-3f)
        : estimate_(GUIDED_PATH_TRACE), ctx_(ctx), guide_(guide), direct_(direct_lighting), seed_(seed),
          tmin_(tmin) {
        set_max_bounce(max_bounce);
    }

    void integrate(PrimaryRayBundle& b, FrameBuffer* frame) override {
        evaluate(b.rays);
        prefilled_ = true;
        MonteCarloIntegrator::integrate(b, frame);
        prefilled_ = false;
    }
    void on_frame_end() override { ++frames_done_; }

protected:
    void radiance(PrimaryRay& pr) override {
        if (prefilled_) return;                       // set by integrate's batch
        PrimaryRays one(1, pr);
        evaluate(one);
        pr.val = one[0].val;
    }

private:
    // every ray's radiance into its val, one engine call per estimate
    void evaluate(PrimaryRays& rays) {
        const size_t n = rays.size();
        if (n == 0) return;
        std::vector<float> rec(6 * n), rgb(3 * n);
        for (size_t i = 0; i < n; ++i) AggregateGpu::pack(rays[i].r, &rec[6 * i]);
        if (estimate_ == GATHER) {
            std::vector<rzk_hit> h(n);
            std::vector<float> recv(RZK_RECV_FLOATS * n);
            std::vector<int32_t> mat(n);
            check(rzk_trace_closest(ctx_, agg_->handle(), rec.data(), (int64_t)n, tmin_, RAY_TMAX, h.data()),
                  "rzk_trace_closest");
            check(rzk_fill_receivers(ctx_, agg_->handle(), rec.data(), h.data(), (int64_t)n, recv.data(), mat.data(),
                                     nullptr),
                  "rzk_fill_receivers");
            check(rzk_gather(ctx_, agg_->handle(), vpls_, recv.data(), mat.data(), (int64_t)n, clamp_mode_, clamp_,
                             tmin_, num_emitted_, rgb.data()),
                  "rzk_gather");
        } else {
            // samples per pixel of this bundle, and each ray's rank within its pixel
            std::map<std::pair<int, int>, int> seen;
            std::vector<int32_t> pix(2 * n), smp(n);
            int per_pixel = 1;
            for (size_t i = 0; i < n; ++i) per_pixel = std::max(per_pixel, ++seen[{rays[i].s.x, rays[i].s.y}]);
            seen.clear();
            for (size_t i = 0; i < n; ++i) {
                pix[2 * i] = rays[i].s.x;
                pix[2 * i + 1] = rays[i].s.y;
                smp[i] = frames_done_ * per_pixel + seen[{rays[i].s.x, rays[i].s.y}]++;
            }
            rzk_pt_params p{};
            p.spp = 1;
            p.max_bounce = get_max_bounce();
            p.direct_lighting = direct_ ? 1 : 0;
            p.guide_mode = 1;
            p.seed = seed_;
            p.tmin = tmin_;
            check(rzk_render_guided_rays(ctx_, guide_, &p, rec.data(), pix.data(), smp.data(), (int64_t)n, rgb.data()),
                  "rzk_render_guided_rays");
        }
        for (size_t i = 0; i < n; ++i) rays[i].val = Rgb(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]);
    }
    void check(int rc, const char* what) const {
        if (rc != RZK_OK) throw std::runtime_error(std::string(what) + ": " + rzk_ctx_last_error(ctx_));
    }

    Estimate estimate_;
    rzk_ctx* ctx_;
    const AggregateGpu* agg_ = nullptr;
    const rzk_vpls* vpls_ = nullptr;
    Float num_emitted_ = 1.0f;
    int clamp_mode_ = RZK_CLAMP_FORM_FACTOR;
    Float clamp_ = 0.01f;
    const rzk_guide* guide_ = nullptr;
    bool direct_ = false;
    uint32_t seed_ = 2013;
    Float tmin_;
    bool prefilled_ = false;
    int frames_done_ = 0;
};

}  // namespace Renzoku
