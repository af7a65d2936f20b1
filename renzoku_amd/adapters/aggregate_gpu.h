// aggregate_gpu.h — AggregateGpu: Renzoku's Aggregate (aggregate.h:15-91)
// backed by the MI355X engine through the C ABI of include/rzk.h.  The three
// virtual hit overloads answer one ray each (the reference's call pattern);
// hit_batch / visible_batch answer many per call, which is how integrators
// should drive it.  Errors from the engine become std::runtime_error, as
// Aggregate::sample throws (aggregate.cc:138-152).
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/rzk.h"
#include "renzoku_iface.h"

namespace Renzoku {

class AggregateGpu : public Aggregate {
public:
    // triangles flattened in surface order (aggregate_bvh.cc:60-75), as rzk_scene_desc describes
    AggregateGpu(rzk_ctx* ctx, const rzk_scene_desc& desc) : ctx_(ctx) {
        check(rzk_scene_create(ctx_, &desc, &scene_), "rzk_scene_create");
    }
    ~AggregateGpu() override { rzk_scene_destroy(scene_); }
    AggregateGpu(const AggregateGpu&) = delete;
    AggregateGpu& operator=(const AggregateGpu&) = delete;

    // closest hit (aggregate_bvh.cc:402-448)
    bool hit(const Ray& r, Float tmin, Float tmax, Float time, HitRecord& rec) const override {
        (void)time;
        float ray[6];
        pack(r, ray);
        rzk_hit h;
        check(rzk_trace_closest(ctx_, scene_, ray, 1, tmin, tmax, &h), "rzk_trace_closest");
        if (h.tri < 0) return false;
        rec.hit = true;
        rec.t = h.t;
        rec.beta = h.beta;
        rec.gamma = h.gamma;
        rec.triangle = h.tri;
        return true;
    }
    // any hit (aggregate_bvh.cc:452-478)
    bool hit(const Ray& r, Float tmin, Float tmax, Float time) const override {
        (void)time;
        float ray[6];
        pack(r, ray);
        uint8_t any = 0;
        check(rzk_trace_any(ctx_, scene_, ray, 1, tmin, tmax, &any), "rzk_trace_any");
        return any != 0;
    }
    // segment visibility, true = blocked (aggregate.cc:84-90)
    bool hit(const Vec3& p, const Vec3& q, Float tmin, Float time) const override {
        (void)time;
        float a[3] = {p.x(), p.y(), p.z()}, b[3] = {q.x(), q.y(), q.z()};
        uint8_t blocked = 0;
        check(rzk_trace_segments(ctx_, scene_, a, b, 1, tmin, &blocked), "rzk_trace_segments");
        return blocked != 0;
    }

    // batched forms: n rays ({o, d} records) or n segments per call
    void hit_batch(const std::vector<Ray>& rays, Float tmin, Float tmax, std::vector<HitRecord>& out) const {
        std::vector<float> rec(6 * rays.size());
        for (size_t i = 0; i < rays.size(); ++i) pack(rays[i], &rec[6 * i]);
        std::vector<rzk_hit> h(rays.size());
        check(rzk_trace_closest(ctx_, scene_, rec.data(), (int64_t)rays.size(), tmin, tmax, h.data()),
              "rzk_trace_closest");
        out.assign(rays.size(), HitRecord());
        for (size_t i = 0; i < rays.size(); ++i) {
            if (h[i].tri < 0) continue;
            out[i].hit = true;
            out[i].t = h[i].t;
            out[i].beta = h[i].beta;
            out[i].gamma = h[i].gamma;
            out[i].triangle = h[i].tri;
        }
    }
    void visible_batch(const std::vector<Vec3>& p, const std::vector<Vec3>& q, Float tmin,
                       std::vector<uint8_t>& blocked) const {
        std::vector<float> a(3 * p.size()), b(3 * q.size());
        for (size_t i = 0; i < p.size(); ++i) {
            a[3 * i] = p[i].x(); a[3 * i + 1] = p[i].y(); a[3 * i + 2] = p[i].z();
            b[3 * i] = q[i].x(); b[3 * i + 1] = q[i].y(); b[3 * i + 2] = q[i].z();
        }
        blocked.assign(p.size(), 0);
        check(rzk_trace_segments(ctx_, scene_, a.data(), b.data(), (int64_t)p.size(), tmin, blocked.data()),
              "rzk_trace_segments");
    }

    rzk_scene* handle() const { return scene_; }

private:
    static void pack(const Ray& r, float* out) {
        Vec3 o = r.org(), d = r.dir();
        out[0] = o.x(); out[1] = o.y(); out[2] = o.z();
        out[3] = d.x(); out[4] = d.y(); out[5] = d.z();
    }
    void check(int rc, const char* what) const {
        if (rc != RZK_OK) throw std::runtime_error(std::string(what) + ": " + rzk_ctx_last_error(ctx_));
    }
    rzk_ctx* ctx_;
    rzk_scene* scene_ = nullptr;
};

}  // namespace Renzoku
