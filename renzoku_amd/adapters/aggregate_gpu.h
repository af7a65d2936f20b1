// aggregate_gpu.h — AggregateGpu: Renzoku's Aggregate (aggregate.h:15-91)
// backed by the MI355X engine through the C ABI of include/rzk.h.  The three
// virtual hit overloads answer one ray each (the reference's call pattern);
// hit_batch / visible_batch answer many per call, which is how integrators
// should drive it.  A closest hit fills the whole HitRecord as
// Triangle::fill_hit_record + Surface::fill_hit_record do (triangle.cc:277-296,
// surface.cc:62-96): t, beta, gamma, geometric and shading normals (flipped per
// the surface flags), the triangle's tangent, the uvn frame, and the surface's
// material / light / env_light / shape / surface pointers from the binding table
// the caller passes (the reference's Surface objects, in surface order).  Texture
// coordinates (uv) are not read on the traced path and stay zero.  Errors from
// the engine become std::runtime_error, as Aggregate::sample throws
// (aggregate.cc:138-152).
#pragma once

#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/rzk.h"
#include "renzoku_iface.h"

namespace Renzoku {

// what Surface::fill_hit_record copies into a record, for one surface
struct SurfaceBinding {
    Surface* surface = nullptr;
    Material* material = nullptr;
    AreaLight* light = nullptr;
    EnvLight* env_light = nullptr;
    Shape* shape = nullptr;
};

class AggregateGpu : public Aggregate {
public:
    // triangles flattened in surface order (aggregate_bvh.cc:60-75), as rzk_scene_desc describes;
    // bindings: one per surface (desc.num_surfaces), or empty (pointers stay null);
    // scene_flags: RZK_SCENE_BUILD_DEVICE builds the trees on the GPU (rzk_scene_create_ex)
    AggregateGpu(rzk_ctx* ctx, const rzk_scene_desc& desc, std::vector<SurfaceBinding> bindings = {},
                 int scene_flags = 0)
        : ctx_(ctx), bindings_(std::move(bindings)) {
        if (!bindings_.empty() && (int)bindings_.size() != desc.num_surfaces)
            throw std::runtime_error("AggregateGpu: one surface binding per surface");
        tri_surface_.assign(desc.tri_surface, desc.tri_surface + desc.num_triangles);
        tangents_.assign(desc.tangents, desc.tangents + 3 * (size_t)desc.num_triangles);
        for (size_t i = 0; i < bindings_.size(); ++i)
            if (bindings_[i].material) material_ids_[bindings_[i].material] = desc.surf_material[i];
        check(rzk_scene_create_ex(ctx_, &desc, scene_flags, &scene_), "rzk_scene_create_ex");
    }
    ~AggregateGpu() override { rzk_scene_destroy(scene_); }
    AggregateGpu(const AggregateGpu&) = delete;
    AggregateGpu& operator=(const AggregateGpu&) = delete;

    // closest hit (aggregate_bvh.cc:402-448)
    bool hit(const Ray& r, Float tmin, Float tmax, Float time, HitRecord& rec) const override {
        (void)time;
        std::vector<Ray> one(1, r);
        std::vector<HitRecord> out;
        hit_batch(one, tmin, tmax, out);
        if (!out[0].hit) return false;
        rec = out[0];
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

    // batched closest hits: n rays per call, full records; triangles (optional)
    // gets each hit's triangle index in the flattened surface order (-1: miss)
    void hit_batch(const std::vector<Ray>& rays, Float tmin, Float tmax, std::vector<HitRecord>& out,
                   std::vector<int>* triangles = nullptr) const {
        const size_t n = rays.size();
        std::vector<float> rec(6 * n), recv(RZK_RECV_FLOATS * n);
        for (size_t i = 0; i < n; ++i) pack(rays[i], &rec[6 * i]);
        std::vector<rzk_hit> h(n);
        std::vector<int32_t> mat(n), light(n);
        out.assign(n, HitRecord());
        if (triangles) triangles->assign(n, -1);
        if (n == 0) return;
        check(rzk_trace_closest(ctx_, scene_, rec.data(), (int64_t)n, tmin, tmax, h.data()), "rzk_trace_closest");
        check(rzk_fill_receivers(ctx_, scene_, rec.data(), h.data(), (int64_t)n, recv.data(), mat.data(),
                                 light.data()),
              "rzk_fill_receivers");
        for (size_t i = 0; i < n; ++i) {
            if (h[i].tri < 0) continue;
            const float* rc = &recv[RZK_RECV_FLOATS * i];   // p, n, n_shading, wo, uvn.u, uvn.v, uvn.n
            HitRecord& o = out[i];
            o.hit = true;
            o.t = h[i].t;
            o.beta = h[i].beta;
            o.gamma = h[i].gamma;
            o.normal = Vec3(rc[3], rc[4], rc[5]);
            o.shading_normal = Vec3(rc[6], rc[7], rc[8]);
            o.uvn = Onb(Vec3(rc[12], rc[13], rc[14]), Vec3(rc[15], rc[16], rc[17]), Vec3(rc[18], rc[19], rc[20]));
            const float* tg = &tangents_[3 * (size_t)h[i].tri];
            o.tangent = Vec3(tg[0], tg[1], tg[2]);
            if (!bindings_.empty()) {
                const SurfaceBinding& s = bindings_[tri_surface_[h[i].tri]];
                o.material = s.material;
                o.light = s.light;
                o.env_light = s.env_light;
                o.shape = s.shape;
                o.surface = s.surface;
            }
            if (triangles) (*triangles)[i] = h[i].tri;
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
        if (p.empty()) return;
        check(rzk_trace_segments(ctx_, scene_, a.data(), b.data(), (int64_t)p.size(), tmin, blocked.data()),
              "rzk_trace_segments");
    }

    // the engine's material index of a bound surface's material (-1: none)
    int material_id(const Material* m) const {
        if (!m) return -1;
        auto it = material_ids_.find(m);
        if (it == material_ids_.end()) throw std::runtime_error("AggregateGpu: material of no bound surface");
        return it->second;
    }

    rzk_ctx* context() const { return ctx_; }
    rzk_scene* handle() const { return scene_; }

    // a ray record ({o, d} as Ray's constructor left them), the layout every trace entry point reads
    static void pack(const Ray& r, float* out) {
        Vec3 o = r.org(), d = r.dir();
        out[0] = o.x(); out[1] = o.y(); out[2] = o.z();
        out[3] = d.x(); out[4] = d.y(); out[5] = d.z();
    }

private:
    void check(int rc, const char* what) const {
        if (rc != RZK_OK) throw std::runtime_error(std::string(what) + ": " + rzk_ctx_last_error(ctx_));
    }
    rzk_ctx* ctx_;
    rzk_scene* scene_ = nullptr;
    std::vector<SurfaceBinding> bindings_;
    std::vector<int32_t> tri_surface_;
    std::vector<float> tangents_;
    std::map<const Material*, int> material_ids_;
};

}  // namespace Renzoku
