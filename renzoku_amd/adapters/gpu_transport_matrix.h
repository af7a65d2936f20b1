// gpu_transport_matrix.h — Renzoku's VPL evaluator (IVirtualPointLightEvaluator,
// vpl_interface.h:126-139, as VirtualPointLight implements it, vpl.cc:405-440)
// and light transport matrix (LightTransportMatrix, light_transport_matrix.h:57-80,
// fill :16-50) backed by the MI355X engine.
//
//  * GpuVplEvaluator answers the reference's per-pair calls, radiance (the
//    gather_point value and visibility) and incident_radiance, one engine call
//    each; callers with many pairs should go through GpuLightTransportMatrix.
//  * GpuLightTransportMatrix::fill(scene, evaluator, vpls, pixels) builds the
//    whole pixels x VPLs matrix in one pass: pixel-centre primary rays
//    (Camera::shoot_ray(pixel.y, pixel.x)), rows whose ray misses or lands on a
//    light left black, then evaluator->radiance of every (row, VPL) pair, with
//    the evaluator's clamp.  The evaluator must be a GpuVplEvaluator over the
//    same aggregate: the pairs are evaluated on the device, not through the
//    virtual per-pair call.
// Materials map to the engine's material table through the AggregateGpu's
// surface bindings; a VPL without a material is one on a light.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/rzk.h"
#include "aggregate_gpu.h"
#include "renzoku_iface.h"

namespace Renzoku {

inline void put3(float* d, const Vec3& v) { d[0] = v.x(); d[1] = v.y(); d[2] = v.z(); }

// the engine's VPL record (RZK_VPL_FLOATS: position, normal, frame u, v, n,
// incident direction, power) and material index
inline void pack_vpl(const AggregateGpu& agg, const BrdfPointLight& l, float* rec, int32_t& mat) {
    const LocalGeometry dg = l.get_dg();
    put3(rec, dg.p);
    put3(rec + 3, dg.n);
    put3(rec + 6, dg.uvn.u());
    put3(rec + 9, dg.uvn.v());
    put3(rec + 12, dg.uvn.n());
    put3(rec + 15, l.get_wi());
    const Rgb pw = l.power();
    rec[18] = pw.red(); rec[19] = pw.green(); rec[20] = pw.blue();
    mat = agg.material_id(l.get_material());
}

// the engine's receiver record (RZK_RECV_FLOATS: p, n, shading n, wo, shading
// u, v, n) and material index
inline void pack_receiver(const AggregateGpu& agg, const Receiver& r, float* rec, int32_t& mat) {
    put3(rec, r.p);
    put3(rec + 3, r.n);
    put3(rec + 6, r.shading_n);
    put3(rec + 9, r.wo);
    put3(rec + 12, r.shading_uvn.u());
    put3(rec + 15, r.shading_uvn.v());
    put3(rec + 18, r.shading_uvn.n());
    mat = agg.material_id(r.m);
}

class GpuVplEvaluator : public IVirtualPointLightEvaluator {
public:
    // clamp: gather_point's (clamp_threshold), incident_clamp: incident_radiance's
    GpuVplEvaluator(const AggregateGpu* agg, int clamp_mode, Float clamp, Float incident_clamp, Float tmin = 1e-3f)
        : agg_(agg), clamp_mode_(clamp_mode), clamp_(clamp), incident_clamp_(incident_clamp), tmin_(tmin) {}

    using IVirtualPointLightEvaluator::radiance;
    Rgb radiance(Scene* scene, const Receiver& r, const BrdfPointLight& light, bool& visible) override {
        (void)scene;
        float v = 0.f;
        Rgb out = pair(r, light, RZK_TERM_RADIANCE, clamp_, &v);
        visible = v != 0.f;
        return out;
    }
    Rgb incident_radiance(Scene* scene, const Receiver& r, const BrdfPointLight& light) override {
        (void)scene;
        return pair(r, light, RZK_TERM_INCIDENT, incident_clamp_, nullptr);
    }

    const AggregateGpu* aggregate() const { return agg_; }
    int clamp_mode() const { return clamp_mode_; }
    Float clamp() const { return clamp_; }
    Float tmin() const { return tmin_; }

    // every (receiver, VPL) pair of a resident VPL set: M column-major Rgb, V visibility
    void radiance_matrix(const std::vector<float>& recv, const std::vector<int32_t>& recv_mat, const rzk_vpls* vpls,
                         float* M, float* V) const {
        check(rzk_transport_matrix(agg_->context(), agg_->handle(), vpls, recv.data(), recv_mat.data(),
                                   (int32_t)recv_mat.size(), RZK_TERM_RADIANCE, clamp_mode_, clamp_, tmin_, M, V),
              "rzk_transport_matrix");
    }

private:
    Rgb pair(const Receiver& r, const BrdfPointLight& light, int term, Float thr, float* vis) const {
        float rec[RZK_RECV_FLOATS], vrec[RZK_VPL_FLOATS];
        int32_t rmat = -1, vmat = -1;
        pack_receiver(*agg_, r, rec, rmat);
        pack_vpl(*agg_, light, vrec, vmat);
        rzk_ctx* ctx = agg_->context();
        rzk_vpls* one = nullptr;
        check(rzk_vpls_create(ctx, vrec, &vmat, 1, &one), "rzk_vpls_create");
        float m[3] = {0.f, 0.f, 0.f}, v = 0.f;
        const int rc = rzk_transport_matrix(ctx, agg_->handle(), one, rec, &rmat, 1, term, clamp_mode_, thr, tmin_, m, &v);
        rzk_vpls_destroy(one);
        check(rc, "rzk_transport_matrix");
        if (vis) *vis = v;
        return Rgb(m[0], m[1], m[2]);
    }
    void check(int rc, const char* what) const {
        if (rc != RZK_OK) throw std::runtime_error(std::string(what) + ": " + rzk_ctx_last_error(agg_->context()));
    }

    const AggregateGpu* agg_;
    int clamp_mode_;
    Float clamp_, incident_clamp_, tmin_;
};

class GpuLightTransportMatrix : public LightTransportMatrix {
public:
    // camera: the engine's camera record (RZK_CAM_FLOATS) of the scene's camera
    explicit GpuLightTransportMatrix(const float* camera) : cam_(camera, camera + RZK_CAM_FLOATS) {}

    void fill(Scene* scene, IVirtualPointLightEvaluator* renderer, BrdfPointLights& vpls, Pixels& pixels) override {
        (void)scene;
        auto* ev = dynamic_cast<GpuVplEvaluator*>(renderer);
        if (!ev) throw std::runtime_error("GpuLightTransportMatrix::fill needs a GpuVplEvaluator");
        const AggregateGpu& agg = *ev->aggregate();
        rzk_ctx* ctx = agg.context();
        const size_t rows = pixels.size(), cols = vpls.size();
        m.resize((int)rows, (int)cols);
        if (rows == 0 || cols == 0) return;
        // primary hits at the pixel centres; rows that miss or land on a light stay black
        std::vector<int32_t> px(rows), py(rows), mat(rows), light(rows);
        for (size_t i = 0; i < rows; ++i) {
            px[i] = (int32_t)pixels[i].x();
            py[i] = (int32_t)pixels[i].y();
        }
        std::vector<rzk_hit> hits(rows);
        std::vector<float> recv(RZK_RECV_FLOATS * rows);
        check(ctx, rzk_camera_primary(ctx, agg.handle(), cam_.data(), (int64_t)rows, px.data(), py.data(), nullptr,
                                      hits.data(), recv.data(), mat.data(), light.data()),
              "rzk_camera_primary");
        for (size_t i = 0; i < rows; ++i)
            if (hits[i].tri < 0 || light[i] >= 0) mat[i] = -1;
        std::vector<float> vrec(RZK_VPL_FLOATS * cols);
        std::vector<int32_t> vmat(cols);
        for (size_t k = 0; k < cols; ++k) pack_vpl(agg, vpls[k], &vrec[RZK_VPL_FLOATS * k], vmat[k]);
        rzk_vpls* set = nullptr;
        check(ctx, rzk_vpls_create(ctx, vrec.data(), vmat.data(), (int64_t)cols, &set), "rzk_vpls_create");
        std::vector<float> M(3 * rows * cols), V(rows * cols);
        try {
            ev->radiance_matrix(recv, mat, set, M.data(), V.data());
        } catch (...) {
            rzk_vpls_destroy(set);
            throw;
        }
        rzk_vpls_destroy(set);
        for (size_t k = 0; k < cols; ++k)
            for (size_t i = 0; i < rows; ++i) {
                const float* e = &M[3 * (k * rows + i)];
                m((int)i, (int)k) = Rgb(e[0], e[1], e[2]);
            }
    }

private:
    static void check(rzk_ctx* ctx, int rc, const char* what) {
        if (rc != RZK_OK) throw std::runtime_error(std::string(what) + ": " + rzk_ctx_last_error(ctx));
    }
    std::vector<float> cam_;
};

}  // namespace Renzoku
