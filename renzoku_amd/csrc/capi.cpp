// capi.cpp — the C ABI (include/rzk.h): contexts, scenes, VPL sets, traversal,
// receivers and the gather.  Every entry point converts exceptions into status
// codes and records the message on the context.
#pragma clang fp contract(off)

#include <algorithm>
#include <cstring>
#include <memory>

#include "rzk_internal.h"

using namespace rzk;

// launchers defined in traverse.hip / shade.hip / cluster.hip that are not in
// rzk_internal.h
void launch_trace_stats(rzk_ctx* ctx, const SceneView& sv, int kind, const float* d_rays, int64_t n,
                        float tmin, float tmax, void* d_scratch_out, int* d_stats);
void launch_alphas_full(rzk_ctx* ctx, int rows, int64_t cols, const float* d_R, float* d_alphas,
                        float* d_mags, float* d_norms, float* d_u, float* d_v, float* d_c0, float* d_mT);
void launch_assign_full(rzk_ctx* ctx, int rows, int64_t cols, const float* d_R, int ncl, const int* d_centers,
                        const float* d_weights, const float* d_mags, float* d_K, float* d_Knorm,
                        unsigned long long* d_key, int* d_assign);

// ---------------------------------------------------------------------------
// small RAII device buffer
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    void alloc(size_t count) {
        free();
        n = count;
        if (count) RZK_HIP(hipMalloc((void**)&p, sizeof(T) * count));
    }
    void free() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { free(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    void upload(rzk_ctx* ctx, const T* h, size_t count) {
        if (count) RZK_HIP(hipMemcpyAsync(p, h, sizeof(T) * count, hipMemcpyHostToDevice, ctx->stream));
    }
    void download(rzk_ctx* ctx, T* h, size_t count) const {
        if (count) RZK_HIP(hipMemcpyAsync(h, p, sizeof(T) * count, hipMemcpyDeviceToHost, ctx->stream));
    }
};

static void sync(rzk_ctx* ctx) { RZK_HIP(hipStreamSynchronize(ctx->stream)); }

template <typename F>
static int guarded(rzk_ctx* ctx, F&& f) {
    try {
        f();
        if (ctx) ctx->last_error.clear();
        return RZK_OK;
    } catch (const RzkError& e) {
        if (ctx) ctx->last_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        if (ctx) ctx->last_error = "host allocation failed";
        return RZK_E_OOM;
    } catch (const std::exception& e) {
        if (ctx) ctx->last_error = e.what();
        return RZK_E_INVALID;
    }
}

SceneView scene_view(const rzk_scene* s) {
    SceneView v;
    v.nodes = s->d_nodes;
    v.nodes4 = s->d_nodes4;
    v.tris = s->d_tris;
    v.tris4 = s->d_tris4;
    v.tri_n = s->d_tri_n;
    v.tri_tan = s->d_tri_tan;
    v.surf = s->d_surf;
    v.mat = s->d_mat;
    v.root_ref = s->bvh.root_ref;
    v.depth = s->bvh.depth > 0 ? s->bvh.depth : 1;
    v.root4_ref = s->bvh.root4_ref;
    v.ltris = s->d_ltris;
    v.lights = s->d_lights;
    v.nlights = (int)s->h_lights.size();
    v.stack4 = s->bvh.stack4;
    v.spill = nullptr;
    v.spill_pitch = 0;
    for (int i = 0; i < 3; ++i) { v.rbmin[i] = s->bvh.root_box[i]; v.rbmax[i] = s->bvh.root_box[3 + i]; }
    return v;
}

void* ctx_scratch(rzk_ctx* ctx, size_t bytes) {
    if (bytes > ctx->scratch_bytes) {
        if (ctx->scratch) RZK_HIP(hipFree(ctx->scratch));
        ctx->scratch = nullptr;
        ctx->scratch_bytes = 0;
        RZK_HIP(hipMalloc(&ctx->scratch, bytes));
        ctx->scratch_bytes = bytes;
    }
    return ctx->scratch;
}

void ctx_retain(rzk_ctx* ctx) { ctx->refs.fetch_add(1, std::memory_order_relaxed); }

void ctx_release(rzk_ctx* ctx) {
    if (!ctx || ctx->refs.fetch_sub(1, std::memory_order_acq_rel) != 1) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

void bind_spill(rzk_ctx* ctx, SceneView& sv, int64_t threads, int entries, int lds_entries) {
    int64_t extra = (entries > 0 ? entries : sv.stack4) - lds_entries;
    if (extra <= 0 || threads <= 0) { sv.spill = nullptr; sv.spill_pitch = 0; return; }
    sv.spill = static_cast<int*>(ctx_scratch(ctx, sizeof(int) * (size_t)extra * (size_t)threads));
    sv.spill_pitch = threads;
}

extern "C" {

// ---------------------------------------------------------------------------
int rzk_ctx_create(int device, rzk_ctx** out) {
    if (!out) return RZK_E_INVALID;
    *out = nullptr;
    std::unique_ptr<rzk_ctx> ctx(new rzk_ctx());
    ctx->device = device;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= device || device < 0) return RZK_E_HIP;
    if (hipSetDevice(device) != hipSuccess) return RZK_E_HIP;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) return RZK_E_HIP;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->num_cus = prop.multiProcessorCount;
    // the no-allocation gather's scratch address on this device (resolved once per device)
    try {
        gather_noalloc_setup(device);
    } catch (const RzkError&) {
        (void)hipStreamDestroy(ctx->stream);
        return RZK_E_HIP;
    }
    *out = ctx.release();
    return RZK_OK;
}

int rzk_ctx_destroy(rzk_ctx* ctx) {
    ctx_release(ctx);
    return RZK_OK;
}

const char* rzk_ctx_last_error(const rzk_ctx* ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

void* rzk_ctx_stream(rzk_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int rzk_ctx_synchronize(rzk_ctx* ctx) {
    return guarded(ctx, [&] { sync(ctx); });
}

int rzk_device_alloc(rzk_ctx* ctx, int64_t bytes, void** out) {
    return guarded(ctx, [&] {
        RZK_CHECK(out && bytes >= 0, RZK_E_INVALID, "bad allocation request");
        *out = nullptr;
        if (bytes > 0 && hipMalloc(out, (size_t)bytes) != hipSuccess) throw RzkError{RZK_E_OOM, "hipMalloc failed"};
    });
}

int rzk_device_free(rzk_ctx* ctx, void* ptr) {
    return guarded(ctx, [&] { if (ptr) RZK_HIP(hipFree(ptr)); });
}

int rzk_copy_h2d(rzk_ctx* ctx, void* dst, const void* src, int64_t bytes) {
    return guarded(ctx, [&] {
        if (bytes > 0) RZK_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, ctx->stream));
        sync(ctx);
    });
}

int rzk_copy_d2h(rzk_ctx* ctx, void* dst, const void* src, int64_t bytes) {
    return guarded(ctx, [&] {
        if (bytes > 0) RZK_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, ctx->stream));
        sync(ctx);
    });
}

// ---------------------------------------------------------------------------
int rzk_scene_create(rzk_ctx* ctx, const rzk_scene_desc* d, rzk_scene** out) {
    return rzk_scene_create_ex(ctx, d, 0, out);
}

int rzk_scene_create_ex(rzk_ctx* ctx, const rzk_scene_desc* d, int flags, rzk_scene** out) {
    return guarded(ctx, [&] {
        RZK_CHECK((flags & ~RZK_SCENE_BUILD_DEVICE) == 0, RZK_E_INVALID, "unknown scene flags");
        RZK_CHECK(ctx && d && out, RZK_E_INVALID, "null argument");
        *out = nullptr;
        RZK_CHECK(d->num_triangles >= 0 && d->num_surfaces >= 0 && d->num_materials >= 0, RZK_E_INVALID,
                  "negative counts");
        RZK_CHECK(d->num_triangles == 0 || (d->positions && d->tangents && d->tri_surface), RZK_E_INVALID,
                  "missing triangle arrays");
        for (int k = 0; k < d->num_triangles; ++k)
            RZK_CHECK(d->tri_surface[k] >= 0 && d->tri_surface[k] < d->num_surfaces, RZK_E_INVALID,
                      "triangle surface index out of range");
        for (int s = 0; s < d->num_surfaces; ++s) {
            RZK_CHECK(d->surf_material[s] < d->num_materials, RZK_E_INVALID, "surface material out of range");
            RZK_CHECK(d->surf_light[s] < d->num_lights, RZK_E_INVALID, "surface light out of range");
        }
        RZK_HIP(hipSetDevice(ctx->device));
        std::unique_ptr<rzk_scene> s(new rzk_scene());
        s->ctx = ctx;
        int n = d->num_triangles;
        if ((flags & RZK_SCENE_BUILD_DEVICE) && n >= 2)
            bvh_build_device(ctx, n, d->positions, d->tangents, d->tri_surface, s.get());
        else
            bvh_build_host(n, d->positions, s->bvh);
        s->nsurf = d->num_surfaces;
        s->nmat = d->num_materials;
        s->nlight = d->num_lights;
        s->h_tri_surface.assign(d->tri_surface, d->tri_surface + n);
        s->h_surf_material.assign(d->surf_material, d->surf_material + s->nsurf);
        s->h_surf_light.assign(d->surf_light, d->surf_light + s->nsurf);
        s->h_surf_flags.assign(d->surf_flags, d->surf_flags + s->nsurf);
        s->h_mat.assign(d->materials, d->materials + (size_t)RZK_MAT_FLOATS * s->nmat);
        if (d->num_lights > 0 && d->light_radiance)
            s->h_light_radiance.assign(d->light_radiance, d->light_radiance + 3 * d->num_lights);
        // per original triangle: geometric normal (triangle.cc:17-18) + surface id, tangent
        // (a device build computes them on the device, with the same arithmetic)
        const bool frames_on_host = !s->bvh.device_built;
        std::vector<float4> tn(frames_on_host && n > 0 ? n : 1), tt(frames_on_host && n > 0 ? n : 1);
        for (int k = 0; frames_on_host && k < n; ++k) {
            const float* p = d->positions + 9 * k;
            v3 p0 = ld3(p), p1 = ld3(p + 3), p2 = ld3(p + 6);
            v3 nn = unit(cross(p0 - p1, p0 - p2));
            int sid = d->tri_surface[k];
            float sf;
            std::memcpy(&sf, &sid, 4);
            tn[k] = make_float4(nn.x, nn.y, nn.z, sf);
            tt[k] = make_float4(d->tangents[3 * k], d->tangents[3 * k + 1], d->tangents[3 * k + 2], 0.f);
        }
        std::vector<int4> surf(s->nsurf > 0 ? s->nsurf : 1);
        for (int i = 0; i < s->nsurf; ++i)
            surf[i] = make_int4(d->surf_material[i], d->surf_light[i], d->surf_flags[i], 0);
        auto up = [&](void** dst, const void* src, size_t bytes) {
            if (bytes == 0) bytes = 16;
            RZK_HIP(hipMalloc(dst, bytes));
            if (src) RZK_HIP(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
        };
        if (!s->bvh.device_built) {
            up((void**)&s->d_nodes, s->bvh.nodes.data(), sizeof(BvhNodeGpu) * s->bvh.nodes.size());
            up((void**)&s->d_nodes4, s->bvh.nodes4.data(), sizeof(Bvh4NodeGpu) * s->bvh.nodes4.size());
            up((void**)&s->d_tris, s->bvh.tris.data(), sizeof(TriGpu) * s->bvh.tris.size());
            up((void**)&s->d_tris4, s->bvh.tris4.data(), sizeof(TriGpu) * s->bvh.tris4.size());
        }
        if (frames_on_host) {
            up((void**)&s->d_tri_n, tn.data(), sizeof(float4) * n);
            up((void**)&s->d_tri_tan, tt.data(), sizeof(float4) * n);
        }
        up((void**)&s->d_surf, surf.data(), sizeof(int4) * s->nsurf);
        up((void**)&s->d_mat, s->h_mat.empty() ? nullptr : s->h_mat.data(), sizeof(float) * s->h_mat.size());
        // light tables: emitting triangles grouped per light in surface order;
        // triangle choice by cumulative area ratio (quad.cc:75-87), light choice
        // by a DiscretePdf over radiance magnitude x area (Scene::sample_light)
        for (int li = 0; li < s->nlight; ++li) {
            LightGpu L;
            L.le[0] = s->h_light_radiance.empty() ? 0.f : s->h_light_radiance[3 * li];
            L.le[1] = s->h_light_radiance.empty() ? 0.f : s->h_light_radiance[3 * li + 1];
            L.le[2] = s->h_light_radiance.empty() ? 0.f : s->h_light_radiance[3 * li + 2];
            L.first = (int)s->h_ltris.size();
            float total = 0.f;
            std::vector<float> areas;
            for (int k = 0; k < n; ++k) {
                int sid = d->tri_surface[k];
                if (d->surf_light[sid] != li) continue;
                const float* p = d->positions + 9 * k;
                v3 p0 = ld3(p), p1 = ld3(p + 3), p2 = ld3(p + 6);
                v3 cr = cross(p0 - p1, p0 - p2);
                float area = length(cr) * 0.5f;                // triangle.cc:15
                v3 nn = unit(cr);
                if (d->surf_flags[sid] & RZK_SURF_FLIPPED) nn = -nn;   // surface.cc:98-102
                LightTri t;
                t.p0 = make_float4(p0.x, p0.y, p0.z, 0.f);
                t.p1 = make_float4(p1.x, p1.y, p1.z, 0.f);
                t.p2 = make_float4(p2.x, p2.y, p2.z, 0.f);
                t.n = make_float4(nn.x, nn.y, nn.z, area);
                t.light = li;
                t.tri = k;
                s->h_ltris.push_back(t);
                areas.push_back(area);
                total += area;
            }
            L.count = (int)s->h_ltris.size() - L.first;
            float cum = 0.f;
            for (int q = 0; q < L.count; ++q) {
                LightTri& t = s->h_ltris[L.first + q];
                t.ratio = areas[q] / total;
                cum += t.ratio;
                t.cum = q == L.count - 1 ? 1.0f : cum;
            }
            L.sel_cdf = total * rgb_value(mk(L.le[0], L.le[1], L.le[2]));   // weight, normalised below
            s->h_lights.push_back(L);
        }
        {
            float sum = 0.f;
            for (auto& L : s->h_lights) sum += L.sel_cdf;
            float cum = 0.f;
            for (size_t i = 0; i < s->h_lights.size(); ++i) {
                LightGpu& L = s->h_lights[i];
                L.sel_pdf = sum > 0.f ? L.sel_cdf / sum : 0.f;
                cum += L.sel_pdf;
                L.sel_cdf = i + 1 == s->h_lights.size() ? 1.0f : cum;
            }
        }
        s->nltri = (int)s->h_ltris.size();
        up((void**)&s->d_ltris, s->h_ltris.empty() ? nullptr : s->h_ltris.data(), sizeof(LightTri) * s->h_ltris.size());
        up((void**)&s->d_lights, s->h_lights.empty() ? nullptr : s->h_lights.data(), sizeof(LightGpu) * s->h_lights.size());
        ctx_retain(ctx);
        *out = s.release();
    });
}

int rzk_scene_destroy(rzk_scene* s) {
    if (!s) return RZK_OK;
    (void)hipSetDevice(s->ctx->device);
    void* ptrs[] = {s->d_nodes, s->d_nodes4, s->d_tris, s->d_tris4, s->d_tri_n, s->d_tri_tan, s->d_surf, s->d_mat,
                    s->d_ltris, s->d_lights};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    rzk_ctx* ctx = s->ctx;
    delete s;
    ctx_release(ctx);
    return RZK_OK;
}

int rzk_scene_get_info(const rzk_scene* s, rzk_scene_info* out) {
    if (!s || !out) return RZK_E_INVALID;
    out->num_triangles = s->bvh.ntri;
    out->num_interior_nodes = s->bvh.num_interior;
    out->depth = s->bvh.depth;
    for (int i = 0; i < 6; ++i) out->box[i] = s->bvh.box[i];
    out->device_bytes = (int64_t)sizeof(BvhNodeGpu) * s->bvh.num_interior +
                        (int64_t)sizeof(Bvh4NodeGpu) * (int64_t)s->bvh.num_nodes4 +
                        (int64_t)2 * sizeof(TriGpu) * s->bvh.ntri + (int64_t)2 * sizeof(float4) * s->bvh.ntri;
    return RZK_OK;
}

int rzk_scene_export_bvh(const rzk_scene* s, int32_t* sorted_tri, uint32_t* sorted_morton, int32_t* node_child,
                         int32_t* node_range, float* node_box, int32_t* num_nodes) {
    if (!s) return RZK_E_INVALID;
    if (s->bvh.device_built && (node_child || node_range || node_box)) return RZK_E_STATE;
    int nn = s->bvh.ntri > 0 ? 2 * s->bvh.ntri - 1 : 0;
    if (num_nodes) *num_nodes = nn;
    for (int i = 0; i < s->bvh.ntri; ++i) {
        if (sorted_tri) sorted_tri[i] = s->bvh.sorted_tri[i];
        if (sorted_morton) sorted_morton[i] = s->bvh.sorted_morton[i];
    }
    for (int j = 0; j < nn; ++j) {
        if (node_child) { node_child[2 * j] = s->bvh.bfs_child[2 * j]; node_child[2 * j + 1] = s->bvh.bfs_child[2 * j + 1]; }
        if (node_range) { node_range[2 * j] = s->bvh.bfs_range[2 * j]; node_range[2 * j + 1] = s->bvh.bfs_range[2 * j + 1]; }
        if (node_box) std::memcpy(node_box + 6 * j, &s->bvh.bfs_box[6 * j], 6 * sizeof(float));
    }
    return RZK_OK;
}

int rzk_scene_export_trees(const rzk_scene* s, void* nodes, void* tris, void* nodes4, void* tris4,
                           int32_t* num_nodes4) {
    if (!s) return RZK_E_INVALID;
    return guarded(s->ctx, [&] {
        if (num_nodes4) *num_nodes4 = s->bvh.num_nodes4;
        if (s->bvh.ntri == 0) return;
        auto get = [&](void* dst, const void* src, size_t bytes) {
            if (dst && bytes) RZK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s->ctx->stream));
        };
        get(nodes, s->d_nodes, sizeof(BvhNodeGpu) * (size_t)s->bvh.num_interior);
        get(tris, s->d_tris, sizeof(TriGpu) * (size_t)s->bvh.ntri);
        get(nodes4, s->d_nodes4, sizeof(Bvh4NodeGpu) * (size_t)s->bvh.num_nodes4);
        get(tris4, s->d_tris4, sizeof(TriGpu) * (size_t)s->bvh.ntri);
        RZK_HIP(hipStreamSynchronize(s->ctx->stream));
    });
}

// ---------------------------------------------------------------------------
int rzk_trace_closest_dev(rzk_ctx* ctx, const rzk_scene* s, const float* d_rays, int64_t n, float tmin,
                          float tmax, rzk_hit* d_out, uint64_t* d_stats) {
    return guarded(ctx, [&] {
        RZK_CHECK(s && n >= 0 && (n == 0 || (d_rays && d_out)), RZK_E_INVALID, "bad trace arguments");
        RZK_CHECK(s->bvh.ntri > 0 || n == 0, RZK_E_INVALID, "empty scene");
        launch_trace_closest(ctx, scene_view(s), d_rays, n, tmin, tmax, reinterpret_cast<float4*>(d_out), (int*)d_stats);
    });
}

int rzk_trace_any_dev(rzk_ctx* ctx, const rzk_scene* s, const float* d_rays, int64_t n, float tmin, float tmax,
                      uint8_t* d_out, uint64_t* d_stats) {
    return guarded(ctx, [&] {
        RZK_CHECK(s && n >= 0 && (n == 0 || (d_rays && d_out)), RZK_E_INVALID, "bad trace arguments");
        RZK_CHECK(s->bvh.ntri > 0 || n == 0, RZK_E_INVALID, "empty scene");
        launch_trace_any(ctx, scene_view(s), d_rays, n, tmin, tmax, d_out, (int*)d_stats);
    });
}

int rzk_trace_segments_dev(rzk_ctx* ctx, const rzk_scene* s, const float* d_p, const float* d_q, int64_t n,
                           float tmin, uint8_t* d_out, uint64_t* d_stats) {
    return guarded(ctx, [&] {
        RZK_CHECK(s && n >= 0 && (n == 0 || (d_p && d_q && d_out)), RZK_E_INVALID, "bad trace arguments");
        RZK_CHECK(s->bvh.ntri > 0 || n == 0, RZK_E_INVALID, "empty scene");
        launch_trace_segments(ctx, scene_view(s), d_p, d_q, n, tmin, d_out, (int*)d_stats);
    });
}

// host-buffer forms: an empty scene hits nothing
int rzk_trace_closest(rzk_ctx* ctx, const rzk_scene* s, const float* rays, int64_t n, float tmin, float tmax,
                      rzk_hit* out) {
    return guarded(ctx, [&] {
        RZK_CHECK(s && n >= 0 && (n == 0 || (rays && out)), RZK_E_INVALID, "bad trace arguments");
        if (n == 0) return;
        if (s->bvh.ntri == 0) {
            for (int64_t i = 0; i < n; ++i) out[i] = rzk_hit{0.f, 0.f, 0.f, -1};
            return;
        }
        DevBuf<float> dr(6 * n);
        DevBuf<float4> dh(n);
        dr.upload(ctx, rays, 6 * n);
        launch_trace_closest(ctx, scene_view(s), dr.p, n, tmin, tmax, dh.p, nullptr);
        dh.download(ctx, reinterpret_cast<float4*>(out), n);
        sync(ctx);
    });
}

int rzk_trace_any(rzk_ctx* ctx, const rzk_scene* s, const float* rays, int64_t n, float tmin, float tmax,
                  uint8_t* out) {
    return guarded(ctx, [&] {
        RZK_CHECK(s && n >= 0 && (n == 0 || (rays && out)), RZK_E_INVALID, "bad trace arguments");
        if (n == 0) return;
        if (s->bvh.ntri == 0) { std::memset(out, 0, n); return; }
        DevBuf<float> dr(6 * n);
        DevBuf<uint8_t> dh(n);
        dr.upload(ctx, rays, 6 * n);
        launch_trace_any(ctx, scene_view(s), dr.p, n, tmin, tmax, dh.p, nullptr);
        dh.download(ctx, out, n);
        sync(ctx);
    });
}

int rzk_trace_segments(rzk_ctx* ctx, const rzk_scene* s, const float* p, const float* q, int64_t n, float tmin,
                       uint8_t* out) {
    return guarded(ctx, [&] {
        RZK_CHECK(s && n >= 0 && (n == 0 || (p && q && out)), RZK_E_INVALID, "bad trace arguments");
        if (n == 0) return;
        if (s->bvh.ntri == 0) { std::memset(out, 0, n); return; }
        DevBuf<float> dp(3 * n), dq(3 * n);
        DevBuf<uint8_t> dh(n);
        dp.upload(ctx, p, 3 * n);
        dq.upload(ctx, q, 3 * n);
        launch_trace_segments(ctx, scene_view(s), dp.p, dq.p, n, tmin, dh.p, nullptr);
        dh.download(ctx, out, n);
        sync(ctx);
    });
}

int rzk_trace_stats(rzk_ctx* ctx, const rzk_scene* s, int kind, const float* rays, int64_t n, float tmin,
                    float tmax, int32_t* out_stats) {
    return guarded(ctx, [&] {
        RZK_CHECK(s && n >= 0 && (kind == 0 || kind == 1) && (n == 0 || (rays && out_stats)), RZK_E_INVALID,
                  "bad stats arguments");
        if (n == 0) return;
        if (s->bvh.ntri == 0) { std::memset(out_stats, 0, sizeof(int32_t) * 2 * n); return; }
        DevBuf<float> dr(6 * n);
        DevBuf<float4> dh(n);
        DevBuf<int> ds(2 * n);
        dr.upload(ctx, rays, 6 * n);
        launch_trace_stats(ctx, scene_view(s), kind, dr.p, n, tmin, tmax, dh.p, ds.p);
        ds.download(ctx, out_stats, 2 * n);
        sync(ctx);
    });
}

int rzk_fill_receivers(rzk_ctx* ctx, const rzk_scene* s, const float* rays, const rzk_hit* hits, int64_t n,
                       float* recv, int32_t* recv_mat, int32_t* recv_light) {
    return guarded(ctx, [&] {
        RZK_CHECK(s && n >= 0 && (n == 0 || (rays && hits && recv && recv_mat)), RZK_E_INVALID,
                  "bad receiver arguments");
        if (n == 0) return;
        DevBuf<float> dr(6 * n), dv(RZK_RECV_FLOATS * n);
        DevBuf<float4> dh(n);
        DevBuf<int> dm(n), dl(n);
        dr.upload(ctx, rays, 6 * n);
        dh.upload(ctx, reinterpret_cast<const float4*>(hits), n);
        launch_fill_receivers(ctx, scene_view(s), dr.p, dh.p, n, dv.p, dm.p, dl.p);
        dv.download(ctx, recv, RZK_RECV_FLOATS * n);
        dm.download(ctx, recv_mat, n);
        if (recv_light) dl.download(ctx, recv_light, n);
        sync(ctx);
    });
}

int rzk_camera_primary(rzk_ctx* ctx, const rzk_scene* s, const float* camera, int64_t n, const int32_t* px,
                       const int32_t* py, float* rays, rzk_hit* hits, float* recv, int32_t* recv_mat,
                       int32_t* recv_light) {
    return guarded(ctx, [&] {
        RZK_CHECK(s && camera && n >= 0 && (n == 0 || (px && py)), RZK_E_INVALID, "bad camera arguments");
        if (n == 0) return;
        DevBuf<int> dpx(n), dpy(n), dm(n), dl(n);
        DevBuf<float> dr(6 * n), dv(RZK_RECV_FLOATS * n);
        DevBuf<float4> dh(n);
        dpx.upload(ctx, px, n);
        dpy.upload(ctx, py, n);
        launch_camera_rays(ctx, camera, n, dpx.p, dpy.p, (int)camera[17], dr.p);
        if (s->bvh.ntri > 0) {
            launch_trace_closest(ctx, scene_view(s), dr.p, n, 1e-3f, 1e6f, dh.p, nullptr);
        } else {
            RZK_HIP(hipMemsetAsync(dh.p, 0xff, sizeof(float4) * n, ctx->stream));
        }
        launch_fill_receivers(ctx, scene_view(s), dr.p, dh.p, n, dv.p, dm.p, dl.p);
        if (rays) dr.download(ctx, rays, 6 * n);
        if (hits) dh.download(ctx, reinterpret_cast<float4*>(hits), n);
        if (recv) dv.download(ctx, recv, RZK_RECV_FLOATS * n);
        if (recv_mat) dm.download(ctx, recv_mat, n);
        if (recv_light) dl.download(ctx, recv_light, n);
        sync(ctx);
    });
}

// ---------------------------------------------------------------------------
int rzk_vpls_create(rzk_ctx* ctx, const float* vpl, const int32_t* vpl_mat, int64_t n, rzk_vpls** out) {
    return guarded(ctx, [&] {
        RZK_CHECK(out && n >= 0 && (n == 0 || (vpl && vpl_mat)), RZK_E_INVALID, "bad VPL arguments");
        *out = nullptr;
        std::unique_ptr<rzk_vpls> v(new rzk_vpls());
        v->ctx = ctx;
        v->n = n;
        v->h_vpl.assign(vpl, vpl + RZK_VPL_FLOATS * n);
        v->h_mat.assign(vpl_mat, vpl_mat + n);
        size_t nb = (size_t)(n > 0 ? n : 1);
        RZK_HIP(hipMalloc((void**)&v->d_vpl, sizeof(float) * RZK_VPL_FLOATS * nb));
        RZK_HIP(hipMalloc((void**)&v->d_mat, sizeof(int) * nb));
        RZK_HIP(hipMalloc((void**)&v->d_order, sizeof(int) * nb));
        if (n) {
            // Morton order of the VPL positions: kernels that send a wave of
            // segments to 64 VPLs at once take them in this order, so the
            // segments of a wave fan out into a narrow cone
            v3 lo = ld3(vpl), hi = ld3(vpl);
            for (int64_t i = 0; i < n; ++i) { lo = vmin(lo, ld3(vpl + RZK_VPL_FLOATS * i)); hi = vmax(hi, ld3(vpl + RZK_VPL_FLOATS * i)); }
            v3 ext = hi - lo;
            if (ext.x <= 0.f) ext.x = 1.f;
            if (ext.y <= 0.f) ext.y = 1.f;
            if (ext.z <= 0.f) ext.z = 1.f;
            std::vector<uint64_t> key(n);
            for (int64_t i = 0; i < n; ++i) {
                v3 c = (ld3(vpl + RZK_VPL_FLOATS * i) - lo) / ext;
                key[i] = ((uint64_t)morton3d(c.x, c.y, c.z) << 32) | (uint64_t)i;
            }
            std::sort(key.begin(), key.end());
            std::vector<int> order(n);
            for (int64_t i = 0; i < n; ++i) order[i] = (int)(key[i] & 0xffffffffu);
            RZK_HIP(hipMemcpy(v->d_vpl, vpl, sizeof(float) * RZK_VPL_FLOATS * n, hipMemcpyHostToDevice));
            RZK_HIP(hipMemcpy(v->d_mat, vpl_mat, sizeof(int) * n, hipMemcpyHostToDevice));
            RZK_HIP(hipMemcpy(v->d_order, order.data(), sizeof(int) * n, hipMemcpyHostToDevice));
        }
        ctx_retain(ctx);
        *out = v.release();
    });
}

int rzk_vpls_destroy(rzk_vpls* v) {
    if (!v) return RZK_OK;
    if (v->d_vpl) (void)hipFree(v->d_vpl);
    if (v->d_mat) (void)hipFree(v->d_mat);
    if (v->d_order) (void)hipFree(v->d_order);
    rzk_ctx* ctx = v->ctx;
    delete v;
    ctx_release(ctx);
    return RZK_OK;
}

// ---------------------------------------------------------------------------
int rzk_visibility(rzk_ctx* ctx, const rzk_scene* s, const rzk_vpls* v, const float* recv, const int32_t* recv_mat,
                   int64_t npts, float tmin, uint64_t* out_bits) {
    return guarded(ctx, [&] {
        RZK_CHECK(s && v && npts >= 0 && (npts == 0 || (recv && recv_mat && out_bits)), RZK_E_INVALID,
                  "bad visibility arguments");
        int64_t ngroups = (npts + 63) / 64;
        if (npts == 0 || v->n == 0) return;
        RZK_CHECK(s->bvh.ntri > 0, RZK_E_INVALID, "empty scene");
        DevBuf<float> dv(RZK_RECV_FLOATS * npts);
        DevBuf<int> dm(npts);
        DevBuf<uint64_t> db(ngroups * v->n);
        dv.upload(ctx, recv, RZK_RECV_FLOATS * npts);
        dm.upload(ctx, recv_mat, npts);
        launch_visibility(ctx, scene_view(s), dv.p, dm.p, npts, nullptr, v->d_vpl, v->n, tmin, db.p, nullptr);
        db.download(ctx, out_bits, ngroups * v->n);
        sync(ctx);
    });
}

int rzk_gather(rzk_ctx* ctx, const rzk_scene* s, const rzk_vpls* v, const float* recv, const int32_t* recv_mat,
               int64_t npts, int clamp_mode, float clamp, float tmin, float n_emitted, float* out_rgb) {
    return guarded(ctx, [&] {
        RZK_CHECK(s && v && npts >= 0 && (npts == 0 || (recv && recv_mat && out_rgb)), RZK_E_INVALID,
                  "bad gather arguments");
        RZK_CHECK(clamp_mode >= 0 && clamp_mode <= 3, RZK_E_INVALID, "bad clamp mode");
        if (npts == 0) return;
        RZK_CHECK(s->bvh.ntri > 0, RZK_E_INVALID, "empty scene");
        int64_t ngroups = (npts + 63) / 64;
        DevBuf<float> dv(RZK_RECV_FLOATS * npts), dout(3 * npts);
        DevBuf<int> dm(npts);
        DevBuf<uint64_t> db(ngroups * (v->n > 0 ? v->n : 1));
        dv.upload(ctx, recv, RZK_RECV_FLOATS * npts);
        dm.upload(ctx, recv_mat, npts);
        SceneView sv = scene_view(s);
        launch_visibility(ctx, sv, dv.p, dm.p, npts, nullptr, v->d_vpl, v->n, tmin, db.p, nullptr);
        launch_gather_shade(ctx, sv, dv.p, dm.p, npts, nullptr, v->d_vpl, v->d_mat, v->n, db.p, clamp_mode, clamp,
                            n_emitted, nullptr, nullptr, dout.p);
        dout.download(ctx, out_rgb, 3 * npts);
        sync(ctx);
    });
}

int rzk_reduced_matrix(rzk_ctx* ctx, const rzk_scene* s, const rzk_vpls* v, const float* recv,
                       const int32_t* recv_mat, int32_t rows, int clamp_mode, float clamp, float tmin, float* out_R,
                       float* out_V) {
    return guarded(ctx, [&] {
        RZK_CHECK(s && v && rows >= 0 && (rows == 0 || (recv && recv_mat && out_R && out_V)), RZK_E_INVALID,
                  "bad reduced matrix arguments");
        int64_t n = (int64_t)rows * v->n;
        if (n == 0) return;
        RZK_CHECK(s->bvh.ntri > 0, RZK_E_INVALID, "empty scene");
        DevBuf<float> dv(RZK_RECV_FLOATS * rows), dR(3 * n), dV(n);
        DevBuf<int> dm(rows);
        dv.upload(ctx, recv, RZK_RECV_FLOATS * rows);
        dm.upload(ctx, recv_mat, rows);
        launch_reduced_matrix(ctx, scene_view(s), dv.p, dm.p, rows, v->d_order, v->d_vpl, v->d_mat, v->n, clamp_mode, clamp, tmin,
                              dR.p, dV.p, nullptr);
        dR.download(ctx, out_R, 3 * n);
        dV.download(ctx, out_V, n);
        sync(ctx);
    });
}

int rzk_transport_matrix(rzk_ctx* ctx, const rzk_scene* s, const rzk_vpls* v, const float* recv,
                         const int32_t* recv_mat, int32_t rows, int term, int clamp_mode, float clamp, float tmin,
                         float* out_M, float* out_V) {
    return guarded(ctx, [&] {
        RZK_CHECK(s && v && rows >= 0 && (rows == 0 || (recv && recv_mat && out_M && out_V)), RZK_E_INVALID,
                  "bad transport matrix arguments");
        RZK_CHECK(clamp_mode >= RZK_CLAMP_NONE && clamp_mode <= RZK_CLAMP_RADIANCE, RZK_E_INVALID, "bad clamp mode");
        RZK_CHECK(term == RZK_TERM_RADIANCE || term == RZK_TERM_INCIDENT, RZK_E_INVALID, "bad transport term");
        int64_t n = (int64_t)rows * v->n;
        if (n == 0) return;
        RZK_CHECK(s->bvh.ntri > 0, RZK_E_INVALID, "empty scene");
        DevBuf<float> dv(RZK_RECV_FLOATS * rows), dM(3 * n), dV(n);
        DevBuf<int> dm(rows);
        dv.upload(ctx, recv, RZK_RECV_FLOATS * rows);
        dm.upload(ctx, recv_mat, rows);
        launch_reduced_matrix(ctx, scene_view(s), dv.p, dm.p, rows, v->d_order, v->d_vpl, v->d_mat, v->n, clamp_mode,
                              clamp, tmin, dM.p, dV.p, nullptr, term == RZK_TERM_RADIANCE ? 1 : 2);
        dM.download(ctx, out_M, 3 * n);
        dV.download(ctx, out_V, n);
        sync(ctx);
    });
}

int rzk_alphas(rzk_ctx* ctx, int32_t rows, int64_t cols, const float* R, float* out_alphas, float* out_mags) {
    return guarded(ctx, [&] {
        RZK_CHECK(rows > 0 && cols >= 0 && (cols == 0 || (R && out_alphas)), RZK_E_INVALID, "bad alpha arguments");
        if (cols == 0) return;
        DevBuf<float> dR(3 * (size_t)rows * cols), da(cols), dm(cols), dn(cols), du(rows), dvv(cols), dc(1),
            dmT((size_t)rows * cols);
        dR.upload(ctx, R, 3 * (size_t)rows * cols);
        launch_alphas_full(ctx, rows, cols, dR.p, da.p, dm.p, dn.p, du.p, dvv.p, dc.p, dmT.p);
        da.download(ctx, out_alphas, cols);
        if (out_mags) dm.download(ctx, out_mags, cols);
        sync(ctx);
    });
}

int rzk_cluster_assign(rzk_ctx* ctx, int32_t rows, int64_t cols, const float* R, int32_t ncl, const int32_t* centers,
                       const float* weights, int32_t* out_assign) {
    return guarded(ctx, [&] {
        RZK_CHECK(rows > 0 && cols >= 0 && ncl > 0 && (cols == 0 || (R && centers && weights && out_assign)),
                  RZK_E_INVALID, "bad assignment arguments");
        for (int j = 0; j < ncl; ++j)
            RZK_CHECK(centers[j] >= 0 && centers[j] < cols, RZK_E_INVALID, "centre out of range");
        if (cols == 0) return;
        size_t n3 = 3 * (size_t)rows;
        DevBuf<float> dR(n3 * cols), dw(ncl), dK(n3 * ncl), dKn(ncl), dn(cols), dm(cols), da(cols), du(rows),
            dvv(cols), dc(1), dmT((size_t)rows * cols);
        DevBuf<int> dc_(ncl), das(cols);
        DevBuf<unsigned long long> dkey(cols);
        dR.upload(ctx, R, n3 * cols);
        dw.upload(ctx, weights, ncl);
        dc_.upload(ctx, centers, ncl);
        launch_alphas_full(ctx, rows, cols, dR.p, da.p, dm.p, dn.p, du.p, dvv.p, dc.p, dmT.p);
        launch_assign_full(ctx, rows, cols, dR.p, ncl, dc_.p, dw.p, dm.p, dK.p, dKn.p, dkey.p, das.p);
        das.download(ctx, out_assign, cols);
        sync(ctx);
    });
}

}  // extern "C"
