// traverse.hip — closest-hit, any-hit and segment traversal kernels, hit
// records and camera rays.
#pragma clang fp contract(off)

#include <algorithm>

#include "rzk_device.h"

namespace rzk {

constexpr int kTraceBlock = 128;

// kStats: 0 = none, 1 = per-ray counts (int2 per ray), 2 = launch totals
template <int kStats>
__device__ __forceinline__ void record_stats(int* d_stats, int64_t i, const TravStats& st) {
    if (kStats == 1) {
        d_stats[2 * i] = st.nodes;
        d_stats[2 * i + 1] = st.tris;
    } else if (kStats == 2) {
        add_work_totals(reinterpret_cast<unsigned long long*>(d_stats), st);
    }
}

template <int kStats>
__global__ __launch_bounds__(kTraceBlock) void k_trace_closest(SceneView sv, const float* __restrict__ rays,
                                                               int64_t n, float tmin, float tmax,
                                                               float4* __restrict__ hits,
                                                               int* __restrict__ d_stats) {
    extern __shared__ int s_stack[];
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    TravStats st;
    if (i < n) {
        const float* rp = rays + RZK_RAY_FLOATS * i;
        Ray r = ray_from_od(mk(rp[0], rp[1], rp[2]), mk(rp[3], rp[4], rp[5]));
        float t = 0.f, b = 0.f, g = 0.f;
        int tri = traverse<false>(sv, r, tmin, tmax, s_stack + threadIdx.x, blockDim.x, t, b, g, st);
        hits[i] = tri >= 0 ? make_float4(t, b, g, __int_as_float(tri))
                           : make_float4(0.f, 0.f, 0.f, __int_as_float(-1));
    }
    record_stats<kStats>(d_stats, i, st);
}

// 64-bit words of one wave's stack of the tiled closest-hit walk: depth lane
// masks and depth node refs.  One pending entry per tree level at most.
__host__ __device__ constexpr size_t tile_stack_u64(int depth) { return (size_t)depth + ((size_t)depth + 1) / 2; }

// Closest hits of a W x H image of rays (row-major), one wave per 8x8 pixel
// tile.  The wave walks the reference tree in the reference's depth-first order
// (aggregate_bvh.cc:402-448: both child boxes tested against the current tmax,
// right child first, left deferred) with one shared schedule: the node index is
// wave-uniform, so nodes and triangles are fetched once per wave through the
// scalar cache, and the wave stack holds (node, lanes) pairs.  A lane takes
// part only in the nodes its own walk would visit, with its own tmax, in the
// same order (the depth-first order of the whole tree restricted to the lane's
// visits is the lane's own order), so t, beta, gamma and the triangle id,
// ties included, are those of k_trace_closest.  LDS per wave: depth masks
// (uint64) followed by depth node refs (tile_stack_u64 words).
template <int kStats>
__global__ __launch_bounds__(kTraceBlock) void k_trace_closest_tiles(SceneView sv, const float* __restrict__ rays,
                                                                     int width, int height, float tmin,
                                                                     float tmax0, float4* __restrict__ hits,
                                                                     int* __restrict__ d_stats) {
    extern __shared__ unsigned long long s_wave[];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned long long* wmask = s_wave + (size_t)wave * tile_stack_u64(sv.depth);
    int* wnode = reinterpret_cast<int*>(wmask + sv.depth);
    const int tiles_x = (width + 7) / 8;
    const int64_t ntiles = (int64_t)tiles_x * ((height + 7) / 8);
    const int64_t tile = (int64_t)blockIdx.x * (kTraceBlock / 64) + wave;
    int64_t pix = -1;
    if (tile < ntiles) {
        int px = (int)(tile % tiles_x) * 8 + (lane & 7), py = (int)(tile / tiles_x) * 8 + (lane >> 3);
        if (px < width && py < height) pix = (int64_t)py * width + px;
    }
    TravStats st;
    Ray r = ray_from_od(mk(0.f, 0.f, 1.f), mk(0.f, 0.f, 1.f));
    if (pix >= 0) {
        const float* rp = rays + RZK_RAY_FLOATS * pix;
        r = ray_from_od(mk(rp[0], rp[1], rp[2]), mk(rp[3], rp[4], rp[5]));
    }
    float tmax = tmax0, tb = 0.f, bb = 0.f, gb = 0.f;
    int best = -1;
    bool live = pix >= 0 && box_hit(mk(sv.rbmin[0], sv.rbmin[1], sv.rbmin[2]),
                                    mk(sv.rbmax[0], sv.rbmax[1], sv.rbmax[2]), r, tmin, tmax);
    unsigned long long mask = __ballot(live);
    if (mask != 0ull) {
        int node = sv.root_ref;
        int sp = 0;
        for (;;) {
            node = __builtin_amdgcn_readfirstlane(node);
            const bool in = (mask >> lane) & 1ull;
            bool descend = false;
            if (node < 0) {
                const kconst float* tf = (const kconst float*)(sv.tris + (~node));
                v3 p0 = mk(tf[0], tf[1], tf[2]), e1 = mk(tf[4], tf[5], tf[6]), e2 = mk(tf[8], tf[9], tf[10]);
                int id = __float_as_int(tf[3]);
                if (in) {
                    st.tris++;
                    float t, b, g;
                    if (tri_hit(p0, e1, e2, r, tmin, tmax, t, b, g)) {
                        best = id;
                        tmax = t; tb = t; bb = b; gb = g;
                    }
                }
            } else {
                const kconst float* nf = (const kconst float*)(sv.nodes + node);
                const kconst int* ni = (const kconst int*)(sv.nodes + node);
                v3 llo = mk(nf[0], nf[1], nf[2]), lhi = mk(nf[3], nf[4], nf[5]);
                v3 rlo = mk(nf[6], nf[7], nf[8]), rhi = mk(nf[9], nf[10], nf[11]);
                int left = ni[12], right = ni[13];
                if (in) st.nodes++;
                bool hl = box_hit(llo, lhi, r, tmin, tmax) & in;     // idle lanes hold a finite ray
                bool hr = box_hit(rlo, rhi, r, tmin, tmax) & in;
                unsigned long long ml = __ballot(hl), mr = __ballot(hr);
                if (mr != 0ull) {
                    if (ml != 0ull) {
                        wmask[sp] = ml;
                        wnode[sp] = left;
                        ++sp;
                    }
                    node = right;
                    mask = mr;
                    descend = true;
                } else if (ml != 0ull) {
                    node = left;
                    mask = ml;
                    descend = true;
                }
            }
            if (!descend) {
                if (sp == 0) break;
                --sp;
                node = wnode[sp];
                mask = wmask[sp];
            }
        }
    }
    if (pix >= 0)
        hits[pix] = best >= 0 ? make_float4(tb, bb, gb, __int_as_float(best))
                              : make_float4(0.f, 0.f, 0.f, __int_as_float(-1));
    record_stats<kStats == 2 ? 2 : 0>(d_stats, pix, st);
}

template <int kStats>
__global__ __launch_bounds__(kTraceBlock) void k_trace_any(SceneView sv, const float* __restrict__ rays,
                                                           int64_t n, float tmin, float tmax,
                                                           uint8_t* __restrict__ out,
                                                           int* __restrict__ d_stats) {
    extern __shared__ int s_stack[];
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    TravStats st;
    if (i < n) {
        const float* rp = rays + RZK_RAY_FLOATS * i;
        Ray r = ray_from_od(mk(rp[0], rp[1], rp[2]), mk(rp[3], rp[4], rp[5]));
        if (kStats == 1) {
            // per-ray counts follow the reference's binary visit order
            float t, b, g;
            out[i] = traverse<true>(sv, r, tmin, tmax, s_stack + threadIdx.x, blockDim.x, t, b, g, st) >= 0;
        } else {
            AnyStack stk = any_stack(sv, s_stack);
            out[i] = any_hit4(sv, r, tmin, tmax, stk, st);
        }
    }
    record_stats<kStats>(d_stats, i, st);
}

template <int kStats>
__global__ __launch_bounds__(kTraceBlock) void k_trace_segments(SceneView sv, const float* __restrict__ p,
                                                                const float* __restrict__ q, int64_t n,
                                                                float tmin, uint8_t* __restrict__ out,
                                                                int* __restrict__ d_stats) {
    extern __shared__ int s_stack[];
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    TravStats st;
    if (i < n) {
        AnyStack stk = any_stack(sv, s_stack);
        out[i] = segment_blocked(sv, ld3(p + 3 * i), ld3(q + 3 * i), tmin, stk, st);
    }
    record_stats<kStats>(d_stats, i, st);
}

__global__ void k_fill_receivers(SceneView sv, const float* __restrict__ rays,
                                 const float4* __restrict__ hits, int64_t n, float* __restrict__ recv,
                                 int* __restrict__ rmat, int* __restrict__ rlight) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* rp = rays + RZK_RAY_FLOATS * i;
    Ray r = ray_from_od(mk(rp[0], rp[1], rp[2]), mk(rp[3], rp[4], rp[5]));
    float4 h = hits[i];
    float rec[RZK_RECV_FLOATS];
    int m, l;
    fill_receiver(sv, r, h.x, h.y, h.z, __float_as_int(h.w), rec, m, l);
    for (int k = 0; k < RZK_RECV_FLOATS; ++k) recv[RZK_RECV_FLOATS * i + k] = rec[k];
    rmat[i] = m;
    if (rlight) rlight[i] = l;
}

struct CamParams { float v[RZK_CAM_FLOATS]; };

// Camera::shoot_ray(i, j), camera.cc:52-65, followed by Ray::Ray (which
// normalises the already normalised direction once more)
__device__ __forceinline__ void camera_ray(const CamParams& c, int i, int j, float* out) {
    v3 eye = mk(c.v[0], c.v[1], c.v[2]);
    Onb b;
    b.u = mk(c.v[3], c.v[4], c.v[5]);
    b.v = mk(c.v[6], c.v[7], c.v[8]);
    b.n = mk(c.v[9], c.v[10], c.v[11]);
    float height = c.v[18];
    float pxf = (float)j + 0.5f;
    float pyf = (height - 1.0f - (float)i) + 0.5f;
    float fx = (pxf - c.v[12]) / c.v[14];
    float fy = (pyf - c.v[13]) / c.v[15];
    v3 d = unit(local_to_world(b, mk(fx, fy, -c.v[16])));
    Ray r = ray_ctor(eye, d);
    st3(out, r.o);
    st3(out + 3, r.d);
}

__global__ void k_camera_rays(CamParams c, int64_t n, const int* __restrict__ px,
                              const int* __restrict__ py, float* __restrict__ rays) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    float out[6];
    camera_ray(c, py[k], px[k], out);
    for (int t = 0; t < 6; ++t) rays[6 * k + t] = out[t];
}

__global__ void k_camera_image_rays(CamParams c, int width, int64_t n, float* __restrict__ rays) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    float out[6];
    camera_ray(c, (int)(k / width), (int)(k % width), out);
    for (int t = 0; t < 6; ++t) rays[6 * k + t] = out[t];
}

}  // namespace rzk

using namespace rzk;

static unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

void launch_trace_closest(rzk_ctx* ctx, const SceneView& sv, const float* d_rays, int64_t n,
                          float tmin, float tmax, float4* d_hits, int* d_stats) {
    if (n <= 0) return;
    size_t lds = sizeof(int) * kTraceBlock * (size_t)sv.depth;
    if (d_stats)
        hipLaunchKernelGGL(k_trace_closest<2>, dim3(grid_for(n, kTraceBlock)), dim3(kTraceBlock), lds,
                           ctx->stream, sv, d_rays, n, tmin, tmax, d_hits, d_stats);
    else
        hipLaunchKernelGGL(k_trace_closest<0>, dim3(grid_for(n, kTraceBlock)), dim3(kTraceBlock), lds,
                           ctx->stream, sv, d_rays, n, tmin, tmax, d_hits, d_stats);
    RZK_HIP(hipGetLastError());
}

void launch_trace_closest_image(rzk_ctx* ctx, const SceneView& sv, const float* d_rays, int width, int height,
                                float tmin, float tmax, float4* d_hits, int* d_stats) {
    const int64_t tiles = (int64_t)((width + 7) / 8) * ((height + 7) / 8);
    if (tiles <= 0) return;
    const int waves = kTraceBlock / 64;
    size_t lds = sizeof(unsigned long long) * waves * tile_stack_u64(sv.depth);
    unsigned grid = (unsigned)((tiles + waves - 1) / waves);
    if (d_stats)
        hipLaunchKernelGGL(k_trace_closest_tiles<2>, dim3(grid), dim3(kTraceBlock), lds, ctx->stream, sv, d_rays,
                           width, height, tmin, tmax, d_hits, d_stats);
    else
        hipLaunchKernelGGL(k_trace_closest_tiles<0>, dim3(grid), dim3(kTraceBlock), lds, ctx->stream, sv, d_rays,
                           width, height, tmin, tmax, d_hits, d_stats);
    RZK_HIP(hipGetLastError());
}

void launch_trace_any(rzk_ctx* ctx, const SceneView& sv_in, const float* d_rays, int64_t n,
                      float tmin, float tmax, uint8_t* d_out, int* d_stats) {
    if (n <= 0) return;
    SceneView sv = sv_in;
    unsigned grid = grid_for(n, kTraceBlock);
    bind_spill(ctx, sv, (int64_t)grid * kTraceBlock);
    size_t lds = sizeof(int) * kTraceBlock * (size_t)kLdsStack;
    if (d_stats)
        hipLaunchKernelGGL(k_trace_any<2>, dim3(grid), dim3(kTraceBlock), lds,
                           ctx->stream, sv, d_rays, n, tmin, tmax, d_out, d_stats);
    else
        hipLaunchKernelGGL(k_trace_any<0>, dim3(grid), dim3(kTraceBlock), lds,
                           ctx->stream, sv, d_rays, n, tmin, tmax, d_out, d_stats);
    RZK_HIP(hipGetLastError());
}

void launch_trace_segments(rzk_ctx* ctx, const SceneView& sv_in, const float* d_p, const float* d_q,
                           int64_t n, float tmin, uint8_t* d_out, int* d_stats) {
    if (n <= 0) return;
    SceneView sv = sv_in;
    unsigned grid = grid_for(n, kTraceBlock);
    bind_spill(ctx, sv, (int64_t)grid * kTraceBlock);
    size_t lds = sizeof(int) * kTraceBlock * (size_t)kLdsStack;
    if (d_stats)
        hipLaunchKernelGGL(k_trace_segments<2>, dim3(grid), dim3(kTraceBlock), lds,
                           ctx->stream, sv, d_p, d_q, n, tmin, d_out, d_stats);
    else
        hipLaunchKernelGGL(k_trace_segments<0>, dim3(grid), dim3(kTraceBlock), lds,
                           ctx->stream, sv, d_p, d_q, n, tmin, d_out, d_stats);
    RZK_HIP(hipGetLastError());
}

// per-ray statistics (kind 0 closest, 1 any)
void launch_trace_stats(rzk_ctx* ctx, const SceneView& sv, int kind, const float* d_rays, int64_t n,
                        float tmin, float tmax, void* d_scratch_out, int* d_stats) {
    if (n <= 0) return;
    size_t lds = sizeof(int) * kTraceBlock * (size_t)std::max(sv.depth, kLdsStack);
    if (kind == 0)
        hipLaunchKernelGGL(k_trace_closest<1>, dim3(grid_for(n, kTraceBlock)), dim3(kTraceBlock), lds,
                           ctx->stream, sv, d_rays, n, tmin, tmax, (float4*)d_scratch_out, d_stats);
    else
        hipLaunchKernelGGL(k_trace_any<1>, dim3(grid_for(n, kTraceBlock)), dim3(kTraceBlock), lds,
                           ctx->stream, sv, d_rays, n, tmin, tmax, (uint8_t*)d_scratch_out, d_stats);
    RZK_HIP(hipGetLastError());
}

void launch_fill_receivers(rzk_ctx* ctx, const SceneView& sv, const float* d_rays,
                           const float4* d_hits, int64_t n, float* d_recv, int* d_mat, int* d_light) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_receivers, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, sv, d_rays,
                       d_hits, n, d_recv, d_mat, d_light);
    RZK_HIP(hipGetLastError());
}

void launch_camera_rays(rzk_ctx* ctx, const float* cam, int64_t n, const int* d_px, const int* d_py,
                        int width, float* d_rays) {
    (void)width;
    if (n <= 0) return;
    CamParams c;
    for (int i = 0; i < RZK_CAM_FLOATS; ++i) c.v[i] = cam[i];
    hipLaunchKernelGGL(k_camera_rays, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, c, n, d_px,
                       d_py, d_rays);
    RZK_HIP(hipGetLastError());
}

void launch_camera_image_rays(rzk_ctx* ctx, const float* cam, int width, int height, float* d_rays) {
    int64_t n = (int64_t)width * height;
    if (n <= 0) return;
    CamParams c;
    for (int i = 0; i < RZK_CAM_FLOATS; ++i) c.v[i] = cam[i];
    hipLaunchKernelGGL(k_camera_image_rays, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, c, width,
                       n, d_rays);
    RZK_HIP(hipGetLastError());
}

