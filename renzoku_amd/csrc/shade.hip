// shade.hip — VPL gather kernels: pair visibility, gather shading, the fused
// gather used by whole-image renders, the LightSlice reduced matrix and the
// per-pixel cluster-representative gather.
#pragma clang fp contract(off)

#include "rzk_device.h"

namespace rzk {

constexpr int kVplTile = 64;         // VPL records staged in LDS per step
constexpr int kVplRec = 22;          // 21 floats + material (int bits)
constexpr int kShadeBlock = 128;

__device__ __forceinline__ void load_record(float* dst, const float* src, int nfl) {
    for (int i = 0; i < nfl; ++i) dst[i] = src[i];
}

// Cooperative load of VPL records [k0, k0 + cnt) into LDS.
__device__ __forceinline__ void stage_vpls(float* s_vpl, const float* __restrict__ vpl,
                                           const int* __restrict__ vmat, int64_t k0, int cnt) {
    for (int e = threadIdx.x; e < cnt * kVplRec; e += blockDim.x) {
        int k = e / kVplRec, f = e - k * kVplRec;
        s_vpl[e] = f < RZK_VPL_FLOATS ? vpl[(k0 + k) * RZK_VPL_FLOATS + f] : __int_as_float(vmat[k0 + k]);
    }
}

// ---- pair visibility: one wave = 64 receivers x one VPL ------------------
// bits[k * ngroups + g] bit l: segment (receiver 64g+l) -> VPL k blocked (or
// the receiver has no material / lies past the end)
__global__ __launch_bounds__(kShadeBlock) void k_visibility(SceneView sv, const float* __restrict__ recv,
                                                           const int* __restrict__ rmat, int64_t npts,
                                                           const float* __restrict__ vpl, int64_t nvpl,
                                                           float tmin, uint64_t* __restrict__ bits,
                                                           int* __restrict__ d_stats) {
    extern __shared__ int s_stack[];
    int64_t ngroups = (npts + 63) / 64;
    int64_t wave = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    int lane = threadIdx.x & 63;
    TravStats st;
    bool blocked = true;
    bool live = wave < ngroups * nvpl;
    int64_t k = 0, g = 0;
    if (live) {
        k = wave / ngroups;
        g = wave - k * ngroups;
        int64_t i = g * 64 + lane;
        if (i < npts && rmat[i] >= 0) {
            const float* rc = recv + RZK_RECV_FLOATS * i;
            const float* vp = vpl + RZK_VPL_FLOATS * k;
            blocked = segment_blocked(sv, mk(rc[0], rc[1], rc[2]), mk(vp[0], vp[1], vp[2]), tmin,
                                      s_stack + threadIdx.x, blockDim.x, st);
        }
    }
    uint64_t b = __ballot(blocked);
    if (live && lane == 0) bits[k * ngroups + g] = b;
    if (d_stats) add_work_totals(reinterpret_cast<unsigned long long*>(d_stats), st);
}

// ---- gather shading over all VPLs with precomputed visibility -------------
// Σ_k (visible ? gather_point : 0) in VPL order, then / num_emitted
__global__ __launch_bounds__(kShadeBlock) void k_gather_shade(const float* __restrict__ mat,
                                                             const float* __restrict__ recv,
                                                             const int* __restrict__ rmat, int64_t npts,
                                                             const float* __restrict__ vpl,
                                                             const int* __restrict__ vmat, int64_t nvpl,
                                                             const uint64_t* __restrict__ bits,
                                                             int mode, float thr, float n_emitted,
                                                             float* __restrict__ out) {
    __shared__ float s_vpl[kVplTile * kVplRec];
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t ngroups = (npts + 63) / 64;
    float rc[RZK_RECV_FLOATS];
    int m = -1;
    if (i < npts) {
        load_record(rc, recv + RZK_RECV_FLOATS * i, RZK_RECV_FLOATS);
        m = rmat[i];
    }
    v3 acc = mk(0.f, 0.f, 0.f);
    int64_t g = i / 64;
    int lane = (int)(i & 63);
    for (int64_t k0 = 0; k0 < nvpl; k0 += kVplTile) {
        int cnt = (int)((nvpl - k0) < kVplTile ? (nvpl - k0) : kVplTile);
        __syncthreads();
        stage_vpls(s_vpl, vpl, vmat, k0, cnt);
        __syncthreads();
        if (i < npts && m >= 0) {
            for (int kk = 0; kk < cnt; ++kk) {
                uint64_t w = bits[(k0 + kk) * ngroups + g];
                if ((w >> lane) & 1ull) continue;
                float vp[kVplRec];
                load_record(vp, s_vpl + kk * kVplRec, kVplRec);
                v3 lo = gather_shade(mat, rc, m, vp, __float_as_int(vp[21]), mode, thr);
                acc = acc + lo * 1.0f;
            }
        }
    }
    if (i < npts) {
        out[3 * i] = acc.x / n_emitted;
        out[3 * i + 1] = acc.y / n_emitted;
        out[3 * i + 2] = acc.z / n_emitted;
    }
}

// ---- fused gather: per receiver, for each VPL in order, trace the segment
// and shade when visible (one wave = 64 neighbouring receivers, so the 64
// segments of a step share an end point and stay coherent) -----------------
__global__ __launch_bounds__(kShadeBlock) void k_gather_fused(SceneView sv, const float* __restrict__ recv,
                                                             const int* __restrict__ rmat, int64_t npts,
                                                             const int* __restrict__ order,
                                                             const float* __restrict__ vpl,
                                                             const int* __restrict__ vmat, int64_t nvpl,
                                                             int mode, float thr, float tmin,
                                                             float n_emitted, float* __restrict__ out,
                                                             int* __restrict__ d_stats) {
    extern __shared__ int s_dyn[];
    float* s_vpl = reinterpret_cast<float*>(s_dyn);
    int* s_stack = s_dyn + kVplTile * kVplRec;
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t i = t < npts ? (order ? order[t] : t) : -1;
    float rc[RZK_RECV_FLOATS];
    int m = -1;
    if (i >= 0) {
        load_record(rc, recv + RZK_RECV_FLOATS * i, RZK_RECV_FLOATS);
        m = rmat[i];
    }
    TravStats st;
    v3 acc = mk(0.f, 0.f, 0.f);
    for (int64_t k0 = 0; k0 < nvpl; k0 += kVplTile) {
        int cnt = (int)((nvpl - k0) < kVplTile ? (nvpl - k0) : kVplTile);
        __syncthreads();
        stage_vpls(s_vpl, vpl, vmat, k0, cnt);
        __syncthreads();
        if (m >= 0) {
            for (int kk = 0; kk < cnt; ++kk) {
                float vp[kVplRec];
                load_record(vp, s_vpl + kk * kVplRec, kVplRec);
                if (segment_blocked(sv, mk(rc[0], rc[1], rc[2]), mk(vp[0], vp[1], vp[2]), tmin,
                                    s_stack + threadIdx.x, blockDim.x, st))
                    continue;
                v3 lo = gather_shade(sv.mat, rc, m, vp, __float_as_int(vp[21]), mode, thr);
                acc = acc + lo * 1.0f;
            }
        }
    }
    if (i >= 0) {
        out[3 * i] = acc.x / n_emitted;
        out[3 * i + 1] = acc.y / n_emitted;
        out[3 * i + 2] = acc.z / n_emitted;
    }
    if (d_stats) add_work_totals(reinterpret_cast<unsigned long long*>(d_stats), st);
}

// ---- LightSlice reduced matrix (lightslice.cc:721-748): pair (row r, VPL k),
// thread index = k * rows + r so the column-major outputs are written in order
__global__ __launch_bounds__(kShadeBlock) void k_reduced_matrix(SceneView sv, const float* __restrict__ recv,
                                                               const int* __restrict__ rmat, int rows,
                                                               const float* __restrict__ vpl,
                                                               const int* __restrict__ vmat, int64_t nvpl,
                                                               int mode, float thr, float tmin,
                                                               float* __restrict__ R, float* __restrict__ V,
                                                               int* __restrict__ d_stats) {
    extern __shared__ int s_stack[];
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    TravStats st;
    if (t < (int64_t)rows * nvpl) {
        int64_t k = t / rows;
        int r = (int)(t - k * rows);
        float rv = 0.f, g = 0.f, b = 0.f, vis = 0.f;
        if (rmat[r] >= 0) {
            float rc[RZK_RECV_FLOATS], vp[RZK_VPL_FLOATS];
            load_record(rc, recv + RZK_RECV_FLOATS * r, RZK_RECV_FLOATS);
            load_record(vp, vpl + RZK_VPL_FLOATS * k, RZK_VPL_FLOATS);
            bool visible = !segment_blocked(sv, ld3(rc), ld3(vp), tmin, s_stack + threadIdx.x, blockDim.x, st);
            v3 inc = incident_radiance(sv.mat, rc, vp, vmat[k], mode, thr);
            vis = (float)visible;
            rv = inc.x * vis; g = inc.y * vis; b = inc.z * vis;
        }
        R[3 * t] = rv; R[3 * t + 1] = g; R[3 * t + 2] = b;
        V[t] = vis;
    }
    if (d_stats) add_work_totals(reinterpret_cast<unsigned long long*>(d_stats), st);
}

// ---- LightSlice::gather_clusters for every pixel (lightslice.cc:1315-1345):
// Σ_c weight_c * gather_point(rep_c) over visible representatives, in cluster
// order, then / num_emitted.  Pixels are walked in 8x8 tiles per wave.
__global__ __launch_bounds__(kShadeBlock) void k_slice_gather(SceneView sv, const float* __restrict__ recv,
                                                             const int* __restrict__ rmat, int width,
                                                             int height, const float* __restrict__ vpl,
                                                             const int* __restrict__ vmat,
                                                             const int* __restrict__ reps,
                                                             const float* __restrict__ rep_w, int nreps,
                                                             int mode, float thr, float tmin,
                                                             float n_emitted, float* __restrict__ out,
                                                             int* __restrict__ d_stats) {
    extern __shared__ int s_dyn[];
    float* s_vpl = reinterpret_cast<float*>(s_dyn);           // kVplTile records + weight
    int* s_stack = s_dyn + kVplTile * (kVplRec + 1);
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int tiles_x = (width + 7) / 8;
    int64_t tile = t / 64;
    int lane = (int)(t & 63);
    int px = (int)(tile % tiles_x) * 8 + (lane & 7);
    int py = (int)(tile / tiles_x) * 8 + (lane >> 3);
    int64_t i = (px < width && py < height) ? (int64_t)py * width + px : -1;
    float rc[RZK_RECV_FLOATS];
    int m = -1;
    if (i >= 0) {
        load_record(rc, recv + RZK_RECV_FLOATS * i, RZK_RECV_FLOATS);
        m = rmat[i];
    }
    TravStats st;
    v3 acc = mk(0.f, 0.f, 0.f);
    for (int c0 = 0; c0 < nreps; c0 += kVplTile) {
        int cnt = (nreps - c0) < kVplTile ? (nreps - c0) : kVplTile;
        __syncthreads();
        for (int e = threadIdx.x; e < cnt * (kVplRec + 1); e += blockDim.x) {
            int c = e / (kVplRec + 1), f = e - c * (kVplRec + 1);
            int k = reps[c0 + c];
            float v;
            if (f < RZK_VPL_FLOATS) v = vpl[(int64_t)k * RZK_VPL_FLOATS + f];
            else if (f == RZK_VPL_FLOATS) v = __int_as_float(vmat[k]);
            else v = rep_w[c0 + c];
            s_vpl[e] = v;
        }
        __syncthreads();
        if (m >= 0) {
            for (int cc = 0; cc < cnt; ++cc) {
                float vp[kVplRec + 1];
                load_record(vp, s_vpl + cc * (kVplRec + 1), kVplRec + 1);
                if (segment_blocked(sv, mk(rc[0], rc[1], rc[2]), mk(vp[0], vp[1], vp[2]), tmin,
                                    s_stack + threadIdx.x, blockDim.x, st))
                    continue;
                v3 lo = gather_shade(sv.mat, rc, m, vp, __float_as_int(vp[21]), mode, thr);
                acc = acc + lo * vp[22];
            }
        }
    }
    if (i >= 0) {
        out[3 * i] = acc.x / n_emitted;
        out[3 * i + 1] = acc.y / n_emitted;
        out[3 * i + 2] = acc.z / n_emitted;
    }
    if (d_stats) add_work_totals(reinterpret_cast<unsigned long long*>(d_stats), st);
}

}  // namespace rzk

using namespace rzk;

static unsigned blocks_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

void launch_visibility(rzk_ctx* ctx, const SceneView& sv, const float* d_recv, const int* d_rmat,
                       int64_t npts, const float* d_vpl, int64_t nvpl, float tmin, uint64_t* d_bits,
                       int* d_stats) {
    int64_t waves = ((npts + 63) / 64) * nvpl;
    if (waves <= 0) return;
    size_t lds = sizeof(int) * kShadeBlock * (size_t)sv.depth;
    hipLaunchKernelGGL(k_visibility, dim3(blocks_for(waves, kShadeBlock / 64)), dim3(kShadeBlock), lds,
                       ctx->stream, sv, d_recv, d_rmat, npts, d_vpl, nvpl, tmin, d_bits, d_stats);
    RZK_HIP(hipGetLastError());
}

void launch_gather_shade(rzk_ctx* ctx, const SceneView& sv, const float* d_recv, const int* d_rmat,
                         int64_t npts, const float* d_vpl, const int* d_vmat, int64_t nvpl,
                         const uint64_t* d_bits, int clamp_mode, float clamp, float n_emitted,
                         float* d_out) {
    if (npts <= 0) return;
    hipLaunchKernelGGL(k_gather_shade, dim3(blocks_for(npts, kShadeBlock)), dim3(kShadeBlock), 0,
                       ctx->stream, sv.mat, d_recv, d_rmat, npts, d_vpl, d_vmat, nvpl, d_bits, clamp_mode,
                       clamp, n_emitted, d_out);
    RZK_HIP(hipGetLastError());
}

void launch_gather_fused(rzk_ctx* ctx, const SceneView& sv, const float* d_recv, const int* d_rmat,
                         int64_t npts, const int* d_order, const float* d_vpl, const int* d_vmat,
                         int64_t nvpl, int clamp_mode, float clamp, float tmin, float n_emitted,
                         float* d_out, int* d_stats) {
    if (npts <= 0) return;
    size_t lds = sizeof(float) * kVplTile * kVplRec + sizeof(int) * kShadeBlock * (size_t)sv.depth;
    hipLaunchKernelGGL(k_gather_fused, dim3(blocks_for(npts, kShadeBlock)), dim3(kShadeBlock), lds,
                       ctx->stream, sv, d_recv, d_rmat, npts, d_order, d_vpl, d_vmat, nvpl, clamp_mode,
                       clamp, tmin, n_emitted, d_out, d_stats);
    RZK_HIP(hipGetLastError());
}

void launch_reduced_matrix(rzk_ctx* ctx, const SceneView& sv, const float* d_recv, const int* d_rmat,
                           int rows, const float* d_vpl, const int* d_vmat, int64_t nvpl,
                           int clamp_mode, float clamp, float tmin, float* d_R, float* d_V,
                           int* d_stats) {
    int64_t n = (int64_t)rows * nvpl;
    if (n <= 0) return;
    size_t lds = sizeof(int) * kShadeBlock * (size_t)sv.depth;
    hipLaunchKernelGGL(k_reduced_matrix, dim3(blocks_for(n, kShadeBlock)), dim3(kShadeBlock), lds,
                       ctx->stream, sv, d_recv, d_rmat, rows, d_vpl, d_vmat, nvpl, clamp_mode, clamp,
                       tmin, d_R, d_V, d_stats);
    RZK_HIP(hipGetLastError());
}

void launch_slice_gather(rzk_ctx* ctx, const SceneView& sv, const float* d_recv, const int* d_rmat,
                         int width, int height, int tile, const float* d_vpl, const int* d_vmat,
                         const int* d_reps, const float* d_rep_w, int nreps, int clamp_mode,
                         float clamp, float tmin, float n_emitted, float* d_out, int* d_stats) {
    (void)tile;
    int64_t tiles = (int64_t)((width + 7) / 8) * ((height + 7) / 8);
    int64_t threads = tiles * 64;
    if (threads <= 0) return;
    size_t lds = sizeof(float) * kVplTile * (kVplRec + 1) + sizeof(int) * kShadeBlock * (size_t)sv.depth;
    hipLaunchKernelGGL(k_slice_gather, dim3(blocks_for(threads, kShadeBlock)), dim3(kShadeBlock), lds,
                       ctx->stream, sv, d_recv, d_rmat, width, height, d_vpl, d_vmat, d_reps, d_rep_w,
                       nreps, clamp_mode, clamp, tmin, n_emitted, d_out, d_stats);
    RZK_HIP(hipGetLastError());
}
