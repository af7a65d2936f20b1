// shade.hip — VPL gather kernels: pair visibility, gather shading, the fused
// gather used by whole-image renders, the LightSlice reduced matrix and the
// per-pixel cluster-representative gather.
#pragma clang fp contract(off)

#include <algorithm>

#include "rzk_device.h"

namespace rzk {

constexpr int kVplTile = 64;         // VPL records staged in LDS per step
constexpr int kVplRec = 22;          // 21 floats + material (int bits)
constexpr int kShadeBlock = 128;

__device__ __forceinline__ void load_record(float* dst, const float* src, int nfl) {
    for (int i = 0; i < nfl; ++i) dst[i] = src[i];
}

// This wave's any-hit stack in LDS (sv.stack4 entries per wave)
__device__ __forceinline__ int* wave_stack(int* base, const SceneView& sv) {
    return base + (threadIdx.x >> 6) * sv.stack4;
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
// the receiver has no material / lies past the end).  order (optional): the
// receiver at position t is order[t] (64 consecutive positions form a wave;
// an 8x8 pixel tile per wave keeps its 64 segments coherent).
__global__ __launch_bounds__(kShadeBlock) void k_visibility(SceneView sv, const float* __restrict__ recv,
                                                           const int* __restrict__ rmat, int64_t npts,
                                                           const int* __restrict__ order,
                                                           const float* __restrict__ vpl, int64_t nvpl,
                                                           float tmin, uint64_t* __restrict__ bits,
                                                           int* __restrict__ d_stats) {
    extern __shared__ int s_stack[];
    int* wstk = wave_stack(s_stack, sv);
    int64_t ngroups = (npts + 63) / 64;
    int64_t wave = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    int lane = threadIdx.x & 63;
    TravStats st;
    bool live = wave < ngroups * nvpl;
    int64_t k = 0, g = 0;
    bool active = false;
    v3 p = mk(0.f, 0.f, 0.f), q = mk(0.f, 0.f, 0.f);
    if (live) {
        k = wave / ngroups;
        g = wave - k * ngroups;
        int64_t t = g * 64 + lane;
        int64_t i = t < npts ? (order ? order[t] : t) : -1;
        if (i >= 0 && rmat[i] >= 0) {
            active = true;
            p = ld3(recv + RZK_RECV_FLOATS * i);
            q = ld3(vpl + RZK_VPL_FLOATS * k);
        }
    }
    bool blocked = segment_blocked_wave(sv, p, q, tmin, active, wstk, st) || !active;
    uint64_t b = __ballot(blocked);
    if (live && lane == 0) bits[k * ngroups + g] = b;
    if (d_stats) add_work_totals(reinterpret_cast<unsigned long long*>(d_stats), st);
}

// ---- gather shading over a VPL range with precomputed visibility ---------
// Σ_k (visible ? gather_point : 0) in VPL order, then / num_emitted.  The
// range may be one chunk of a longer VPL list: acc_in (optional) holds the
// receivers' running sums over the earlier chunks, and acc_out (optional)
// takes the running sums instead of the final radiance, so a list gathered in
// chunks sums in the same order, to the same bits, as in one launch.  order:
// as k_visibility (thread t shades receiver order[t]; bits of group t / 64).
__global__ __launch_bounds__(kShadeBlock) void k_gather_shade(const float* __restrict__ mat,
                                                             const float* __restrict__ recv,
                                                             const int* __restrict__ rmat, int64_t npts,
                                                             const int* __restrict__ order,
                                                             const float* __restrict__ vpl,
                                                             const int* __restrict__ vmat, int64_t nvpl,
                                                             const uint64_t* __restrict__ bits,
                                                             int mode, float thr, float n_emitted,
                                                             const float* __restrict__ acc_in,
                                                             float* __restrict__ acc_out,
                                                             float* __restrict__ out) {
    __shared__ float s_vpl[kVplTile * kVplRec];
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t ngroups = (npts + 63) / 64;
    int64_t i = t < npts ? (order ? order[t] : t) : -1;
    float rc[RZK_RECV_FLOATS];
    int m = -1;
    v3 acc = mk(0.f, 0.f, 0.f);
    if (i >= 0) {
        load_record(rc, recv + RZK_RECV_FLOATS * i, RZK_RECV_FLOATS);
        m = rmat[i];
        if (acc_in) acc = ld3(acc_in + 3 * i);
    }
    int64_t g = t / 64;
    int lane = (int)(t & 63);
    static_assert(kVplTile == 64, "one visibility word of the tile per lane");
    for (int64_t k0 = 0; k0 < nvpl; k0 += kVplTile) {
        int cnt = (int)((nvpl - k0) < kVplTile ? (nvpl - k0) : kVplTile);
        // this wave's visibility words of the tile's VPLs, word kk in lane kk,
        // all in one load; the loop below takes them with readlane
        const uint64_t wv = (lane < cnt && g < ngroups) ? bits[(k0 + lane) * ngroups + g] : ~0ull;
        const uint32_t wlo = (uint32_t)wv, whi = (uint32_t)(wv >> 32);
        __syncthreads();
        stage_vpls(s_vpl, vpl, vmat, k0, cnt);
        __syncthreads();
        if (i >= 0 && m >= 0) {
            for (int kk = 0; kk < cnt; ++kk) {
                const uint64_t w = (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)wlo, kk) |
                                   ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)whi, kk) << 32);
                if ((w >> lane) & 1ull) continue;
                float vp[kVplRec];
                load_record(vp, s_vpl + kk * kVplRec, kVplRec);
                v3 lo = gather_shade(mat, rc, m, vp, __float_as_int(vp[21]), mode, thr);
                acc = acc + lo * 1.0f;
            }
        }
    }
    if (i >= 0) {
        if (acc_out) {
            acc_out[3 * i] = acc.x;
            acc_out[3 * i + 1] = acc.y;
            acc_out[3 * i + 2] = acc.z;
        } else {
            out[3 * i] = acc.x / n_emitted;
            out[3 * i + 1] = acc.y / n_emitted;
            out[3 * i + 2] = acc.z / n_emitted;
        }
    }
}

// ---- LightSlice reduced matrix (lightslice.cc:721-748): pair (row r, VPL k).
// A wave holds one row and 64 VPLs that are consecutive in Morton order
// (k = vorder[kk]), so its 64 segments leave one receiver in a narrow cone.
// Outputs go to the column-major position of the original VPL index.
// kCount: traversal work added to d_stats (the counting pass); otherwise the
// counters are dead code.
constexpr int kRmBlock = 256;
constexpr int kRmXcd = 8;                      // XCDs; workgroup b runs on XCD b % 8
constexpr int kRmPair = 2;                     // blocks per row group, kRmXcd apart
constexpr int kRmRows = kRmPair * (kRmBlock / 64);   // rows sharing a VPL chunk

// kTerm 0: vis x incident radiance (the reduced matrix); 1: the light transport
// matrix's entries (LightTransportMatrix::fill, light_transport_matrix.cc:16-50,
// with VirtualPointLight::radiance as the evaluator, vpl.cc:405-409 ->
// gather_point :661-688): black when blocked, else the clamped outgoing
// radiance toward the receiver's wo; 2: the incident radiance without
// visibility (VirtualPointLight::incident_radiance, vpl.cc:411-440).  V is the
// visibility in every case.
template <bool kCount, int kTerm = 0>
__global__ __launch_bounds__(kRmBlock) void k_reduced_matrix(SceneView sv, const float* __restrict__ recv,
                                                               const int* __restrict__ rmat, int rows,
                                                               const int* __restrict__ vorder,
                                                               const float* __restrict__ vpl,
                                                               const int* __restrict__ vmat, int64_t nvpl,
                                                               int mode, float thr, float tmin,
                                                               float* __restrict__ R, float* __restrict__ V,
                                                               int* __restrict__ d_stats) {
    extern __shared__ int s_stack[];
    int* wstk = wave_stack(s_stack, sv);
    // a wave takes 64 consecutive (Morton-ordered) VPLs of one row.  A group
    // of kRmRows waves takes the same VPLs for kRmRows consecutive rows, so
    // their R / V writes for one VPL land side by side: the group is the
    // waves of blocks b and b + kRmXcd, which the dispatcher places on the
    // same XCD (workgroups go round-robin over the XCDs), so one L2 sees them
    const int64_t b = blockIdx.x;
    const int64_t x = b % kRmXcd, bx = b / kRmXcd;
    const int64_t G = (bx / kRmPair) * kRmXcd + x;               // (row group, chunk) pair
    const int gw = (int)((bx % kRmPair) * (kRmBlock / 64) + (threadIdx.x >> 6));
    const int lane = (int)(threadIdx.x & 63);
    const int64_t nc = (nvpl + 63) >> 6;                         // VPL chunks per row
    const int64_t rg = G / nc, chunk = G - rg * nc;
    const int r_ = (int)(rg * kRmRows + gw);
    const int64_t kk = chunk * 64 + lane;
    TravStats st;
    bool inside = r_ < rows && kk < nvpl;
    int r = 0;
    int64_t k = 0;
    bool active = false;
    v3 p = mk(0.f, 0.f, 0.f), q = mk(0.f, 0.f, 0.f);
    if (inside) {
        r = r_;
        k = vorder ? vorder[kk] : kk;
        active = rmat[r] >= 0;
        if (active) {
            p = ld3(recv + RZK_RECV_FLOATS * r);
            q = ld3(vpl + RZK_VPL_FLOATS * k);
        }
    }
    bool blocked = segment_blocked_wave<kCount>(sv, p, q, tmin, active, wstk, st);
    if (inside) {
        int64_t o = k * rows + r;
        float rv = 0.f, g = 0.f, b = 0.f, vis = 0.f;
        if (active) {
            float rc[RZK_RECV_FLOATS], vp[RZK_VPL_FLOATS];
            load_record(rc, recv + RZK_RECV_FLOATS * r, RZK_RECV_FLOATS);
            load_record(vp, vpl + RZK_VPL_FLOATS * k, RZK_VPL_FLOATS);
            vis = (float)(!blocked);
            if (kTerm == 1) {
                if (!blocked) {
                    const v3 lo = gather_shade(sv.mat, rc, rmat[r], vp, vmat[k], mode, thr);
                    rv = lo.x; g = lo.y; b = lo.z;
                }
            } else if (kTerm == 2) {
                const v3 inc = incident_radiance(sv.mat, rc, vp, vmat[k], mode, thr);
                rv = inc.x; g = inc.y; b = inc.z;
            } else {
                v3 inc = incident_radiance(sv.mat, rc, vp, vmat[k], mode, thr);
                rv = inc.x * vis; g = inc.y * vis; b = inc.z * vis;
            }
        }
        R[3 * o] = rv; R[3 * o + 1] = g; R[3 * o + 2] = b;
        V[o] = vis;
    }
    if (kCount) add_work_totals(reinterpret_cast<unsigned long long*>(d_stats), st);
}

// Pixel of lane `lane` of 8x8 tile `tile` (-1 outside the image)
__device__ __forceinline__ int64_t tile_pixel(int64_t tile, int lane, int width, int height) {
    int tiles_x = (width + 7) / 8;
    int px = (int)(tile % tiles_x) * 8 + (lane & 7);
    int py = (int)(tile / tiles_x) * 8 + (lane >> 3);
    return (px < width && py < height) ? (int64_t)py * width + px : -1;
}

// ---- LightSlice::gather_clusters, phase 1: visibility of every (pixel,
// representative) pair.  One wave = one 8x8 pixel tile and a share of the
// representatives (blockIdx.y); it records a 64-bit blocked mask per
// representative: bits[c * ntiles + tile], bit = lane.  Only the segment
// endpoints are live here, which keeps the traversal kernel small (high
// occupancy).  kShared: one representative list for every slice, staged in
// LDS per block; otherwise each lane reads its own slice's list (clusters
// refined per slice).  A pixel without a material, or a slice with fewer
// representatives than c, gets a set bit.
template <bool kCount, bool kShared>
__global__ __launch_bounds__(kShadeBlock) void k_slice_visibility(SceneView sv, const float* __restrict__ recv,
                                                                 const int* __restrict__ rmat, int width,
                                                                 int height, const float* __restrict__ vpl,
                                                                 SliceReps sr, float tmin,
                                                                 uint64_t* __restrict__ bits,
                                                                 int* __restrict__ d_stats, int groups) {
    extern __shared__ int s_dyn[];
    float* s_q = reinterpret_cast<float*>(s_dyn);            // kVplTile positions
    int* wstk = wave_stack(s_dyn + 3 * kVplTile, sv);
    // one flat grid: block B runs on XCD B % 8 (round-robin dispatch); the
    // `groups` representative shares of one tile block take consecutive
    // blocks of one XCD, so they run together over one L2 and share the
    // pixel-side nodes: B = ((tb / 8) * groups + g) * 8 + tb % 8
    const int64_t B = blockIdx.x, q = B / 8;
    const int gidx = (int)(q % groups);
    const int64_t tb = (q / groups) * 8 + (B % 8);
    int64_t t = tb * blockDim.x + threadIdx.x;
    int64_t ntiles = (int64_t)((width + 7) / 8) * ((height + 7) / 8);
    int64_t tile = t / 64;
    int lane = (int)(t & 63);
    int64_t i = tile < ntiles ? tile_pixel(tile, lane, width, height) : -1;
    int m = i >= 0 ? rmat[i] : -1;
    v3 p = m >= 0 ? ld3(recv + RZK_RECV_FLOATS * i) : mk(0.f, 0.f, 0.f);
    const int sl = (!kShared && m >= 0) ? sr.slice_of(i, width) : 0;
    const int mine = m >= 0 ? sr.count(sl) : 0;
    TravStats st;
    // gidx picks this block's share of the representatives
    const int C = sr.width();
    const int per = (C + groups - 1) / groups;
    const int c_begin = gidx * per, c_end = min(C, c_begin + per);
    for (int c0 = c_begin; c0 < c_end; c0 += kVplTile) {
        int cnt = (c_end - c0) < kVplTile ? (c_end - c0) : kVplTile;
        if (kShared) {
            __syncthreads();
            for (int e = threadIdx.x; e < cnt * 3; e += blockDim.x) {
                int c = e / 3, f = e - c * 3;
                s_q[e] = vpl[(int64_t)sr.reps[c0 + c] * RZK_VPL_FLOATS + f];
            }
            __syncthreads();
        }
        for (int cc = 0; cc < cnt; ++cc) {
            const bool active = m >= 0 && c0 + cc < mine;
            v3 q;
            if (kShared) q = mk(s_q[3 * cc], s_q[3 * cc + 1], s_q[3 * cc + 2]);
            else q = active ? ld3(vpl + (int64_t)sr.rep(sl, c0 + cc) * RZK_VPL_FLOATS) : mk(0.f, 0.f, 0.f);
            bool blocked = segment_blocked_wave<kCount>(sv, p, q, tmin, active, wstk, st) || !active;
            uint64_t b = __ballot(blocked);
            if (lane == 0 && tile < ntiles) bits[(int64_t)(c0 + cc) * ntiles + tile] = b;
        }
    }
    if (kCount) add_work_totals(reinterpret_cast<unsigned long long*>(d_stats), st);
}

// ---- phase 2: Σ_c weight_c * gather_point(rep_c) over the visible
// representatives, in cluster order, then / num_emitted (lightslice.cc:1325-1339,
// vpl.cc:488-489).  Same pixel-to-lane mapping as phase 1; kShared as there.
template <bool kShared>
__global__ __launch_bounds__(kShadeBlock) void k_slice_shade(const float* __restrict__ mat,
                                                            const float* __restrict__ recv,
                                                            const int* __restrict__ rmat, int width, int height,
                                                            const float* __restrict__ vpl,
                                                            const int* __restrict__ vmat, SliceReps sr,
                                                            const uint64_t* __restrict__ bits, int mode,
                                                            float thr, float n_emitted, float* __restrict__ out) {
    __shared__ float s_vpl[kVplTile * (kVplRec + 1)];
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t ntiles = (int64_t)((width + 7) / 8) * ((height + 7) / 8);
    int64_t tile = t / 64;
    int lane = (int)(t & 63);
    int64_t i = tile < ntiles ? tile_pixel(tile, lane, width, height) : -1;
    int m = i >= 0 ? rmat[i] : -1;
    float rc[RZK_RECV_FLOATS];
    if (m >= 0) load_record(rc, recv + RZK_RECV_FLOATS * i, RZK_RECV_FLOATS);
    v3 acc = mk(0.f, 0.f, 0.f);
    if (kShared) {
        const int nreps = sr.nreps;
        for (int c0 = 0; c0 < nreps; c0 += kVplTile) {
            int cnt = (nreps - c0) < kVplTile ? (nreps - c0) : kVplTile;
            // this wave's visibility words of the tile's representatives, word
            // cc in lane cc, all in one load; the loop below takes them with readlane
            const uint64_t wv = (lane < cnt && tile < ntiles) ? bits[(int64_t)(c0 + lane) * ntiles + tile] : ~0ull;
            const uint32_t wlo = (uint32_t)wv, whi = (uint32_t)(wv >> 32);
            __syncthreads();
            for (int e = threadIdx.x; e < cnt * (kVplRec + 1); e += blockDim.x) {
                int c = e / (kVplRec + 1), f = e - c * (kVplRec + 1);
                int k = sr.reps[c0 + c];
                float v;
                if (f < RZK_VPL_FLOATS) v = vpl[(int64_t)k * RZK_VPL_FLOATS + f];
                else if (f == RZK_VPL_FLOATS) v = __int_as_float(vmat[k]);
                else v = sr.rep_w[c0 + c];
                s_vpl[e] = v;
            }
            __syncthreads();
            if (m >= 0) {
                for (int cc = 0; cc < cnt; ++cc) {
                    const uint64_t w = (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)wlo, cc) |
                                       ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)whi, cc) << 32);
                    if ((w >> lane) & 1ull) continue;
                    float vp[kVplRec + 1];
                    load_record(vp, s_vpl + cc * (kVplRec + 1), kVplRec + 1);
                    v3 lo = gather_shade(mat, rc, m, vp, __float_as_int(vp[21]), mode, thr);
                    acc = acc + lo * vp[22];
                }
            }
        }
    } else if (m >= 0) {
        const int sl = sr.slice_of(i, width), n = sr.count(sl);
        for (int c = 0; c < n; ++c) {
            uint64_t w = bits[(int64_t)c * ntiles + tile];
            if ((w >> lane) & 1ull) continue;
            const int k = sr.rep(sl, c);
            float vp[RZK_VPL_FLOATS];
            load_record(vp, vpl + (int64_t)k * RZK_VPL_FLOATS, RZK_VPL_FLOATS);
            v3 lo = gather_shade(mat, rc, m, vp, vmat[k], mode, thr);
            acc = acc + lo * sr.weight(sl, c);
        }
    }
    if (i >= 0) {
        out[3 * i] = acc.x / n_emitted;
        out[3 * i + 1] = acc.y / n_emitted;
        out[3 * i + 2] = acc.z / n_emitted;
    }
}

}  // namespace rzk

using namespace rzk;

static unsigned blocks_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

void launch_visibility(rzk_ctx* ctx, const SceneView& sv_in, const float* d_recv, const int* d_rmat,
                       int64_t npts, const int* d_order, const float* d_vpl, int64_t nvpl, float tmin,
                       uint64_t* d_bits, int* d_stats) {
    int64_t waves = ((npts + 63) / 64) * nvpl;
    if (waves <= 0) return;
    SceneView sv = sv_in;
    unsigned grid = blocks_for(waves, kShadeBlock / 64);
    size_t lds = sizeof(int) * (kShadeBlock / 64) * (size_t)sv.stack4;
    hipLaunchKernelGGL(k_visibility, dim3(grid), dim3(kShadeBlock), lds,
                       ctx->stream, sv, d_recv, d_rmat, npts, d_order, d_vpl, nvpl, tmin, d_bits, d_stats);
    RZK_HIP(hipGetLastError());
}

void launch_gather_shade(rzk_ctx* ctx, const SceneView& sv, const float* d_recv, const int* d_rmat,
                         int64_t npts, const int* d_order, const float* d_vpl, const int* d_vmat, int64_t nvpl,
                         const uint64_t* d_bits, int clamp_mode, float clamp, float n_emitted,
                         const float* d_acc_in, float* d_acc_out, float* d_out) {
    if (npts <= 0) return;
    hipLaunchKernelGGL(k_gather_shade, dim3(blocks_for(npts, kShadeBlock)), dim3(kShadeBlock), 0,
                       ctx->stream, sv.mat, d_recv, d_rmat, npts, d_order, d_vpl, d_vmat, nvpl, d_bits, clamp_mode,
                       clamp, n_emitted, d_acc_in, d_acc_out, d_out);
    RZK_HIP(hipGetLastError());
}

void launch_reduced_matrix(rzk_ctx* ctx, const SceneView& sv_in, const float* d_recv, const int* d_rmat,
                           int rows, const int* d_vorder, const float* d_vpl, const int* d_vmat,
                           int64_t nvpl, int clamp_mode, float clamp, float tmin, float* d_R, float* d_V,
                           int* d_stats, int term) {
    if ((int64_t)rows * nvpl <= 0) return;
    SceneView sv = sv_in;
    // (row group, chunk) pairs, each kRmPair blocks kRmXcd apart (see the kernel);
    // the pairs are padded to a multiple of kRmXcd so every block has its partner
    const int64_t pairs = (((int64_t)rows + kRmRows - 1) / kRmRows) * ((nvpl + 63) / 64);
    const int64_t padded = (pairs + kRmXcd - 1) / kRmXcd * kRmXcd;
    unsigned grid = (unsigned)(padded * kRmPair);
    size_t lds = sizeof(int) * (kRmBlock / 64) * (size_t)sv.stack4;
    if (term == 1)
        hipLaunchKernelGGL((k_reduced_matrix<false, 1>), dim3(grid), dim3(kRmBlock), lds, ctx->stream, sv, d_recv,
                           d_rmat, rows, d_vorder, d_vpl, d_vmat, nvpl, clamp_mode, clamp, tmin, d_R, d_V, nullptr);
    else if (term == 2)
        hipLaunchKernelGGL((k_reduced_matrix<false, 2>), dim3(grid), dim3(kRmBlock), lds, ctx->stream, sv, d_recv,
                           d_rmat, rows, d_vorder, d_vpl, d_vmat, nvpl, clamp_mode, clamp, tmin, d_R, d_V, nullptr);
    else if (d_stats)
        hipLaunchKernelGGL(k_reduced_matrix<true>, dim3(grid), dim3(kRmBlock), lds, ctx->stream, sv, d_recv,
                           d_rmat, rows, d_vorder, d_vpl, d_vmat, nvpl, clamp_mode, clamp, tmin, d_R, d_V, d_stats);
    else
        hipLaunchKernelGGL(k_reduced_matrix<false>, dim3(grid), dim3(kRmBlock), lds, ctx->stream, sv, d_recv,
                           d_rmat, rows, d_vorder, d_vpl, d_vmat, nvpl, clamp_mode, clamp, tmin, d_R, d_V, d_stats);
    RZK_HIP(hipGetLastError());
}

void launch_slice_gather(rzk_ctx* ctx, const SceneView& sv_in, const float* d_recv, const int* d_rmat,
                         int width, int height, const float* d_vpl, const int* d_vmat, const SliceReps& sr,
                         int clamp_mode, float clamp, float tmin, float n_emitted, float* d_out, uint64_t* d_bits,
                         int* d_stats) {
    int64_t tiles = (int64_t)((width + 7) / 8) * ((height + 7) / 8);
    int64_t threads = tiles * 64;
    if (threads <= 0) return;
    SceneView sv = sv_in;
    unsigned grid = blocks_for(threads, kShadeBlock);
    size_t lds = sizeof(float) * 3 * kVplTile + sizeof(int) * (kShadeBlock / 64) * (size_t)sv.stack4;
    // a tile's representatives are split over `groups` waves of about four
    // representatives each (each wave writes its own bits): many short waves
    // keep every compute unit full and hide the node-fetch latency of the
    // wave-coherent walk (measured: one wave per tile 23.1 ms, four
    // representatives per wave 7.3 ms for 512x512 pixels x 300 representatives)
    const int groups = std::max(1, std::min((sr.width() + 3) / 4, 65535));
    const bool shared = sr.ld == 0;
    // flat grid of (tile block, share) pairs, the tile blocks padded to a multiple of 8 (see the kernel)
    const int64_t flat = ((int64_t)grid + 7) / 8 * 8 * groups;
    RZK_CHECK(flat < ((int64_t)1 << 31), RZK_E_INVALID, "image too large for one gather launch");
    dim3 g2((unsigned)flat);
    if (d_stats && shared)
        hipLaunchKernelGGL((k_slice_visibility<true, true>), g2, dim3(kShadeBlock), lds, ctx->stream, sv, d_recv, d_rmat,
                           width, height, d_vpl, sr, tmin, d_bits, d_stats, groups);
    else if (d_stats)
        hipLaunchKernelGGL((k_slice_visibility<true, false>), g2, dim3(kShadeBlock), lds, ctx->stream, sv, d_recv,
                           d_rmat, width, height, d_vpl, sr, tmin, d_bits, d_stats, groups);
    else if (shared)
        hipLaunchKernelGGL((k_slice_visibility<false, true>), g2, dim3(kShadeBlock), lds, ctx->stream, sv, d_recv,
                           d_rmat, width, height, d_vpl, sr, tmin, d_bits, d_stats, groups);
    else
        hipLaunchKernelGGL((k_slice_visibility<false, false>), g2, dim3(kShadeBlock), lds, ctx->stream, sv, d_recv,
                           d_rmat, width, height, d_vpl, sr, tmin, d_bits, d_stats, groups);
    RZK_HIP(hipGetLastError());
    if (shared)
        hipLaunchKernelGGL(k_slice_shade<true>, dim3(grid), dim3(kShadeBlock), 0, ctx->stream, sv.mat, d_recv, d_rmat,
                           width, height, d_vpl, d_vmat, sr, d_bits, clamp_mode, clamp, n_emitted, d_out);
    else
        hipLaunchKernelGGL(k_slice_shade<false>, dim3(grid), dim3(kShadeBlock), 0, ctx->stream, sv.mat, d_recv, d_rmat,
                           width, height, d_vpl, d_vmat, sr, d_bits, clamp_mode, clamp, n_emitted, d_out);
    RZK_HIP(hipGetLastError());
}
