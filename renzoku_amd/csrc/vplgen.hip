// vplgen.hip — bounce-1 VPL generation on the device: VplGenerator::spawn with
// max_bounce = 2 (vpl_generator.cc:127-181) over area lights.
//
// Light sample i (SubpathGenerator::sample_direct_vertex, path.cc:171-220 and
// 305-334): a light by Scene::sample_light (its DiscretePdf, strict <), a
// triangle of it by cumulative area ratio and a uniform point on it
// (AreaLight::sample -> surface sample, quad.cc:75-87, triangle.cc:298-308), a
// cosine-weighted direction in the light surface's basis
// (area_light.cc:26-44: init_from_nu(shape normal, shape tangent)), then the
// closest hit on the reference tree.  A hit on a surface with a material that
// is not a light becomes a VPL: receiver geometry, wi = -ray direction,
// power = Le |n_q . wo| / (pdf_light * (pdf_q * pdf_wo)).  Samples are taken
// in order until the requested number of VPLs is stored; the number of light
// samples used is the num_emitted_photons of the gather.
//
// Random numbers: 8 per sample from Philox4x32-10 keyed by the seed, counter =
// sample index, so every sample is independent of batching; the reference
// draws the same 7 numbers (2 for the light choice pair, 1 triangle, 2 point,
// 2 direction) from its global MT stream.  cos/sin of the direction angle come
// from sincos_turn (a few ulp from libm, the same bits on host and device).
//
// Per batch of samples: k_light_vpls traces them (records + valid flags),
// k_vpl_order gives each valid sample its place in sample order (one block
// scanning the flags), k_vpl_scatter writes the first ones still needed.
#pragma clang fp contract(off)

#include <algorithm>
#include <new>
#include <string>
#include <vector>

#include "rzk_device.h"

namespace rzk {

constexpr int kVplBlock = 256;
constexpr int kVplScanBlock = 1024;
constexpr float kVplTmax = 1e6f;        // common.h:48-50 (max_tmax)

template <typename F>
int vpl_guard(rzk_ctx* ctx, F&& f) {
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

// kStats: add the visited interior nodes and tested triangles to d_work[0..1]
template <bool kStats>
__global__ __launch_bounds__(kVplBlock) void k_light_vpls(SceneView sv, uint32_t seed, int64_t base, int n, float tmin,
                                                          float tmax, float* __restrict__ rec, int* __restrict__ rmat,
                                                          int* __restrict__ flag, unsigned long long* __restrict__ d_work) {
    extern __shared__ int s_stack[];
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    TravStats st;
    if (i < n) {
        int ok = 0;
        const VplRng rng = vpl_rng(seed, (uint64_t)(base + i));
        // Scene::sample_light: DiscretePdf over the lights (strict <)
        const float ul = u24(rng.u[0]);
        int k = -1;
        for (int q = 0; q < sv.nlights; ++q)
            if (ul < sv.lights[q].sel_cdf) { k = q; break; }
        if (k >= 0 && sv.lights[k].count > 0) {
            const LightGpu L = sv.lights[k];
            // triangle by cumulative area ratio (t <= ratio), uniform point on it
            const float ut = u24(rng.u[2]);
            int ti = L.first + L.count - 1;
            for (int q = 0; q < L.count; ++q)
                if (ut <= sv.ltris[L.first + q].cum) { ti = L.first + q; break; }
            const LightTri T = sv.ltris[ti];
            const float tmp = sqrtf(1.0f - u24(rng.u[3]));
            const float bb = 1.0f - tmp;
            const float cc = u24(rng.u[4]) * tmp;
            const float aa = 1.0f - bb - cc;
            const v3 q = (aa * mk(T.p0.x, T.p0.y, T.p0.z) + bb * mk(T.p1.x, T.p1.y, T.p1.z)) +
                         cc * mk(T.p2.x, T.p2.y, T.p2.z);
            const v3 qn = mk(T.n.x, T.n.y, T.n.z);
            const float pdf_q = T.ratio * (1.0f / T.n.w);
            // cosine-weighted direction in the surface basis (shape normal, shape tangent)
            float sp, cp;
            sincos_turn(u24(rng.u[5]), sp, cp);
            const float sin_t = sqrtf(u24(rng.u[6]));
            const float x = sin_t * cp, y = sin_t * sp;
            const float z = sqrtf(1.f - sin_t * sin_t);
            const float4 nf = sv.tri_n[T.tri], tf = sv.tri_tan[T.tri];
            const Onb b = onb_from_nu(mk(nf.x, nf.y, nf.z), mk(tf.x, tf.y, tf.z));
            const v3 wo = unit(local_to_world(b, mk(x, y, z)));
            const float pdf_wo = z * kInvPi;
            if (pdf_q > 0.0f && pdf_wo > 0.0f) {
                const float acc = L.sel_pdf * (pdf_q * pdf_wo);
                const float f = fabsf(dot(qn, wo));
                const v3 power = mk((L.le[0] * f) / acc, (L.le[1] * f) / acc, (L.le[2] * f) / acc);
                const Ray r = ray_ctor(q, wo);
                float t = 0.f, tb = 0.f, tg = 0.f;
                const int tri = traverse<false>(sv, r, tmin, tmax, s_stack + threadIdx.x, blockDim.x, t, tb, tg, st);
                if (tri >= 0) {
                    float rc[RZK_RECV_FLOATS];
                    int m = -1, light = -1;
                    fill_receiver(sv, r, t, tb, tg, tri, rc, m, light);
                    if (light < 0 && m >= 0) {
                        float* o = rec + (int64_t)RZK_VPL_FLOATS * i;
                        st3(o + 0, ld3(rc));             // p
                        st3(o + 3, ld3(rc + 6));         // shading normal
                        for (int e = 0; e < 9; ++e) o[6 + e] = rc[12 + e];   // uvn
                        st3(o + 15, -r.d);               // wi = -ray direction
                        st3(o + 18, power);
                        rmat[i] = m;
                        ok = 1;
                    }
                }
            }
        }
        flag[i] = ok;
    }
    if (kStats) add_work_totals(d_work, st);
}

// place of every valid sample among the batch's valid samples (exclusive scan
// of the 0/1 flags, in sample order); total at pos[n].  One block walks the
// flags in coalesced tiles of kVplScanBlock: ballot + popcount within a wave,
// the wave totals through LDS, a running carry across tiles.
__global__ __launch_bounds__(kVplScanBlock) void k_vpl_order(const int* __restrict__ flag, int n, int* __restrict__ pos) {
    constexpr int kWaves = kVplScanBlock / 64;
    __shared__ int s_w[2][kWaves];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    int carry = 0, buf = 0;
    for (int base = 0; base < n; base += kVplScanBlock, buf ^= 1) {
        const int i = base + t;
        const bool f = i < n && flag[i] != 0;
        const uint64_t b = __ballot(f);
        if (lane == 0) s_w[buf][w] = __popcll(b);
        __syncthreads();
        int off = 0, tot = 0;
        for (int k = 0; k < kWaves; ++k) {
            const int v = s_w[buf][k];
            off += k < w ? v : 0;
            tot += v;
        }
        if (i < n) pos[i] = carry + off + __popcll(b & below);
        carry += tot;
    }
    if (t == 0) pos[n] = carry;
}

// the valid samples with place < need go to out[have + place]; the sample
// holding place need - 1 records its index (the last sample used)
__global__ void k_vpl_scatter(const float* __restrict__ rec, const int* __restrict__ rmat, const int* __restrict__ flag,
                              const int* __restrict__ pos, int n, int64_t have, int64_t need, float* __restrict__ out,
                              int* __restrict__ out_mat, int* __restrict__ last) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n || !flag[i] || pos[i] >= need) return;
    const int64_t d = have + pos[i];
    for (int e = 0; e < RZK_VPL_FLOATS; ++e) out[d * RZK_VPL_FLOATS + e] = rec[(int64_t)i * RZK_VPL_FLOATS + e];
    out_mat[d] = rmat[i];
    if (pos[i] == need - 1) *last = i;
}

}  // namespace rzk

using namespace rzk;

extern "C" {

int rzk_generate_vpls(rzk_ctx* ctx, const rzk_scene* s, int64_t num_vpls, uint32_t seed, int64_t max_samples,
                      float tmin, float* out_vpl, int32_t* out_mat, int64_t* out_samples, int64_t* out_work) {
    if (!ctx) return RZK_E_INVALID;
    return vpl_guard(ctx, [&] {
        RZK_CHECK(s && out_vpl && out_mat && out_samples, RZK_E_INVALID, "null argument");
        RZK_CHECK(num_vpls > 0 && max_samples > 0, RZK_E_INVALID, "num_vpls and max_samples must be positive");
        RZK_CHECK(s->nlight > 0 && s->nltri > 0, RZK_E_INVALID, "the scene has no area light to emit from");
        SceneView sv = scene_view(s);
        hipStream_t st = ctx->stream;
        const int B = 1 << 18;                               // light samples per batch, at most
        // one piece of the context's scratch, cut 256-byte aligned:
        // records, materials, flags, places (B + 1) of a batch; the output; the last index; work totals
        size_t off = 0;
        auto cut = [&](size_t bytes) { const size_t o = off; off += (bytes + 255) & ~(size_t)255; return o; };
        const size_t o_rec = cut(sizeof(float) * RZK_VPL_FLOATS * (size_t)B), o_rmat = cut(sizeof(int) * (size_t)B),
                     o_flag = cut(sizeof(int) * (size_t)B), o_pos = cut(sizeof(int) * ((size_t)B + 1)),
                     o_out = cut(sizeof(float) * RZK_VPL_FLOATS * (size_t)num_vpls),
                     o_omat = cut(sizeof(int) * (size_t)num_vpls), o_last = cut(sizeof(int)),
                     o_work = cut(sizeof(unsigned long long) * 2);
        char* base = static_cast<char*>(ctx_scratch(ctx, off));
        float* rec = reinterpret_cast<float*>(base + o_rec);
        float* out = reinterpret_cast<float*>(base + o_out);
        int* rmat = reinterpret_cast<int*>(base + o_rmat);
        int* flag = reinterpret_cast<int*>(base + o_flag);
        int* pos = reinterpret_cast<int*>(base + o_pos);
        int* omat = reinterpret_cast<int*>(base + o_omat);
        int* last = reinterpret_cast<int*>(base + o_last);
        unsigned long long* work = reinterpret_cast<unsigned long long*>(base + o_work);
        if (out_work) RZK_HIP(hipMemsetAsync(work, 0, sizeof(unsigned long long) * 2, st));
        const size_t lds = sizeof(int) * kVplBlock * (size_t)sv.depth;
        int64_t have = 0, used = 0, traced = 0;
        while (have < num_vpls && used < max_samples) {
            // as many samples as the VPLs still needed take at the acceptance seen so far, plus 10%
            // (a sample past the last VPL needed is never used, so the batch size changes nothing)
            const double acc = used > 0 ? std::max(0.02, (double)have / (double)used) : 1.0;
            const int64_t want = (int64_t)((double)(num_vpls - have) / acc * 1.1) + 1024;
            const int n = (int)std::min<int64_t>(std::min<int64_t>(B, max_samples - used), want);
            const unsigned grid = (unsigned)((n + kVplBlock - 1) / kVplBlock);
            if (out_work)
                hipLaunchKernelGGL(k_light_vpls<true>, dim3(grid), dim3(kVplBlock), lds, st, sv, seed, used, n, tmin,
                                   kVplTmax, rec, rmat, flag, work);
            else
                hipLaunchKernelGGL(k_light_vpls<false>, dim3(grid), dim3(kVplBlock), lds, st, sv, seed, used, n, tmin,
                                   kVplTmax, rec, rmat, flag, nullptr);
            hipLaunchKernelGGL(k_vpl_order, dim3(1), dim3(kVplScanBlock), 0, st, flag, n, pos);
            const int64_t need = num_vpls - have;
            hipLaunchKernelGGL(k_vpl_scatter, dim3(grid), dim3(kVplBlock), 0, st, rec, rmat, flag, pos, n, have,
                               need, out, omat, last);
            RZK_HIP(hipGetLastError());
            traced += n;
            int total = 0;
            RZK_HIP(hipMemcpyAsync(&total, pos + n, sizeof(int), hipMemcpyDeviceToHost, st));
            RZK_HIP(hipStreamSynchronize(st));
            if (total >= need) {
                int li = 0;
                RZK_HIP(hipMemcpyAsync(&li, last, sizeof(int), hipMemcpyDeviceToHost, st));
                RZK_HIP(hipStreamSynchronize(st));
                used += (int64_t)li + 1;
                have = num_vpls;
            } else {
                used += n;
                have += total;
            }
        }
        RZK_CHECK(have == num_vpls, RZK_E_STATE,
                  "only " + std::to_string(have) + " VPLs after " + std::to_string(used) + " light samples");
        RZK_HIP(hipMemcpyAsync(out_vpl, out, sizeof(float) * RZK_VPL_FLOATS * (size_t)num_vpls,
                               hipMemcpyDeviceToHost, st));
        RZK_HIP(hipMemcpyAsync(out_mat, omat, sizeof(int) * (size_t)num_vpls, hipMemcpyDeviceToHost, st));
        if (out_work) {
            unsigned long long h[2];
            RZK_HIP(hipMemcpyAsync(h, work, sizeof(h), hipMemcpyDeviceToHost, st));
            RZK_HIP(hipStreamSynchronize(st));
            out_work[0] = traced;                            // closest rays traced (one per light sample traced,
                                                             // with the batch's samples past the last one used)
            out_work[1] = (int64_t)h[0];                     // interior nodes visited
            out_work[2] = (int64_t)h[1];                     // triangles tested
        }
        RZK_HIP(hipStreamSynchronize(st));
        *out_samples = used;
    });
}

}  // extern "C"
