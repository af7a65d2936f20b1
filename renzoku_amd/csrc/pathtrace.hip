// pathtrace.hip — guided path tracing (VirtualPointLight::trace_wellington,
// vpl.cc:934-1041, METRO_RANDOM branch) with next-event estimation
// (FirstBounce::radiance, first_bounce.h:54-71 -> AreaLight::first_bounce,
// area_light.cc:48-142) and the LightSlice direction sampler at the first
// bounce (multi_slice_sample / sample_map / pdf_map, lightslice.cc:1569-2055,
// 2313-2377).
//
// One lane per pixel walks its samples in order; the recursion of
// trace_wellington is unrolled into a throughput product.  Every random
// number is Philox(pixel, sample, bounce, slot), so a pixel's value does not
// depend on which other pixels or requests share its launch.
#pragma clang fp contract(off)

#include <algorithm>

#include "pathtrace.h"
#include "rzk_device.h"

namespace rzk {

// AreaLight::query (area_light.cc:19-25 with AreaLight::pdf, area_light.h:120-131)
__device__ __forceinline__ v3 light_query(const LightGpu& L, v3 n, v3 wo) {
    if (dot(wo, n) < kZeroEps) return mk(0.f, 0.f, 0.f);
    return mk(L.le[0], L.le[1], L.le[2]);
}

__device__ __forceinline__ float balance(float a, float b) { return a / (a + b); }   // mis.h:11-13

// FirstBounce::radiance with no direction sampler: one light chosen by
// Scene::sample_light (DiscretePdf, strict <), then AreaLight::first_bounce:
// light sampling + BSDF sampling combined with the balance heuristic.
__device__ v3 next_event(const SceneView& sv, const float* rc, int m, const PathRng& rng, float tmin,
                         AnyStack& astk, TravStats& st, unsigned long long& nclosest, unsigned long long& nshadow) {
    float ul = u24(rng.u[1]);
    int k = -1;
    for (int i = 0; i < sv.nlights; ++i)
        if (ul < sv.lights[i].sel_cdf) { k = i; break; }
    if (k < 0 || sv.lights[k].count <= 0) return mk(0.f, 0.f, 0.f);
    LightGpu L = sv.lights[k];
    const float* mat = sv.mat + RZK_MAT_FLOATS * m;
    v3 p = ld3(rc), pn = ld3(rc + 6), wo = ld3(rc + 9);
    Onb b; b.u = ld3(rc + 12); b.v = ld3(rc + 15); b.n = ld3(rc + 18);
    v3 Lo = mk(0.f, 0.f, 0.f);
    // light sample: triangle by cumulative area ratio (quad.cc:75-87: t <= ratio),
    // uniform point on it (triangle.cc:298-308)
    float ut = u24(rng.u[2]);
    int ti = L.first + L.count - 1;
    for (int q = 0; q < L.count; ++q)
        if (ut <= sv.ltris[L.first + q].cum) { ti = L.first + q; break; }
    LightTri T = sv.ltris[ti];
    float tmp = sqrtf(1.0f - u24(rng.u[3]));
    float bb = 1.0f - tmp;
    float cc = u24(rng.u[4]) * tmp;
    float aa = 1.0f - bb - cc;
    v3 q = (aa * mk(T.p0.x, T.p0.y, T.p0.z) + bb * mk(T.p1.x, T.p1.y, T.p1.z)) + cc * mk(T.p2.x, T.p2.y, T.p2.z);
    v3 qn = mk(T.n.x, T.n.y, T.n.z);
    float pdf = T.ratio * (1.0f / T.n.w);
    if (pdf > 0.0f) {
        nshadow++;
        if (!segment_blocked(sv, p, q, tmin, astk, st)) {
            v3 wi = unit(q - p);
            v3 Le = light_query(L, qn, -wi);
            float G = form_factor_abs(p, pn, q, qn);
            if (G > 0.0f) {
                v3 brdf = bsdf_eval(mat, b, wo, wi);
                float dpdf = bsdf_pdf(mat, b, wo, wi);
                v3 pq = q - p;
                float w1 = balance(pdf * dot(pq, pq) / fabsf(dot(unit(p - q), qn)), dpdf);
                Lo = Lo + (((w1 * G) / pdf) * Le) * brdf;
            }
        }
    }
    // BSDF sample that may reach the same light
    v3 wi;
    float dpdf;
    bsdf_sample(mat, b, wo, &rng.u[5], wi, dpdf);
    if (dpdf > 0.0f) {
        Ray r2 = ray_ctor(p, wi);
        float t = 0.f, tb = 0.f, tg = 0.f;
        nclosest++;
        int tri = closest_hit4(sv, r2, tmin, 1e6f, astk, t, tb, tg, st);
        if (tri >= 0) {
            float4 nf = sv.tri_n[tri];
            int4 sf = sv.surf[__float_as_int(nf.w)];
            if (sf.y == k) {
                v3 qh = p + t * wi;
                v3 n = mk(nf.x, nf.y, nf.z);
                v3 qhn = unit((n + (n - n) * tb) + (n - n) * tg);
                if (sf.z & RZK_SURF_FLIPPED) qhn = -qhn;
                v3 Le = light_query(L, qhn, -wi);
                float pa = 0.f;                           // Quad::pdf at a point of the hit triangle
                for (int qq = 0; qq < L.count; ++qq) {
                    const LightTri& lt = sv.ltris[L.first + qq];
                    if (lt.tri == tri) { pa = lt.ratio * (1.0f / lt.n.w); break; }
                }
                v3 pq = qh - p;
                float lpdf = pa * dot(pq, pq) / fabsf(dot(unit(p - qh), qhn));
                float w2 = balance(dpdf, lpdf);
                v3 brdf = bsdf_eval(mat, b, wo, wi);
                float cosv = fabsf(dot(wi, pn));
                Lo = Lo + (((w2 * cosv) / dpdf) * Le) * brdf;
            }
        }
    }
    return mk(Lo.x / L.sel_pdf, Lo.y / L.sel_pdf, Lo.z / L.sel_pdf);
}

// pdf_map (lightslice.cc:2352-2377) for a local direction of the receiver frame
__device__ __forceinline__ float map_pdf(const GuideView& gv, int slice, int res, v3 local) {
    if (local.z <= 0.0f) return 0.0f;
    float sx, sy;
    direction_to_square(local, sx, sy);
    int col = (int)(sx * res), row = (int)(sy * res);
    if (col >= res) col = res - 1;
    if (row >= res) row = res - 1;
    float pdf_k = gv.maps[((int64_t)slice * res + row) * res + col] / gv.totals[slice];
    float pdf_unit_square = (float)(res * res) * pdf_k;
    float pdf_unit_disk = pdf_unit_square * kInvPi;
    return pdf_unit_disk * 0.5f;
}

// multi_slice_pdf (lightslice.cc:1930-2055) with the receiver-frame map of
// `slice` when it is valid, else the uniform-hemisphere fallback
__device__ __forceinline__ float guide_pdf(const GuideView& gv, int slice, int res, const Onb& b, v3 sn, v3 wi) {
    if (dot(sn, wi) <= 0.0f) return 0.0f;
    if (slice < 0) return kInv2Pi;
    return map_pdf(gv, slice, res, world_to_local(b, wi));
}

// LightSlice::sample -> multi_slice_sample: sample_map of the slice (receiver
// frame, same_local_wi) or the uniform-hemisphere fallback; pdf re-evaluated
// by multi_slice_pdf; returns the receiver BRDF for wi
__device__ v3 guide_sample(const GuideView& gv, int slice, int res, const float* mat, const Onb& b, v3 sn, v3 wo,
                           const uint32_t* u, v3& wi, float& pdf) {
    if (slice < 0) {
        float phi = kTwoPi * u24(u[0]);
        float theta = acosf(u24(u[1]));
        wi = unit(spherical_to_world(b, theta, phi));
    } else {
        // DiscretePdf2D::sample (pdf.cc:171-186): row by the row cdf, column by that row's cdf
        float ur = u24(u[0]), uc = u24(u[1]);
        const float* rcdf = gv.row_cdf + (int64_t)slice * res;
        int row = -1, col = -1;
        for (int i = 0; i < res; ++i)
            if (ur < rcdf[i]) { row = i; break; }
        if (row >= 0) {
            const float* ccdf = gv.col_cdf + ((int64_t)slice * res + row) * res;
            for (int j = 0; j < res; ++j)
                if (uc < ccdf[j]) { col = j; break; }
        }
        if (row < 0 || col < 0) { wi = mk(0.f, 0.f, 0.f); pdf = 0.f; return mk(0.f, 0.f, 0.f); }
        float sx = ((float)col + u24(u[2])) / (float)res;
        float sy = ((float)row + u24(u[3])) / (float)res;
        wi = local_to_world(b, square_to_direction(sx, sy));
    }
    pdf = guide_pdf(gv, slice, res, b, sn, wi);
    return bsdf_eval(mat, b, wo, wi);
}

// Camera::shoot_ray(const SensorSample&), camera.cc:67-80
__device__ __forceinline__ Ray camera_sample_ray(const float* c, float x, float y) {
    Onb b;
    b.u = mk(c[3], c[4], c[5]); b.v = mk(c[6], c[7], c[8]); b.n = mk(c[9], c[10], c[11]);
    float height = c[18];
    float fx = (x - c[12]) / c[14];
    float fy = ((height - y) - c[13]) / c[15];
    v3 d = unit(local_to_world(b, mk(fx, fy, -c[16])));
    return ray_ctor(mk(c[0], c[1], c[2]), d);
}

// Slice of a pixel: its tile when guide_mode 1 and the tile's map is non-empty,
// else -1 (the uniform-hemisphere fallback)
__device__ __forceinline__ int pixel_slice(const PtParams& pp, const GuideView& gv, int x, int y) {
    if (pp.guide_mode != 1) return -1;
    int sx = min(x / pp.tile, pp.tiles_x - 1), sy = min(y / pp.tile, pp.tiles_y - 1);
    int s = sy * pp.tiles_x + sx;
    return gv.totals[s] > 0.0f ? s : -1;
}

// samples per axis when spp is a square (Sensor's stratified jitter, sensor.cc:25-41), else 0
__device__ __forceinline__ int spp_strata(int spp) {
    int strata = 1;
    while ((strata + 1) * (strata + 1) <= spp) ++strata;
    return strata * strata == spp ? strata : 0;
}

// One path sample of pixel (x, y): trace_wellington unrolled into a
// throughput product (vpl.cc:934-1041).  Every random number comes from
// Philox(seed, pixel, sample, bounce), so the value depends on nothing else.
__device__ v3 trace_path(const SceneView& sv, const float* cam, const PtParams& pp, const GuideView& gv, int x,
                         int y, int slice, int strata, int smp, AnyStack& astk, TravStats& st,
                         unsigned long long& nclosest, unsigned long long& nshadow) {
    uint32_t pixel = (uint32_t)(y * pp.width + x);
    PathRng r0 = path_rng(pp.seed, pixel, (uint32_t)smp, 0u);
    float jx = u24(r0.u[0]), jy = u24(r0.u[1]);
    if (strata > 0) {
        jx = ((float)(smp % strata) + jx) / (float)strata;
        jy = ((float)(smp / strata) + jy) / (float)strata;
    }
    Ray ray = camera_sample_ray(cam, (float)x + jx, (float)y + jy);
    v3 L = mk(0.f, 0.f, 0.f), T = mk(1.f, 1.f, 1.f);
    float th = 0.f, tb = 0.f, tg = 0.f;
    nclosest++;
    int tri = closest_hit4(sv, ray, pp.tmin, 1e6f, astk, th, tb, tg, st);
    for (int bounce = 1; tri >= 0 && bounce <= pp.max_bounce; ++bounce) {
        float rc[RZK_RECV_FLOATS];
        int m, lidx;
        fill_receiver(sv, ray, th, tb, tg, tri, rc, m, lidx);
        if (lidx >= 0) {
            if (bounce == 1 && pp.direct) L = L + T * light_query(sv.lights[lidx], ld3(rc + 6), ld3(rc + 9));
            break;
        }
        if (m < 0) break;
        PathRng rng = path_rng(pp.seed, pixel, (uint32_t)smp, (uint32_t)bounce);
        bool brdf_sampling = u24(rng.u[0]) < 0.5f;
        if (bounce > 1 || pp.direct)
            L = L + T * next_event(sv, rc, m, rng, pp.tmin, astk, st, nclosest, nshadow);
        if (bounce >= pp.max_bounce) break;
        const float* mat = sv.mat + RZK_MAT_FLOATS * m;
        Onb b; b.u = ld3(rc + 12); b.v = ld3(rc + 15); b.n = ld3(rc + 18);
        v3 sn = ld3(rc + 6), wo = ld3(rc + 9), p = ld3(rc);
        int gslice = bounce == 1 ? slice : -1;
        v3 wi, brdf;
        float pdf;
        if (brdf_sampling) brdf = bsdf_sample(mat, b, wo, &rng.u[8], wi, pdf);
        else brdf = guide_sample(gv, gslice, pp.res, mat, b, sn, wo, &rng.u[8], wi, pdf);
        if (!(pdf > 0.0f)) break;
        ray = ray_ctor(p, wi);
        nclosest++;
        tri = closest_hit4(sv, ray, pp.tmin, 1e6f, astk, th, tb, tg, st);
        if (tri < 0) break;
        int4 sf = sv.surf[__float_as_int(sv.tri_n[tri].w)];
        if (sf.y >= 0) break;                               // light: counted by the light sample
        float pdf_other = brdf_sampling ? guide_pdf(gv, gslice, pp.res, b, sn, wi) : bsdf_pdf(mat, b, wo, wi);
        float weight = pdf / (pdf + pdf_other);
        float cosv = fabsf(dot(wi, sn));
        float s = ((weight * cosv) / pdf) * 2.0f;
        T = (T * brdf) * s;
    }
    return L;
}

__device__ __forceinline__ void add_ray_counts(unsigned long long* work, unsigned long long a, unsigned long long c) {
    for (int off = 32; off > 0; off >>= 1) { a += __shfl_xor(a, off); c += __shfl_xor(c, off); }
    if ((threadIdx.x & 63) == 0) { atomicAdd(work, a); atomicAdd(work + 1, c); }
}

// One lane per pixel of the window, all its samples in order, summed in
// sample order (MonteCarloIntegrator::integrate, monte_carlo_integrator.cc:45-70)
__global__ __launch_bounds__(128) void k_guided_pt(SceneView sv, CamView cam, PtParams pp, GuideView gv,
                                                   float* __restrict__ out, unsigned long long* __restrict__ work) {
    extern __shared__ int s_stack[];
    AnyStack astk = any_stack(sv, s_stack);
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    TravStats st;
    unsigned long long nclosest = 0, nshadow = 0;
    int64_t npx = (int64_t)pp.w * pp.h;
    if (t < npx) {
        int x = pp.x0 + (int)(t % pp.w), y = pp.y0 + (int)(t / pp.w);
        int slice = pixel_slice(pp, gv, x, y);
        int strata = spp_strata(pp.spp);
        v3 acc = mk(0.f, 0.f, 0.f);
        for (int smp = 0; smp < pp.spp; ++smp)
            acc = acc + trace_path(sv, cam.v, pp, gv, x, y, slice, strata, smp, astk, st, nclosest,
                                   nshadow);
        int64_t o = (int64_t)y * pp.width + x;
        out[3 * o] = acc.x / (float)pp.spp;
        out[3 * o + 1] = acc.y / (float)pp.spp;
        out[3 * o + 2] = acc.z / (float)pp.spp;
    }
    if (work) add_ray_counts(work, nclosest, nshadow);
}

// ---------------------------------------------------------------------------
// Batched form: one persistent kernel over a work queue of (job, 64-pixel
// block, sample block) items.  A wave takes items_per_dequeue items per
// atomic, finds each item's job by binary search over item_first, and writes
// every sample's radiance; k_batch_reduce then sums each pixel's samples in
// sample order, so a job's image is bit-equal to k_guided_pt's.

__device__ __forceinline__ unsigned long long wave_bcast64(unsigned long long v) {
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((unsigned long long)hi << 32) | lo;
}

__global__ void k_batch_reset(const BatchHeader* __restrict__ hdr) {
    if (threadIdx.x == 0) {
        *hdr->counter = 0ull;
        if (hdr->work) { hdr->work[0] = 0ull; hdr->work[1] = 0ull; }
    }
}

__global__ __launch_bounds__(256) void k_batch_pt(SceneView sv, const BatchHeader* __restrict__ hdrp) {
    extern __shared__ int s_stack[];
    const BatchHeader hdr = *hdrp;
    AnyStack astk = any_stack(sv, s_stack);
    const int lane = threadIdx.x & 63;
    TravStats st;
    unsigned long long nclosest = 0, nshadow = 0;
    const unsigned long long n_items = (unsigned long long)hdr.n_items;
    const unsigned long long q = (unsigned long long)hdr.items_per_dequeue;
    const int last_job = (int)hdr.n_jobs - 1;
    for (;;) {
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(hdr.counter, q);
        base = wave_bcast64(base);
        if (base >= n_items) break;
        unsigned long long end = base + q < n_items ? base + q : n_items;
        for (unsigned long long item = base; item < end; ++item) {
            int lo = 0, hi = last_job;
            while (lo < hi) {
                int mid = (lo + hi + 1) >> 1;
                if ((unsigned long long)hdr.jobs[mid].item_first <= item) lo = mid;
                else hi = mid - 1;
            }
            const JobGpu* J = hdr.jobs + lo;
            int64_t local = (int64_t)(item - (unsigned long long)J->item_first);
            int nsb = J->samp_blocks;
            int64_t pb = local / nsb;
            int sb = (int)(local - pb * nsb);
            int64_t p = pb * 64 + lane;
            const PtParams pp = J->pp;
            if (p < (int64_t)pp.w * pp.h) {
                int x = pp.x0 + (int)(p % pp.w), y = pp.y0 + (int)(p / pp.w);
                const GuideView gv = J->gv;
                int slice = pixel_slice(pp, gv, x, y);
                int s0 = sb * J->samples_per_item;
                int s1 = min(pp.spp, s0 + J->samples_per_item);
                float* dst = hdr.samples + J->sample_off + p * pp.spp * 3;
                for (int smp = s0; smp < s1; ++smp) {
                    v3 L = trace_path(sv, J->cam.v, pp, gv, x, y, slice, J->strata, smp, astk, st,
                                      nclosest, nshadow);
                    dst[3 * smp] = L.x;
                    dst[3 * smp + 1] = L.y;
                    dst[3 * smp + 2] = L.z;
                }
            }
        }
    }
    if (hdr.work) add_ray_counts(hdr.work, nclosest, nshadow);
}

__global__ __launch_bounds__(256) void k_batch_reduce(const BatchHeader* __restrict__ hdrp) {
    const BatchHeader hdr = *hdrp;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int last_job = (int)hdr.n_jobs - 1;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < hdr.n_pixels; i += stride) {
        int lo = 0, hi = last_job;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (hdr.jobs[mid].pix_first <= i) lo = mid;
            else hi = mid - 1;
        }
        const JobGpu* J = hdr.jobs + lo;
        int spp = J->pp.spp;
        const float* src = hdr.samples + J->sample_off + (i - J->pix_first) * spp * 3;
        v3 acc = mk(0.f, 0.f, 0.f);
        for (int smp = 0; smp < spp; ++smp) acc = acc + mk(src[3 * smp], src[3 * smp + 1], src[3 * smp + 2]);
        hdr.out[3 * i] = acc.x / (float)spp;
        hdr.out[3 * i + 1] = acc.y / (float)spp;
        hdr.out[3 * i + 2] = acc.z / (float)spp;
    }
}

// ---------------------------------------------------------------------------
// Served form: a round over shares of many requests.  Each lane adds its
// pixel's samples of one sample block to the running sum in sample order and
// writes the mean with the last sample, so the image is the plain render's.

__global__ void k_serve_reset(const ServeHeader* __restrict__ hdr) {
    if (threadIdx.x == 0) *hdr->counter = 0ull;
}

__global__ __launch_bounds__(256) void k_serve_pt(SceneView sv, const ServeHeader* __restrict__ hdrp) {
    extern __shared__ int s_stack[];
    const ServeHeader hdr = *hdrp;
    AnyStack astk = any_stack(sv, s_stack);
    const int lane = threadIdx.x & 63;
    TravStats st;
    unsigned long long nclosest = 0, nshadow = 0;
    const unsigned long long n_items = (unsigned long long)hdr.n_items;
    const unsigned long long q = (unsigned long long)hdr.items_per_dequeue;
    const int last_job = (int)hdr.n_jobs - 1;
    for (;;) {
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(hdr.counter, q);
        base = wave_bcast64(base);
        if (base >= n_items) break;
        unsigned long long end = base + q < n_items ? base + q : n_items;
        for (unsigned long long item = base; item < end; ++item) {
            int lo = 0, hi = last_job;
            while (lo < hi) {
                int mid = (lo + hi + 1) >> 1;
                if ((unsigned long long)hdr.jobs[mid].item_first <= item) lo = mid;
                else hi = mid - 1;
            }
            const ServeJobGpu* J = hdr.jobs + lo;
            int64_t local = (int64_t)(item - (unsigned long long)J->item_first) + J->item_begin;
            int64_t sb = local / J->pix_blocks;
            int64_t pb = local - sb * J->pix_blocks;
            int64_t p = pb * 64 + lane;
            const PtParams pp = J->pp;
            if (p < (int64_t)pp.w * pp.h) {
                int x = pp.x0 + (int)(p % pp.w), y = pp.y0 + (int)(p / pp.w);
                const GuideView gv = J->gv;
                int slice = pixel_slice(pp, gv, x, y);
                int s0 = (int)sb * J->samples_per_item;
                int s1 = min(pp.spp, s0 + J->samples_per_item);
                float* a = J->acc + 3 * p;
                v3 acc = s0 == 0 ? mk(0.f, 0.f, 0.f) : mk(a[0], a[1], a[2]);
                for (int smp = s0; smp < s1; ++smp)
                    acc = acc + trace_path(sv, J->cam.v, pp, gv, x, y, slice, J->strata, smp, astk, st,
                                           nclosest, nshadow);
                if (s1 == pp.spp) {
                    float* o = J->out + 3 * p;
                    o[0] = acc.x / (float)pp.spp;
                    o[1] = acc.y / (float)pp.spp;
                    o[2] = acc.z / (float)pp.spp;
                } else {
                    a[0] = acc.x;
                    a[1] = acc.y;
                    a[2] = acc.z;
                }
            }
        }
    }
}

}  // namespace rzk

using namespace rzk;

void launch_serve_round(hipStream_t stream, const SceneView& sv, const ServeHeader* d_hdr, int grid, int block) {
    hipLaunchKernelGGL(k_serve_reset, dim3(1), dim3(64), 0, stream, d_hdr);
    size_t lds = sizeof(int) * (size_t)block * (size_t)kLdsStack;
    hipLaunchKernelGGL(k_serve_pt, dim3(grid), dim3(block), lds, stream, sv, d_hdr);
}

void launch_batch(hipStream_t stream, const SceneView& sv, const BatchHeader* d_hdr, int grid, int block,
                  int reduce_grid) {
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, stream, d_hdr);
    size_t lds = sizeof(int) * (size_t)block * (size_t)kLdsStack;
    hipLaunchKernelGGL(k_batch_pt, dim3(grid), dim3(block), lds, stream, sv, d_hdr);
    hipLaunchKernelGGL(k_batch_reduce, dim3(reduce_grid), dim3(256), 0, stream, d_hdr);
}

void launch_guided_pt(rzk_ctx* ctx, const SceneView& sv_in, const float* cam, const PtParams& pp, const GuideView& gv,
                      float* d_out, unsigned long long* d_work) {
    int64_t n = (int64_t)pp.w * pp.h;
    if (n <= 0) return;
    SceneView sv = sv_in;
    const int block = 128;
    unsigned grid = (unsigned)((n + block - 1) / block);
    bind_spill(ctx, sv, (int64_t)grid * block, 2 * sv.stack4);
    CamView cv;
    for (int i = 0; i < RZK_CAM_FLOATS; ++i) cv.v[i] = cam[i];
    size_t lds = sizeof(int) * block * (size_t)kLdsStack;
    hipLaunchKernelGGL(k_guided_pt, dim3(grid), dim3(block), lds, ctx->stream, sv, cv, pp, gv, d_out, d_work);
    RZK_HIP(hipGetLastError());
}
