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

using PtStack = AnyStackT<kPtLdsStack>;

// Waves per SIMD the path kernels' registers are budgeted for (512 / waves
// VGPRs each); the LDS columns (kPtLdsStack + kPtCold ints per lane) fit
// that many.
#ifndef RZK_PT_WAVES
#define RZK_PT_WAVES 4
#endif
constexpr int kPtWaves = RZK_PT_WAVES;
// Lanes per block of the one-lane-per-pixel path kernels: one wave, so the
// uneven lengths of the waves' work are balanced at wave granularity.
constexpr int kPtBlock = 64;

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
                         PtStack& astk, TravStats& st, unsigned long long& nclosest, unsigned long long& nshadow) {
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

#ifndef RZK_SCAN_CHUNK
#define RZK_SCAN_CHUNK 16
#endif

// first i in [0, n) with u < cdf[i], -1 when none (cdf non-decreasing, or all
// NaN).  The predicate !(u < cdf[i]) holds on a prefix, so the answer is its
// length: entries are read kScanChunk at a time, all loads of a chunk in
// flight together, instead of one dependent load per halving step.
constexpr int kScanChunk = RZK_SCAN_CHUNK;
__device__ __forceinline__ int first_above(const float* cdf, int n, float u) {
    int cnt = 0;
    for (int base = 0; base < n; base += kScanChunk) {
        float v[kScanChunk];
#pragma unroll
        for (int i = 0; i < kScanChunk; ++i) v[i] = base + i < n ? cdf[base + i] : INFINITY;
        int c = 0;
#pragma unroll
        for (int i = 0; i < kScanChunk; ++i) c += (u < v[i]) ? 0 : 1;
        cnt += c;
        if (c < kScanChunk) break;
    }
    return cnt < n ? cnt : -1;
}

// The direction step of LightSlice::sample -> multi_slice_sample: sample_map
// of the slice (receiver frame, same_local_wi) or the uniform-hemisphere
// fallback.  False (wi = 0) when the slice's map has no cell to pick.
__device__ __forceinline__ bool guide_sample_dir(const GuideView& gv, int slice, int res, const Onb& b,
                                                 const uint32_t* u, v3& wi) {
    if (slice < 0) {
        float phi = kTwoPi * u24(u[0]);
        float theta = xacos(u24(u[1]));
        wi = unit(spherical_to_world(b, theta, phi));
        return true;
    }
    // DiscretePdf2D::sample (pdf.cc:171-186): row by the row cdf, column by that row's cdf
    // the first index with u < cdf[i]; the predicate is monotone in i
    // (cumulative sums, the last row entry forced to 1 > u), so first_above
    // finds the index the reference's linear scan finds
    float ur = u24(u[0]), uc = u24(u[1]);
    const float* rcdf = gv.row_cdf + (int64_t)slice * res;
    int row = first_above(rcdf, res, ur), col = -1;
    if (row >= 0) col = first_above(gv.col_cdf + ((int64_t)slice * res + row) * res, res, uc);
    if (row < 0 || col < 0) { wi = mk(0.f, 0.f, 0.f); return false; }
    float sx = ((float)col + u24(u[2])) / (float)res;
    float sy = ((float)row + u24(u[3])) / (float)res;
    wi = local_to_world(b, square_to_direction(sx, sy));
    return true;
}

// LightSlice::sample -> multi_slice_sample: the direction, its pdf
// re-evaluated by multi_slice_pdf; returns the receiver BRDF for wi
__device__ __forceinline__ v3 guide_sample(const GuideView& gv, int slice, int res, const float* mat, const Onb& b, v3 sn, v3 wo,
                           const uint32_t* u, v3& wi, float& pdf) {
    if (!guide_sample_dir(gv, slice, res, b, u, wi)) { pdf = 0.f; return mk(0.f, 0.f, 0.f); }
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

// Slice of a pixel: its tile (or, with kd-tree slices, the slice of the
// pixel's gather point) when guide_mode 1 and the slice's map is non-empty,
// else -1 (the uniform-hemisphere fallback)
__device__ __forceinline__ int pixel_slice(const PtParams& pp, const GuideView& gv, int x, int y) {
    if (pp.guide_mode != 1) return -1;
    int s;
    if (pp.pix_slice) {
        s = pp.pix_slice[(int64_t)y * pp.width + x];
        if (s < 0) return -1;
    } else {
        int sx = min(x / pp.tile, pp.tiles_x - 1), sy = min(y / pp.tile, pp.tiles_y - 1);
        s = sy * pp.tiles_x + sx;
    }
    return gv.totals[s] > 0.0f ? s : -1;
}

// samples per axis when spp is a square (Sensor's stratified jitter, sensor.cc:25-41), else 0
__device__ __forceinline__ int spp_strata(int spp) {
    int strata = 1;
    while ((strata + 1) * (strata + 1) <= spp) ++strata;
    return strata * strata == spp ? strata : 0;
}

// One path sample from primary ray `ray` of pixel (x, y): trace_wellington
// unrolled into a throughput product (vpl.cc:934-1041).  The numbers of bounce
// b come from Philox(seed, pixel, sample, b), so the value depends on nothing else.
__device__ v3 trace_from(const SceneView& sv, const PtParams& pp, const GuideView& gv, Ray ray, int x, int y,
                         int slice, int smp, PtStack& astk, TravStats& st, unsigned long long& nclosest,
                         unsigned long long& nshadow) {
    uint32_t pixel = (uint32_t)(y * pp.width + x);
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

// One path sample of pixel (x, y): the camera ray through the sample's jittered
// position (2-D strata when spp is a square), then trace_from.
__device__ v3 trace_path(const SceneView& sv, const float* cam, const PtParams& pp, const GuideView& gv, int x,
                         int y, int slice, int strata, int smp, PtStack& astk, TravStats& st,
                         unsigned long long& nclosest, unsigned long long& nshadow) {
    uint32_t pixel = (uint32_t)(y * pp.width + x);
    PathRng r0 = path_rng(pp.seed, pixel, (uint32_t)smp, 0u);
    float jx = u24(r0.u[0]), jy = u24(r0.u[1]);
    if (strata > 0) {
        jx = ((float)(smp % strata) + jx) / (float)strata;
        jy = ((float)(smp / strata) + jy) / (float)strata;
    }
    Ray ray = camera_sample_ray(cam, (float)x + jx, (float)y + jy);
    return trace_from(sv, pp, gv, ray, x, y, slice, smp, astk, st, nclosest, nshadow);
}

// ---------------------------------------------------------------------------
// The same path samples as a per-lane state machine with one traversal site.
// trace_path runs its four traversals (camera or continuation ray, shadow
// segment, next-event BSDF ray) at four places in the code, so the lanes of a
// wave trace together only while they are at the same step of the same
// bounce, and a lane whose path has ended waits for the longest path of the
// wave.  Here each lane runs its samples as stages between traversal queries:
// every loop turn the lane consumes its last query's result and runs on to
// its next query (starting its next sample when a path ends), and then every
// lane of the wave traverses its own query at the same place.  The arithmetic
// of each stage is trace_path's, in the same order, so every sample gives the
// same bits.  Shadow segments use closest_hit4's any-hit mode (any_hit4's
// answer).  State that outlives a query but is only read while shading (the
// receiver record, the next event's light sample and partial sum, the
// continuation's direction) lives in LDS, one column per lane; the rest stays
// in registers.

enum : int {
    kPsSample = 0,      // start the next sample (no query pending)
    kPsPrimary,         // result: the sample's first closest hit
    kPsShadow,          // result: the next event's shadow segment
    kPsNeeHit,          // result: the next event's BSDF ray
    kPsContHit,         // result: the continuation ray
    kPsNeeSample,       // next: the next event's BSDF sample (no query pending)
    kPsContinue,        // next: the continuation's direction sample (no query pending)
    kPsLightTerm,       // next: the light sample's term, then its shadow segment
};

// LDS columns of one lane (a float each, kPtCold of them)
enum : int {
    kPcP = 0,                           // receiver: point (3), shading normal (3), wo (3),
    kPcSn = kPcP + 3,                   //   frame u, v, n (9); the geometric normal is
    kPcWo = kPcSn + 3,                  //   not read after the hit
    kPcOnb = kPcWo + 3,
    kPcLo = kPcOnb + 9,                 // next event: partial sum (3)
    kPcLterm = kPcLo + 3,               //   the light sample's term (3), added if its segment is clear
    kPcWi2 = kPcLterm + 3,              //   BSDF direction (3), pdf
    kPcDpdf2 = kPcWi2 + 3,
    kPcBrdf2 = kPcLterm,                //   BSDF sample's BRDF (3; the light term is added by then)
    kPcQl = kPcLterm,                   //   light sample point (3), normal (3), area pdf, from the
    kPcQn = kPcWi2,                     //   sample to its term
    kPcLpdf = kPcDpdf2,
    kPcWi = kPcLo,                      // continuation (after the next event): direction (3),
    kPcBrdf = kPcWi + 3,                //   BRDF (3), pdf, the other technique's pdf
    kPcPdf = kPcBrdf + 3,
    kPcPdfO = kPcPdf + 1,
    kPtCold = kPcDpdf2 + 1,
};

struct PtCold {
    float* col;                          // this lane's first column entry
    int pitch;                           // floats between a lane's columns (the block size)
    __device__ float get(int f) const { return col[f * pitch]; }
    __device__ void set(int f, float v) const { col[f * pitch] = v; }
    __device__ v3 get3(int f) const { return mk(get(f), get(f + 1), get(f + 2)); }
    __device__ void set3(int f, v3 v) const { set(f, v.x); set(f + 1, v.y); set(f + 2, v.z); }
    __device__ Onb onb() const {
        Onb b; b.u = get3(kPcOnb); b.v = get3(kPcOnb + 3); b.n = get3(kPcOnb + 6);
        return b;
    }
};

// cold columns after the traversal stack's kPtLdsStack ints per lane
__device__ __forceinline__ PtCold pt_cold(int* s_stack) {
    PtCold c;
    c.pitch = (int)blockDim.x;
    c.col = reinterpret_cast<float*>(s_stack + kPtLdsStack * blockDim.x) + threadIdx.x;
    return c;
}

struct PtLane {
    int stage;
    int smp, bounce;
    int m;                               // the receiver's material
    int k;                               // the next event's light
    bool brdf_sampling;
    v3 L, T;
    int qtri;                            // last query's result
    float qt, qb, qg;
};

struct PtQuery {
    Ray r;
    float tmax;
    bool any;
};

// numbers 4j .. 4j + 3 of path_rng(seed, pixel, sample, bounce)
__device__ __forceinline__ void pt_rng_block(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t bounce,
                                             uint32_t j, uint32_t* u) {
    uint32_t c[4] = {sample, bounce, j, 0x52a7u};
    philox4x32_10(c, seed ^ 0x9a3c5e17u, pixel);
    u[0] = c[0]; u[1] = c[1]; u[2] = c[2]; u[3] = c[3];
}

// where a lane's finished samples go: the pixel's running sum (out == NULL),
// or sample smp's value to out + (smp - smp0) * stride
struct PtSink {
    float* out;
    int64_t stride;
    int smp0;
    v3 acc;
};

// a path sample has ended: its value to the sink, the next sample next
__device__ __forceinline__ void pt_sample_done(PtLane& s, PtSink& k) {
    if (k.out) {
        float* d = k.out + (int64_t)(s.smp - k.smp0) * k.stride;
        d[0] = s.L.x;
        d[1] = s.L.y;
        d[2] = s.L.z;
    } else {
        k.acc = k.acc + s.L;
    }
    ++s.smp;
    s.stage = kPsSample;
}

// next_event's end: L += T * (Lo / the light's selection pdf)
__device__ __forceinline__ void pt_nee_end(const SceneView& sv, const PtCold& c, PtLane& s) {
    const float sp = sv.lights[s.k].sel_pdf;
    const v3 Lo = c.get3(kPcLo);
    s.L = s.L + s.T * mk(Lo.x / sp, Lo.y / sp, Lo.z / sp);
}

// next_event's start (numbers 1..4): the light and its sample.  Returns 1
// with the sample staged in the lane's columns (its term and shadow segment
// come next), 2 when the light sample has no positive pdf (the BSDF sample
// comes next), 0 when no light is chosen (the next event adds nothing and
// ends).
__device__ __forceinline__ int pt_nee_begin(const SceneView& sv, const PtParams& pp, const PtCold& c, uint32_t pixel,
                                            const uint32_t* u0, PtLane& s) {
    float ul = u24(u0[1]);
    int k = -1;
    for (int i = 0; i < sv.nlights; ++i)
        if (ul < sv.lights[i].sel_cdf) { k = i; break; }
    if (k < 0 || sv.lights[k].count <= 0) {
        s.L = s.L + s.T * mk(0.f, 0.f, 0.f);
        return 0;
    }
    s.k = k;
    const LightGpu& Lg = sv.lights[k];
    c.set3(kPcLo, mk(0.f, 0.f, 0.f));
    uint32_t u1[4];
    pt_rng_block(pp.seed, pixel, (uint32_t)s.smp, (uint32_t)s.bounce, 1u, u1);
    float ut = u24(u0[2]);
    int ti = Lg.first + Lg.count - 1;
    for (int qq = 0; qq < Lg.count; ++qq)
        if (ut <= sv.ltris[Lg.first + qq].cum) { ti = Lg.first + qq; break; }
    LightTri T = sv.ltris[ti];
    float tmp = sqrtf(1.0f - u24(u0[3]));
    float bb = 1.0f - tmp;
    float cc = u24(u1[0]) * tmp;
    float aa = 1.0f - bb - cc;
    v3 ql = (aa * mk(T.p0.x, T.p0.y, T.p0.z) + bb * mk(T.p1.x, T.p1.y, T.p1.z)) + cc * mk(T.p2.x, T.p2.y, T.p2.z);
    float lpdf = T.ratio * (1.0f / T.n.w);
    if (lpdf > 0.0f) {
        c.set3(kPcQl, ql);
        c.set3(kPcQn, mk(T.n.x, T.n.y, T.n.z));
        c.set(kPcLpdf, lpdf);
        s.stage = kPsLightTerm;
        return 1;
    }
    return 2;
}

// Runs lane s from its last query's result to its next query (true, query in
// q) or past its last sample (false).  A finished sample's value goes to the
// sink (a running sum in sample order, or one slot per sample).  Samples
// s.smp .. smp_end - 1 of pixel (x, y); cam: the camera (NULL: the lane's one
// sample starts from q.r as given).
__device__ __forceinline__ bool pt_advance(const SceneView& sv, const float* cam, const PtParams& pp, const GuideView& gv,
                           const PtCold& c, int x, int y, int slice, int strata, int smp_end, PtLane& s, PtQuery& q,
                           PtSink& acc, unsigned long long& nclosest, unsigned long long& nshadow) {
    const uint32_t pixel = (uint32_t)(y * pp.width + x);
    for (;;) {
        bool at_hit = false;              // a new closest hit of the path's ray q.r
        switch (s.stage) {
        case kPsSample: {
            if (s.smp >= smp_end) return false;
            if (cam) {
                PathRng r0 = path_rng(pp.seed, pixel, (uint32_t)s.smp, 0u);
                float jx = u24(r0.u[0]), jy = u24(r0.u[1]);
                if (strata > 0) {
                    jx = ((float)(s.smp % strata) + jx) / (float)strata;
                    jy = ((float)(s.smp / strata) + jy) / (float)strata;
                }
                q.r = camera_sample_ray(cam, (float)x + jx, (float)y + jy);
            }
            s.L = mk(0.f, 0.f, 0.f);
            s.T = mk(1.f, 1.f, 1.f);
            nclosest++;
            q.tmax = 1e6f;
            q.any = false;
            s.stage = kPsPrimary;
            return true;
        }
        case kPsPrimary:
            s.bounce = 1;
            at_hit = true;
            break;
        case kPsShadow: {
            // a clear segment adds the light term (zero when G <= 0, which
            // leaves the zero partial sum as it is)
            if (s.qtri < 0) c.set3(kPcLo, c.get3(kPcLo) + c.get3(kPcLterm));
            s.stage = kPsNeeSample;
            break;
        }
        case kPsNeeHit: {
            const int tri = s.qtri;
            if (tri >= 0) {
                float4 nf = sv.tri_n[tri];
                int4 sf = sv.surf[__float_as_int(nf.w)];
                if (sf.y == s.k) {
                    const LightGpu& Lg = sv.lights[s.k];
                    const v3 p = c.get3(kPcP), pn = c.get3(kPcSn);
                    const v3 wi = c.get3(kPcWi2);
                    const float dpdf = c.get(kPcDpdf2);
                    v3 qh = p + s.qt * wi;
                    v3 n = mk(nf.x, nf.y, nf.z);
                    v3 qhn = unit((n + (n - n) * s.qb) + (n - n) * s.qg);
                    if (sf.z & RZK_SURF_FLIPPED) qhn = -qhn;
                    v3 Le = light_query(Lg, qhn, -wi);
                    float pa = 0.f;
                    for (int qq = 0; qq < Lg.count; ++qq) {
                        const LightTri& lt = sv.ltris[Lg.first + qq];
                        if (lt.tri == tri) { pa = lt.ratio * (1.0f / lt.n.w); break; }
                    }
                    v3 pq = qh - p;
                    float lpdf = pa * dot(pq, pq) / fabsf(dot(unit(p - qh), qhn));
                    float w2 = balance(dpdf, lpdf);
                    v3 brdf = c.get3(kPcBrdf2);
                    float cosv = fabsf(dot(wi, pn));
                    c.set3(kPcLo, c.get3(kPcLo) + (((w2 * cosv) / dpdf) * Le) * brdf);
                }
            }
            pt_nee_end(sv, c, s);
            s.stage = kPsContinue;
            break;
        }
        case kPsContHit: {
            // A miss or a light ends the sample (a light is counted by the
            // light sample): the bounce loop's head below ends it on the
            // receiver it finds, after a throughput update that then goes
            // unused, so the hit's surface is read once, there.
            const v3 sn = c.get3(kPcSn);
            const v3 wi = c.get3(kPcWi);
            const float pdf = c.get(kPcPdf);
            const float pdf_other = c.get(kPcPdfO);
            float weight = pdf / (pdf + pdf_other);
            float cosv = fabsf(dot(wi, sn));
            float sc = ((weight * cosv) / pdf) * 2.0f;
            s.T = (s.T * c.get3(kPcBrdf)) * sc;
            ++s.bounce;
            at_hit = true;
            break;
        }
        case kPsNeeSample:
        case kPsContinue:
            break;
        }
        if (at_hit) {
            // the bounce loop's head with the hit of q.r
            if (!(s.qtri >= 0 && s.bounce <= pp.max_bounce)) { pt_sample_done(s, acc); continue; }
            float rc[RZK_RECV_FLOATS];
            int lidx;
            fill_receiver(sv, q.r, s.qt, s.qb, s.qg, s.qtri, rc, s.m, lidx);
            if (lidx >= 0) {
                if (s.bounce == 1 && pp.direct)
                    s.L = s.L + s.T * light_query(sv.lights[lidx], ld3(rc + 6), ld3(rc + 9));
                pt_sample_done(s, acc);
                continue;
            }
            if (s.m < 0) { pt_sample_done(s, acc); continue; }
            c.set3(kPcP, ld3(rc));
            c.set3(kPcSn, ld3(rc + 6));
            c.set3(kPcWo, ld3(rc + 9));
#pragma unroll
            for (int i = 0; i < 9; ++i) c.set(kPcOnb + i, rc[12 + i]);
            uint32_t u0[4];
            pt_rng_block(pp.seed, pixel, (uint32_t)s.smp, (uint32_t)s.bounce, 0u, u0);
            s.brdf_sampling = u24(u0[0]) < 0.5f;
            s.stage = kPsContinue;
            if (s.bounce > 1 || pp.direct) {
                const int nee = pt_nee_begin(sv, pp, c, pixel, u0, s);
                if (nee == 2) s.stage = kPsNeeSample;
            }
        }
        // One site for the light sample's term (AreaLight::first_bounce's
        // light-sampling half), the next event's BSDF sample (numbers 5..7 of
        // block 1, next_event's) and the continuation (numbers 8..11, after
        // the next event): every lane's BSDF pdf and BRDF are evaluated on one
        // path, so a wave whose lanes want any of them runs it once.
        if (s.stage == kPsContinue && s.bounce >= pp.max_bounce) { pt_sample_done(s, acc); continue; }
        {
            const bool light = s.stage == kPsLightTerm;
            const bool nee = s.stage == kPsNeeSample;
            const float* mat = sv.mat + RZK_MAT_FLOATS * s.m;
            const Onb b = c.onb();
            const v3 sn = c.get3(kPcSn), wo = c.get3(kPcWo), p = c.get3(kPcP);
            const int gslice = s.bounce == 1 ? slice : -1;
            uint32_t u[4];
            pt_rng_block(pp.seed, pixel, (uint32_t)s.smp, (uint32_t)s.bounce, nee ? 1u : 2u, u);
            const uint32_t ub[3] = {nee ? u[1] : u[0], nee ? u[2] : u[1], nee ? u[3] : u[2]};
            // the direction by the lane's technique, then, on one path for
            // every lane, the values bsdf_sample / guide_sample and the
            // continuation's hit would compute for it: the BSDF pdf, the
            // guide pdf (continuations only) and the BRDF
            const bool by_bsdf = nee || s.brdf_sampling;
            v3 wi;
            float lam_pdf = 0.0f;
            int kind;
            if (light) {
                wi = unit(c.get3(kPcQl) - p);
                kind = 1;
            } else if (by_bsdf) {
                kind = bsdf_sample_dir(mat, b, wo, ub, wi, lam_pdf);
            } else {
                kind = guide_sample_dir(gv, gslice, pp.res, b, u, wi) ? 1 : 0;
            }
            float pdf = 0.0f, pdf_other = 0.0f, pdf_b = 0.0f;
            v3 brdf = mk(0.f, 0.f, 0.f);
            if (kind != 0) {
                pdf_b = bsdf_pdf(mat, b, wo, wi);
                const float pdf_g = (nee || light) ? 0.0f : guide_pdf(gv, gslice, pp.res, b, sn, wi);
                brdf = bsdf_eval(mat, b, wo, wi);
                pdf = !by_bsdf ? pdf_g : kind == 2 ? lam_pdf : pdf_b;
                pdf_other = s.brdf_sampling ? pdf_g : pdf_b;
            }
            if (light) {
                // the term, added once the segment p -> ql turns out clear
                // (segment_blocked(p, q), the shadow query)
                const v3 ql = c.get3(kPcQl), qn = c.get3(kPcQn);
                const float lpdf = c.get(kPcLpdf);
                const v3 Le = light_query(sv.lights[s.k], qn, -wi);
                const float G = form_factor_abs(p, sn, ql, qn);
                v3 term = mk(0.f, 0.f, 0.f);
                if (G > 0.0f) {
                    v3 pq = ql - p;
                    float w1 = balance(lpdf * dot(pq, pq) / fabsf(dot(unit(p - ql), qn)), pdf_b);
                    term = (((w1 * G) / lpdf) * Le) * brdf;
                }
                c.set3(kPcLterm, term);
                nshadow++;
                q.r = ray_ctor(p, ql - p);
                q.tmax = length(ql - p) - pp.tmin;
                q.any = true;
                s.stage = kPsShadow;
                return true;
            }
            if (!(pdf > 0.0f)) {
                if (nee) {
                    pt_nee_end(sv, c, s);
                    s.stage = kPsContinue;
                } else {
                    pt_sample_done(s, acc);
                }
                continue;
            }
            if (nee) {
                c.set3(kPcWi2, wi);
                c.set(kPcDpdf2, pdf);
                c.set3(kPcBrdf2, brdf);
                s.stage = kPsNeeHit;
            } else {
                c.set3(kPcWi, wi);
                c.set3(kPcBrdf, kind == 2 ? mk(mat[1], mat[2], mat[3]) * kInvPi : brdf);
                c.set(kPcPdf, pdf);
                c.set(kPcPdfO, pdf_other);
                s.stage = kPsContHit;
            }
            nclosest++;
            q.r = ray_ctor(p, wi);
            q.tmax = 1e6f;
            q.any = false;
            return true;
        }
    }
}

__device__ __forceinline__ void add_ray_counts(unsigned long long* work, unsigned long long a, unsigned long long c) {
    for (int off = 32; off > 0; off >>= 1) { a += __shfl_xor(a, off); c += __shfl_xor(c, off); }
    if ((threadIdx.x & 63) == 0) { atomicAdd(work, a); atomicAdd(work + 1, c); }
}

// Lanes still walking the tree at or below which a wave stops walking and lets
// its waiting lanes shade and start their next query; the walkers resume from
// where they stopped on the next pass.
#ifndef RZK_TRAV_REGEN
#define RZK_TRAV_REGEN 8
#endif
constexpr int kTravRegen = RZK_TRAV_REGEN;

// The traversal phase of the path kernels: walk the wave's open queries one
// node at a time until none is open, or until at most kTravRegen are and some
// lane is waiting to shade.  Each query's visits are closest_hit4's, so the
// answers do not depend on when the phase stops.
__device__ __forceinline__ void trav_phase(const SceneView& sv, const Ray& r, float tmin, Trav4& tv,
                                           PtStack& astk, PtLane& ps, TravStats& st, bool busy, bool& trav) {
    for (;;) {
        const unsigned long long tm = __ballot(trav);
        if (tm == 0ull) break;
        if (kTravRegen > 0 && __popcll(tm) <= kTravRegen && __ballot(busy && !trav) != 0ull) break;
        if (trav) trav = trav4_step(sv, r, tmin, tv, astk, ps.qt, ps.qb, ps.qg, st);
    }
}

// One lane per pixel of the window, all its samples in order (the state
// machine above), summed in sample order (MonteCarloIntegrator::integrate,
// monte_carlo_integrator.cc:45-70)
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(kPtWaves))) void k_guided_pt(SceneView sv, CamView cam, PtParams pp, GuideView gv,
                                                   float* __restrict__ out, unsigned long long* __restrict__ work) {
    extern __shared__ int s_stack[];
    PtStack astk = any_stack<kPtLdsStack>(sv, s_stack);
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    TravStats st;
    unsigned long long nclosest = 0, nshadow = 0;
    const int64_t npx = (int64_t)pp.w * pp.h;
    const bool own = t < npx;
    const int x = own ? pp.x0 + (int)(t % pp.w) : 0, y = own ? pp.y0 + (int)(t / pp.w) : 0;
    const int slice = own ? pixel_slice(pp, gv, x, y) : 0;
    const int strata = spp_strata(pp.spp);
    PtSink sink{nullptr, 0, 0, mk(0.f, 0.f, 0.f)};
    const PtCold cold = pt_cold(s_stack);
    PtLane ps;
    ps.stage = kPsSample;
    ps.smp = 0;
    PtQuery q;
    Trav4 tv;
    tv.best = -1;
    bool busy = own, trav = false;
    for (;;) {
        if (busy && !trav) {
            ps.qtri = tv.best;
            busy = pt_advance(sv, cam.v, pp, gv, cold, x, y, slice, strata, pp.spp, ps, q, sink, nclosest, nshadow);
            if (busy) trav = trav4_begin(sv, q.r, pp.tmin, q.tmax, q.any, tv, ps.qt, ps.qb, ps.qg, st);
        }
        if (__ballot(busy) == 0ull) break;
        trav_phase(sv, q.r, pp.tmin, tv, astk, ps, st, busy, trav);
    }
    if (own) {
        const v3 acc = sink.acc;
        int64_t o = (int64_t)y * pp.width + x;
        out[3 * o] = acc.x / (float)pp.spp;
        out[3 * o + 1] = acc.y / (float)pp.spp;
        out[3 * o + 2] = acc.z / (float)pp.spp;
    }
    if (work) {
        add_ray_counts(work, nclosest, nshadow);
        add_work_totals(work + 2, st);        // 4-wide nodes visited, triangles tested (per lane)
    }
}

// One lane per given primary ray: ray i (a ray record) is sample smp[i] of pixel
// (pix[2i], pix[2i+1]); out[3i..] is its radiance, unaveraged
// (MonteCarloIntegrator::radiance(PrimaryRay&), monte_carlo_integrator.cc:40-43).
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(kPtWaves))) void k_guided_pt_rays(SceneView sv, PtParams pp, GuideView gv,
                                                        const float* __restrict__ rays, const int* __restrict__ pix,
                                                        const int* __restrict__ smp, int64_t n,
                                                        float* __restrict__ out) {
    extern __shared__ int s_stack[];
    PtStack astk = any_stack<kPtLdsStack>(sv, s_stack);
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    TravStats st;
    unsigned long long nclosest = 0, nshadow = 0;
    const bool own = i < n;
    const int64_t ii = own ? i : 0;
    const float* rp = rays + 6 * ii;
    const int x = own ? pix[2 * ii] : 0, y = own ? pix[2 * ii + 1] : 0;
    const int sm = own ? smp[ii] : 0;
    const int slice = own ? pixel_slice(pp, gv, x, y) : 0;
    PtSink sink{out + 3 * ii, 3, sm, mk(0.f, 0.f, 0.f)};
    const PtCold cold = pt_cold(s_stack);
    PtLane ps;
    ps.stage = kPsSample;
    ps.smp = sm;
    PtQuery q;
    if (own) q.r = ray_from_od(mk(rp[0], rp[1], rp[2]), mk(rp[3], rp[4], rp[5]));
    Trav4 tv;
    tv.best = -1;
    bool busy = own, trav = false;
    for (;;) {
        if (busy && !trav) {
            ps.qtri = tv.best;
            busy = pt_advance(sv, nullptr, pp, gv, cold, x, y, slice, 0, sm + 1, ps, q, sink, nclosest, nshadow);
            if (busy) trav = trav4_begin(sv, q.r, pp.tmin, q.tmax, q.any, tv, ps.qt, ps.qb, ps.qg, st);
        }
        if (__ballot(busy) == 0ull) break;
        trav_phase(sv, q.r, pp.tmin, tv, astk, ps, st, busy, trav);
    }
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
        *hdr->counter = (unsigned long long)hdr->item_lo;
        if (hdr->work) { hdr->work[0] = 0ull; hdr->work[1] = 0ull; }
    }
}

// the job of work item `item`: the last job (of the ring window job_lo ..
// n_jobs - 1) whose first item is at most `item`
__device__ __forceinline__ int64_t batch_job(const BatchHeader& hdr, unsigned long long item) {
    int64_t lo = hdr.job_lo, hi = hdr.n_jobs - 1;
    while (lo < hi) {
        int64_t mid = (lo + hi + 1) >> 1;
        if ((unsigned long long)hdr.jobs[mid & hdr.job_mask].item_first <= item) lo = mid;
        else hi = mid - 1;
    }
    return lo & hdr.job_mask;
}

// the pixel and samples of item `item` (of job J, parameters pp and gv) for
// this lane; false when the item's pixel block has no pixel for the lane
__device__ __forceinline__ bool batch_lane(const BatchHeader& hdr, const JobGpu* J, const PtParams& pp,
                                           const GuideView& gv, unsigned long long item, int lane, int& x, int& y,
                                           int& slice, int& s1, PtLane& ps, PtSink& sink) {
    const int64_t local = (int64_t)(item - (unsigned long long)J->item_first);
    const int nsb = J->samp_blocks;
    const int64_t pb = local / nsb;
    const int sb = (int)(local - pb * nsb);
    const int64_t p = pb * 64 + lane;
    const int64_t npx = (int64_t)pp.w * pp.h;
    if (p >= npx) return false;
    x = pp.x0 + (int)(p % pp.w);
    y = pp.y0 + (int)(p / pp.w);
    slice = pixel_slice(pp, gv, x, y);
    const int s0 = sb * J->samples_per_item;
    s1 = min(pp.spp, s0 + J->samples_per_item);
    ps.stage = kPsSample;
    ps.smp = s0;
    sink.out = hdr.samples + J->sample_off + p * 3;
    sink.stride = npx * 3;
    sink.smp0 = 0;
    return true;
}

// item `item` for this lane: its job, pixel and samples; false when the
// item's pixel block has no pixel for the lane
__device__ __forceinline__ bool batch_item(const BatchHeader& hdr, unsigned long long item, int lane,
                                           const JobGpu*& J, int& x, int& y, int& slice, int& s1, PtLane& ps,
                                           PtSink& sink) {
    J = hdr.jobs + batch_job(hdr, item);
    return batch_lane(hdr, J, J->pp, J->gv, item, lane, x, y, slice, s1, ps, sink);
}

// a wave-uniform value held in scalar registers: each 32-bit word of src
// read from the wave's first active lane (every lane holds the same value)
template <class T>
__device__ __forceinline__ T wave_uniform(const T& src) {
    static_assert(sizeof(T) % 4 == 0, "whole 32-bit words");
    T out;
    const int* s = reinterpret_cast<const int*>(&src);
    int* d = reinterpret_cast<int*>(&out);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) d[i] = __builtin_amdgcn_readfirstlane(s[i]);
    return out;
}

// One item per dequeue: the wave's lanes run the 64 pixels of one item, so
// the item's job is the same for every lane and its parameters are held in
// scalar registers.  Lanes walk and shade in phases as in k_guided_pt.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kPtWaves))) void k_batch_pt_one(
    SceneView sv, const BatchHeader* __restrict__ hdrp) {
    extern __shared__ int s_stack[];
    const BatchHeader hdr = *hdrp;
    PtStack astk = any_stack<kPtLdsStack>(sv, s_stack);
    const PtCold cold = pt_cold(s_stack);
    const int lane = threadIdx.x & 63;
    TravStats st;
    unsigned long long nclosest = 0, nshadow = 0;
    const unsigned long long n_items = (unsigned long long)hdr.n_items;
    for (;;) {
        unsigned long long item = 0;
        if (lane == 0) item = atomicAdd(hdr.counter, 1ull);
        item = wave_bcast64(item);
        if (item >= n_items) break;
        const JobGpu* J = hdr.jobs + wave_uniform(batch_job(hdr, item));
        const PtParams pp = wave_uniform(J->pp);
        const GuideView gv = wave_uniform(J->gv);
        const CamView cam = wave_uniform(J->cam);
        const int strata = wave_uniform(J->strata);
        int x = 0, y = 0, slice = -1, s1 = 0;
        PtLane ps;
        PtQuery q;
        PtSink sink{nullptr, 0, 0, mk(0.f, 0.f, 0.f)};
        bool busy = batch_lane(hdr, J, pp, gv, item, lane, x, y, slice, s1, ps, sink), trav = false;
        Trav4 tv;
        tv.best = -1;
        for (;;) {
            if (busy && !trav) {
                ps.qtri = tv.best;
                busy = pt_advance(sv, cam.v, pp, gv, cold, x, y, slice, strata, s1, ps, q, sink, nclosest, nshadow);
                if (busy) trav = trav4_begin(sv, q.r, pp.tmin, q.tmax, q.any, tv, ps.qt, ps.qb, ps.qg, st);
            }
            if (__ballot(busy) == 0ull) break;
            trav_phase(sv, q.r, pp.tmin, tv, astk, ps, st, busy, trav);
        }
    }
    if (hdr.work) add_ray_counts(hdr.work, nclosest, nshadow);
}

// A wave takes items_per_dequeue items per atomic; each lane then runs its
// pixel of those items one after the other through the path state machine,
// so a lane whose paths end early goes on to its next item while the others
// are still tracing.  Every sample's radiance goes to its slot of the job's
// per-sample buffer.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kPtWaves))) void k_batch_pt(
    SceneView sv, const BatchHeader* __restrict__ hdrp) {
    extern __shared__ int s_stack[];
    const BatchHeader hdr = *hdrp;
    PtStack astk = any_stack<kPtLdsStack>(sv, s_stack);
    const PtCold cold = pt_cold(s_stack);
    const int lane = threadIdx.x & 63;
    TravStats st;
    unsigned long long nclosest = 0, nshadow = 0;
    const unsigned long long n_items = (unsigned long long)hdr.n_items;
    const unsigned long long qd = (unsigned long long)hdr.items_per_dequeue;
    for (;;) {
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(hdr.counter, qd);
        base = wave_bcast64(base);
        if (base >= n_items) break;
        const unsigned long long end = base + qd < n_items ? base + qd : n_items;
        unsigned long long item = base;
        const JobGpu* J = nullptr;
        int x = 0, y = 0, slice = -1, s1 = 0;
        bool have = false;
        PtLane ps;
        PtQuery q;
        PtSink sink{nullptr, 0, 0, mk(0.f, 0.f, 0.f)};
        Trav4 tv;
        tv.best = -1;
        float tmin = 0.f;
        bool busy = true, trav = false;
        for (;;) {
            if (busy && !trav) {
                ps.qtri = tv.best;
                busy = false;
                while (item < end) {
                    if (!have) {
                        have = batch_item(hdr, item, lane, J, x, y, slice, s1, ps, sink);
                        if (!have) { ++item; continue; }
                    }
                    busy = pt_advance(sv, J->cam.v, J->pp, J->gv, cold, x, y, slice, J->strata, s1, ps, q, sink,
                                      nclosest, nshadow);
                    if (busy) break;
                    have = false;
                    ++item;
                }
                if (busy) {
                    tmin = J->pp.tmin;
                    trav = trav4_begin(sv, q.r, tmin, q.tmax, q.any, tv, ps.qt, ps.qb, ps.qg, st);
                }
            }
            if (__ballot(busy) == 0ull) break;
            trav_phase(sv, q.r, tmin, tv, astk, ps, st, busy, trav);
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
        const int64_t npx = (int64_t)J->pp.w * J->pp.h;
        const float* src = hdr.samples + J->sample_off + (i - J->pix_first) * 3;
        v3 acc = mk(0.f, 0.f, 0.f);
        for (int smp = 0; smp < spp; ++smp, src += npx * 3) acc = acc + mk(src[0], src[1], src[2]);
        hdr.out[3 * i] = acc.x / (float)spp;
        hdr.out[3 * i + 1] = acc.y / (float)spp;
        hdr.out[3 * i + 2] = acc.z / (float)spp;
    }
}

// ---------------------------------------------------------------------------
// Wavefront engine.  A path lives in a slot; each iteration runs
//   closest (extend + next-event BSDF rays) -> any (shadow segments) ->
//   shade2 (next-event resolution: L += T * NEE) -> shade1 (the bounce at the
//   extended hit: miss/light tests, T <- T', receiver, next-event setup,
//   continuation sampling) -> generate (free slots take new paths).
// The arithmetic is trace_path's, split at its traversal calls, so every path
// gives the same bits; the kernels are small enough to run at high occupancy.
// Queue appends are wave-aggregated: one atomic per wave, ranks by ballot
// prefix.

enum : int { kWfActive = 1, kWfLo1 = 2, kWfLo2 = 4, kWfFinishAfterNee = 8 };
enum : int { kQExt = 0, kQNee = 1, kQAny = 2, kQRes = 3 };

__device__ __forceinline__ int wave_lane() { return (int)(threadIdx.x & 63); }

// every lane of the wave calls this; returns the slot of `pred` lanes in the queue
__device__ __forceinline__ int wave_append(int* counter, bool pred) {
    unsigned long long mask = __ballot(pred);
    if (mask == 0ull) return -1;
    int lane = wave_lane();
    int leader = __ffsll((long long)mask) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, __popcll(mask));
    base = __shfl(base, leader);
    int rank = __popcll(mask & ((1ull << lane) - 1ull));
    return pred ? base + rank : -1;
}

__device__ __forceinline__ void wave_sub_active(unsigned long long* active, bool pred) {
    unsigned long long mask = __ballot(pred);
    if (mask != 0ull && wave_lane() == __ffsll((long long)mask) - 1)
        atomicAdd(active, (unsigned long long)(-(long long)__popcll(mask)));
}

__device__ __forceinline__ v3 ld_p(const float* a, int64_t P, int64_t i) { return mk(a[i], a[P + i], a[2 * P + i]); }
__device__ __forceinline__ void st_p(float* a, int64_t P, int64_t i, v3 v) {
    a[i] = v.x; a[P + i] = v.y; a[2 * P + i] = v.z;
}

// a finished path: its radiance into the per-sample buffer, the slot freed
__device__ __forceinline__ void wf_finish(const WfView& v, const WfJobs& jh, int64_t slot) {
    const JobGpu& J = jh.jobs[v.job[slot] & (jh.job_cap - 1)];
    int64_t p = v.pix[slot];
    int smp = v.smp[slot];
    float* dst = jh.samples + J.sample_off + ((int64_t)smp * J.pp.w * J.pp.h + p) * 3;
    v3 L = ld_p(v.L, v.P, slot);
    dst[0] = L.x; dst[1] = L.y; dst[2] = L.z;
    v.flags[slot] = 0;
}

// finished paths per job, one atomic per distinct job in the wave (lanes of a
// wave mostly share one job); every lane of the wave calls this
__device__ __forceinline__ void wave_count_done(const WfJobs& jh, int job, bool fin) {
    if (!jh.job_done) return;
    unsigned long long rem = __ballot(fin);
    const int lane = wave_lane();
    while (rem != 0ull) {
        int leader = __ffsll((long long)rem) - 1;
        int lj = __shfl(job, leader);
        unsigned long long m = __ballot(fin && job == lj);
        if (lane == leader) atomicAdd(jh.job_done + (lj & (jh.job_cap - 1)), (unsigned int)__popcll(m));
        rem &= ~m;
    }
}

__global__ void k_wf_begin(WfView v, int out) {
    if (threadIdx.x < 4) v.counts[out * 4 + threadIdx.x] = 0;
    if (threadIdx.x == 4) v.free_counts[out] = 0;
}

__global__ __launch_bounds__(256) void k_wf_closest(SceneView sv, WfView v, const WfJobs* __restrict__ jhp, int in) {
    extern __shared__ int s_stack[];
    const WfJobs jh = *jhp;
    AnyStack astk = any_stack(sv, s_stack);
    TravStats st;
    unsigned long long nclosest = 0;
    const int n1 = v.counts[in * 4 + kQExt], n2 = v.counts[in * 4 + kQNee];
    const int64_t n = (int64_t)n1 + n2, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        bool nee = i >= n1;
        int64_t slot = nee ? v.q[in][kQNee][i - n1] : v.q[in][kQExt][i];
        float tmin = jh.jobs[v.job[slot] & (jh.job_cap - 1)].pp.tmin;
        Ray r = nee ? ray_from_od(ld_p(v.np, v.P, slot), ld_p(v.nd, v.P, slot))
                    : ray_from_od(ld_p(v.ro, v.P, slot), ld_p(v.rd, v.P, slot));
        float t = 0.f, b = 0.f, g = 0.f;
        nclosest++;
        int tri = closest_hit4(sv, r, tmin, 1e6f, astk, t, b, g, st);
        float4 h = make_float4(t, b, g, __int_as_float(tri));
        if (nee) v.nhit[slot] = h;
        else v.hit[slot] = h;
    }
    if (jh.work) add_ray_counts(jh.work, nclosest, 0ull);
}

__global__ __launch_bounds__(256) void k_wf_any(SceneView sv, WfView v, const WfJobs* __restrict__ jhp, int in) {
    extern __shared__ int s_stack[];
    const WfJobs jh = *jhp;
    AnyStack astk = any_stack(sv, s_stack);
    TravStats st;
    unsigned long long nshadow = 0;
    const int64_t n = v.counts[in * 4 + kQAny], stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        int64_t slot = v.q[in][kQAny][i];
        float tmin = jh.jobs[v.job[slot] & (jh.job_cap - 1)].pp.tmin;
        nshadow++;
        v.blocked[slot] = segment_blocked(sv, ld_p(v.np, v.P, slot), ld_p(v.nq, v.P, slot), tmin, astk, st) ? 1 : 0;
    }
    if (jh.work) add_ray_counts(jh.work, 0ull, nshadow);
}

// next-event resolution (the tail of next_event after its two traversals)
__global__ __launch_bounds__(256) void k_wf_shade2(SceneView sv, WfView v, const WfJobs* __restrict__ jhp, int in) {
    const WfJobs jh = *jhp;
    const int64_t n = v.counts[in * 4 + kQRes], stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x; i0 < n; i0 += stride) {
        int64_t i = i0 + threadIdx.x;
        bool fin = false;
        int64_t slot = 0;
        if (i < n) {
            slot = v.q[in][kQRes][i];
            int fl = v.flags[slot];
            int k = v.nk[slot];
            v3 nee = mk(0.f, 0.f, 0.f);
            if (k >= 0) {
                LightGpu Lk = sv.lights[k];
                v3 Lo = mk(0.f, 0.f, 0.f);
                if ((fl & kWfLo1) && !v.blocked[slot]) Lo = Lo + ld_p(v.lo1, v.P, slot);
                if (fl & kWfLo2) {
                    float4 h = v.nhit[slot];
                    int tri = __float_as_int(h.w);
                    if (tri >= 0) {
                        float4 nf = sv.tri_n[tri];
                        int4 sf = sv.surf[__float_as_int(nf.w)];
                        if (sf.y == k) {
                            v3 p = ld_p(v.np, v.P, slot), wi = ld_p(v.nwi, v.P, slot);
                            float t = h.x, tb = h.y, tg = h.z;
                            v3 qh = p + t * wi;
                            v3 nn = mk(nf.x, nf.y, nf.z);
                            v3 qhn = unit((nn + (nn - nn) * tb) + (nn - nn) * tg);
                            if (sf.z & RZK_SURF_FLIPPED) qhn = -qhn;
                            v3 Le = light_query(Lk, qhn, -wi);
                            float pa = 0.f;
                            for (int qq = 0; qq < Lk.count; ++qq) {
                                const LightTri& lt = sv.ltris[Lk.first + qq];
                                if (lt.tri == tri) { pa = lt.ratio * (1.0f / lt.n.w); break; }
                            }
                            v3 pq = qh - p;
                            float lpdf = pa * dot(pq, pq) / fabsf(dot(unit(p - qh), qhn));
                            float dpdf = v.ndpdf[slot];
                            float w2 = balance(dpdf, lpdf);
                            v3 brdf = ld_p(v.nbr, v.P, slot);
                            float cosv = v.ncos[slot];
                            Lo = Lo + (((w2 * cosv) / dpdf) * Le) * brdf;
                        }
                    }
                }
                nee = mk(Lo.x / Lk.sel_pdf, Lo.y / Lk.sel_pdf, Lo.z / Lk.sel_pdf);
            }
            v3 T = ld_p(v.T, v.P, slot);
            st_p(v.L, v.P, slot, ld_p(v.L, v.P, slot) + T * nee);
            if (fl & kWfFinishAfterNee) {
                wf_finish(v, jh, slot);
                fin = true;
            } else {
                v.flags[slot] = fl & ~(kWfLo1 | kWfLo2);
            }
        }
        int at = wave_append(v.free_counts + (in ^ 1), fin);
        if (fin) v.freeq[in ^ 1][at] = (int)slot;
        wave_count_done(jh, fin ? v.job[slot] : 0, fin);
        wave_sub_active(v.ctrl + 1, fin);
    }
}

// the bounce at the extended hit (the body of trace_path's loop)
__global__ __launch_bounds__(256) void k_wf_shade1(SceneView sv, WfView v, const WfJobs* __restrict__ jhp, int in) {
    const WfJobs jh = *jhp;
    const int out = in ^ 1;
    const int64_t n = v.counts[in * 4 + kQExt], stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x; i0 < n; i0 += stride) {
        int64_t i = i0 + threadIdx.x;
        bool e_ext = false, e_nee = false, e_any = false, e_res = false, fin = false;
        int64_t slot = 0;
        if (i < n) {
            slot = v.q[in][kQExt][i];
            const JobGpu& J = jh.jobs[v.job[slot] & (jh.job_cap - 1)];
            const PtParams pp = J.pp;
            const int bounce = v.bounce[slot];
            float4 h = v.hit[slot];
            int tri = __float_as_int(h.w);
            bool done = tri < 0;
            if (!done && bounce > 1) {
                int4 sf = sv.surf[__float_as_int(sv.tri_n[tri].w)];
                if (sf.y >= 0) done = true;                 // light: counted by the light sample
                else st_p(v.T, v.P, slot, ld_p(v.Tc, v.P, slot));
            }
            if (!done && bounce > pp.max_bounce) done = true;
            int fl = kWfActive;
            if (!done) {
                Ray ray = ray_from_od(ld_p(v.ro, v.P, slot), ld_p(v.rd, v.P, slot));
                float rc[RZK_RECV_FLOATS];
                int m, lidx;
                fill_receiver(sv, ray, h.x, h.y, h.z, tri, rc, m, lidx);
                v3 T = ld_p(v.T, v.P, slot);
                if (lidx >= 0) {
                    if (bounce == 1 && pp.direct)
                        st_p(v.L, v.P, slot, ld_p(v.L, v.P, slot) + T * light_query(sv.lights[lidx], ld3(rc + 6),
                                                                                   ld3(rc + 9)));
                    done = true;
                } else if (m < 0) {
                    done = true;
                } else {
                    int64_t p = v.pix[slot];
                    int x = pp.x0 + (int)(p % pp.w), y = pp.y0 + (int)(p / pp.w);
                    uint32_t pixel = (uint32_t)(y * pp.width + x);
                    PathRng rng = path_rng(pp.seed, pixel, (uint32_t)v.smp[slot], (uint32_t)bounce);
                    bool brdf_sampling = u24(rng.u[0]) < 0.5f;
                    const float* mat = sv.mat + RZK_MAT_FLOATS * m;
                    Onb b; b.u = ld3(rc + 12); b.v = ld3(rc + 15); b.n = ld3(rc + 18);
                    v3 sn = ld3(rc + 6), wo = ld3(rc + 9), p3 = ld3(rc);
                    bool nee = bounce > 1 || pp.direct;
                    if (nee) {
                        e_res = true;
                        // next_event up to its traversals
                        float ul = u24(rng.u[1]);
                        int k = -1;
                        for (int li = 0; li < sv.nlights; ++li)
                            if (ul < sv.lights[li].sel_cdf) { k = li; break; }
                        if (k >= 0 && sv.lights[k].count <= 0) k = -1;
                        v.nk[slot] = k;
                        if (k >= 0) {
                            LightGpu Lk = sv.lights[k];
                            v3 pn = sn;
                            st_p(v.np, v.P, slot, p3);
                            float ut = u24(rng.u[2]);
                            int ti = Lk.first + Lk.count - 1;
                            for (int qi = 0; qi < Lk.count; ++qi)
                                if (ut <= sv.ltris[Lk.first + qi].cum) { ti = Lk.first + qi; break; }
                            LightTri Tl = sv.ltris[ti];
                            float tmp = sqrtf(1.0f - u24(rng.u[3]));
                            float bb = 1.0f - tmp;
                            float cc = u24(rng.u[4]) * tmp;
                            float aa = 1.0f - bb - cc;
                            v3 q = (aa * mk(Tl.p0.x, Tl.p0.y, Tl.p0.z) + bb * mk(Tl.p1.x, Tl.p1.y, Tl.p1.z)) +
                                   cc * mk(Tl.p2.x, Tl.p2.y, Tl.p2.z);
                            v3 qn = mk(Tl.n.x, Tl.n.y, Tl.n.z);
                            float lpdf = Tl.ratio * (1.0f / Tl.n.w);
                            if (lpdf > 0.0f) {
                                e_any = true;
                                st_p(v.nq, v.P, slot, q);
                                v3 wi = unit(q - p3);
                                v3 Le = light_query(Lk, qn, -wi);
                                float G = form_factor_abs(p3, pn, q, qn);
                                if (G > 0.0f) {
                                    v3 brdf = bsdf_eval(mat, b, wo, wi);
                                    float dpdf = bsdf_pdf(mat, b, wo, wi);
                                    v3 pq = q - p3;
                                    float w1 = balance(lpdf * dot(pq, pq) / fabsf(dot(unit(p3 - q), qn)), dpdf);
                                    st_p(v.lo1, v.P, slot, (((w1 * G) / lpdf) * Le) * brdf);
                                    fl |= kWfLo1;
                                }
                            }
                            v3 wi2;
                            float dpdf2;
                            bsdf_sample(mat, b, wo, &rng.u[5], wi2, dpdf2);
                            if (dpdf2 > 0.0f) {
                                e_nee = true;
                                Ray r2 = ray_ctor(p3, wi2);
                                st_p(v.nd, v.P, slot, r2.d);
                                st_p(v.nwi, v.P, slot, wi2);
                                v.ndpdf[slot] = dpdf2;
                                st_p(v.nbr, v.P, slot, bsdf_eval(mat, b, wo, wi2));
                                v.ncos[slot] = fabsf(dot(wi2, pn));
                                fl |= kWfLo2;
                            }
                        }
                    }
                    bool cont = false;
                    if (bounce < pp.max_bounce) {
                        int slice = pixel_slice(pp, J.gv, x, y);
                        int gslice = bounce == 1 ? slice : -1;
                        v3 wi, brdf;
                        float pdf;
                        if (brdf_sampling) brdf = bsdf_sample(mat, b, wo, &rng.u[8], wi, pdf);
                        else brdf = guide_sample(J.gv, gslice, pp.res, mat, b, sn, wo, &rng.u[8], wi, pdf);
                        if (pdf > 0.0f) {
                            cont = true;
                            Ray nr = ray_ctor(p3, wi);
                            st_p(v.ro, v.P, slot, nr.o);
                            st_p(v.rd, v.P, slot, nr.d);
                            float pdf_other = brdf_sampling ? guide_pdf(J.gv, gslice, pp.res, b, sn, wi)
                                                            : bsdf_pdf(mat, b, wo, wi);
                            float weight = pdf / (pdf + pdf_other);
                            float cosv = fabsf(dot(wi, sn));
                            float s = ((weight * cosv) / pdf) * 2.0f;
                            st_p(v.Tc, v.P, slot, (T * brdf) * s);
                            v.bounce[slot] = bounce + 1;
                            e_ext = true;
                        }
                    }
                    if (!cont) {
                        if (nee) fl |= kWfFinishAfterNee;
                        else done = true;
                    }
                }
            }
            if (done) {
                wf_finish(v, jh, slot);
                fin = true;
            } else {
                v.flags[slot] = fl;
            }
        }
        int at;
        if ((at = wave_append(v.counts + out * 4 + kQExt, e_ext)) >= 0) v.q[out][kQExt][at] = (int)slot;
        if ((at = wave_append(v.counts + out * 4 + kQNee, e_nee)) >= 0) v.q[out][kQNee][at] = (int)slot;
        if ((at = wave_append(v.counts + out * 4 + kQAny, e_any)) >= 0) v.q[out][kQAny][at] = (int)slot;
        if ((at = wave_append(v.counts + out * 4 + kQRes, e_res)) >= 0) v.q[out][kQRes][at] = (int)slot;
        if ((at = wave_append(v.free_counts + out, fin)) >= 0) v.freeq[out][at] = (int)slot;
        wave_count_done(jh, fin ? v.job[slot] : 0, fin);
        wave_sub_active(v.ctrl + 1, fin);
    }
}

// slots freed in the previous iteration take the next paths (sample-major
// within a job); primary rays go to the extend queue.  Slots left without a
// path stay free for good: the paths are exhausted.
__global__ __launch_bounds__(256) void k_wf_generate(WfView v, const WfJobs* __restrict__ jhp, int in) {
    const WfJobs jh = *jhp;
    const int out = in ^ 1;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t mask_j = jh.job_cap - 1;
    const int64_t nfree = v.free_counts[in];
    for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x; i0 < nfree; i0 += stride) {
        int64_t fi = i0 + threadIdx.x;
        bool free_ = fi < nfree;
        int64_t slot = free_ ? v.freeq[in][fi] : 0;
        unsigned long long mask = __ballot(free_);
        if (mask == 0ull) continue;
        int lane = wave_lane();
        int leader = __ffsll((long long)mask) - 1;
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(v.ctrl, (unsigned long long)__popcll(mask));
        base = __shfl(base, leader);
        unsigned long long g = base + (unsigned long long)__popcll(mask & ((1ull << lane) - 1ull));
        bool take = free_ && g < (unsigned long long)jh.total_paths && jh.n_jobs > jh.job_lo;
        if (take) {
            int64_t lo = jh.job_lo, hi = jh.n_jobs - 1;
            while (lo < hi) {
                int64_t mid = (lo + hi + 1) >> 1;
                if ((unsigned long long)jh.jobs[mid & mask_j].item_first <= g) lo = mid;
                else hi = mid - 1;
            }
            const JobGpu& J = jh.jobs[lo & mask_j];
            const PtParams pp = J.pp;
            int64_t npx = (int64_t)pp.w * pp.h;
            int64_t local = (int64_t)(g - (unsigned long long)J.item_first);
            int smp = (int)(local / npx);
            int64_t p = local - (int64_t)smp * npx;
            int x = pp.x0 + (int)(p % pp.w), y = pp.y0 + (int)(p / pp.w);
            uint32_t pixel = (uint32_t)(y * pp.width + x);
            PathRng r0 = path_rng(pp.seed, pixel, (uint32_t)smp, 0u);
            float jx = u24(r0.u[0]), jy = u24(r0.u[1]);
            int strata = J.strata;
            if (strata > 0) {
                jx = ((float)(smp % strata) + jx) / (float)strata;
                jy = ((float)(smp / strata) + jy) / (float)strata;
            }
            Ray ray = camera_sample_ray(J.cam.v, (float)x + jx, (float)y + jy);
            v.job[slot] = (int)lo;
            v.pix[slot] = (int)p;
            v.smp[slot] = smp;
            v.bounce[slot] = 1;
            v.flags[slot] = kWfActive;
            st_p(v.ro, v.P, slot, ray.o);
            st_p(v.rd, v.P, slot, ray.d);
            st_p(v.T, v.P, slot, mk(1.f, 1.f, 1.f));
            st_p(v.L, v.P, slot, mk(0.f, 0.f, 0.f));
        }
        int at = wave_append(v.counts + out * 4 + kQExt, take);
        if (take) v.q[out][kQExt][at] = (int)slot;
        // a free slot without a path now waits for the next iteration's paths
        bool wait = free_ && !take;
        at = wave_append(v.free_counts + out, wait);
        if (wait) v.freeq[out][at] = (int)slot;
        unsigned long long tm = __ballot(take);
        if (tm != 0ull && lane == __ffsll((long long)tm) - 1) atomicAdd(v.ctrl + 1, (unsigned long long)__popcll(tm));
    }
}

// paths handed out past the table's end were not taken: hand them out again
__global__ void k_wf_clamp(WfView v, const WfJobs* __restrict__ jhp) {
    if (threadIdx.x == 0 && v.ctrl[0] > (unsigned long long)jhp->total_paths)
        v.ctrl[0] = (unsigned long long)jhp->total_paths;
}

// a finished job: per-pixel mean of its samples, summed in sample order
__global__ void k_job_reduce(const JobGpu* __restrict__ J, const float* __restrict__ samples,
                             float* __restrict__ out) {
    const int64_t npx = (int64_t)J->pp.w * J->pp.h;
    const int spp = J->pp.spp;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npx; p += (int64_t)gridDim.x * blockDim.x) {
        const float* src = samples + J->sample_off + p * 3;
        v3 acc = mk(0.f, 0.f, 0.f);
        for (int smp = 0; smp < spp; ++smp, src += npx * 3) acc = acc + mk(src[0], src[1], src[2]);
        out[3 * p] = acc.x / (float)spp;
        out[3 * p + 1] = acc.y / (float)spp;
        out[3 * p + 2] = acc.z / (float)spp;
    }
}

}  // namespace rzk

using namespace rzk;

void launch_wf_iteration(hipStream_t stream, const SceneView& sv, const WfView& v, const WfJobs* d_jh, int in,
                         int grid, int block) {
    size_t lds = sizeof(int) * (size_t)block * (size_t)kLdsStack;
    hipLaunchKernelGGL(k_wf_begin, dim3(1), dim3(64), 0, stream, v, in ^ 1);
    hipLaunchKernelGGL(k_wf_closest, dim3(grid), dim3(block), lds, stream, sv, v, d_jh, in);
    hipLaunchKernelGGL(k_wf_any, dim3(grid), dim3(block), lds, stream, sv, v, d_jh, in);
    hipLaunchKernelGGL(k_wf_shade2, dim3(grid), dim3(block), 0, stream, sv, v, d_jh, in);
    hipLaunchKernelGGL(k_wf_shade1, dim3(grid), dim3(block), 0, stream, sv, v, d_jh, in);
    hipLaunchKernelGGL(k_wf_generate, dim3(grid), dim3(block), 0, stream, v, d_jh, in);
    hipLaunchKernelGGL(k_wf_clamp, dim3(1), dim3(64), 0, stream, v, d_jh);
}

void launch_job_reduce(hipStream_t stream, const JobGpu* d_job, const float* samples, float* out) {
    hipLaunchKernelGGL(k_job_reduce, dim3(512), dim3(256), 0, stream, d_job, samples, out);
}

void launch_batch_reduce(hipStream_t stream, const BatchHeader* d_hdr, int grid) {
    hipLaunchKernelGGL(k_batch_reduce, dim3(grid), dim3(256), 0, stream, d_hdr);
}


void launch_batch_items(hipStream_t stream, const SceneView& sv, const BatchHeader* d_hdr, int grid, int block,
                        int items_per_dequeue) {
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, stream, d_hdr);
    size_t lds = sizeof(int) * (size_t)block * ((size_t)kPtLdsStack + kPtCold);
    if (items_per_dequeue == 1) hipLaunchKernelGGL(k_batch_pt_one, dim3(grid), dim3(block), lds, stream, sv, d_hdr);
    else hipLaunchKernelGGL(k_batch_pt, dim3(grid), dim3(block), lds, stream, sv, d_hdr);
}

void launch_batch(hipStream_t stream, const SceneView& sv, const BatchHeader* d_hdr, int grid, int block,
                  int items_per_dequeue, int reduce_grid) {
    launch_batch_items(stream, sv, d_hdr, grid, block, items_per_dequeue);
    hipLaunchKernelGGL(k_batch_reduce, dim3(reduce_grid), dim3(256), 0, stream, d_hdr);
}

void launch_guided_pt_rays(rzk_ctx* ctx, const SceneView& sv_in, const PtParams& pp, const GuideView& gv,
                           const float* d_rays, const int* d_pix, const int* d_smp, int64_t n, float* d_out) {
    if (n <= 0) return;
    SceneView sv = sv_in;
    const int block = kPtBlock;
    unsigned grid = (unsigned)((n + block - 1) / block);
    bind_spill(ctx, sv, (int64_t)grid * block, 2 * sv.stack4, kPtLdsStack);
    size_t lds = sizeof(int) * block * ((size_t)kPtLdsStack + kPtCold);
    hipLaunchKernelGGL(k_guided_pt_rays, dim3(grid), dim3(block), lds, ctx->stream, sv, pp, gv, d_rays, d_pix, d_smp,
                       n, d_out);
    RZK_HIP(hipGetLastError());
}

void launch_guided_pt(rzk_ctx* ctx, const SceneView& sv_in, const float* cam, const PtParams& pp, const GuideView& gv,
                      float* d_out, unsigned long long* d_work) {
    int64_t n = (int64_t)pp.w * pp.h;
    if (n <= 0) return;
    SceneView sv = sv_in;
    const int block = kPtBlock;
    unsigned grid = (unsigned)((n + block - 1) / block);
    bind_spill(ctx, sv, (int64_t)grid * block, 2 * sv.stack4, kPtLdsStack);
    CamView cv;
    for (int i = 0; i < RZK_CAM_FLOATS; ++i) cv.v[i] = cam[i];
    size_t lds = sizeof(int) * block * ((size_t)kPtLdsStack + kPtCold);
    hipLaunchKernelGGL(k_guided_pt, dim3(grid), dim3(block), lds, ctx->stream, sv, cv, pp, gv, d_out, d_work);
    RZK_HIP(hipGetLastError());
}
