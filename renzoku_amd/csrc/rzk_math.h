// rzk_math.h — fp32 geometry and shading math shared by the host BVH builder
// and the gfx950 kernels.
//
// Every expression keeps Renzoku's evaluation order (fp32, left to right, no
// fused multiply-add: the reference is built -O2 for x86-64 without -mfma,
// reference Makefile:10-11).  This translation unit family is compiled with
// -ffp-contract=off and the pragma below, and with correctly rounded fp32
// division and square root, so that ray/box/triangle results are bit-identical
// to the reference's.
#pragma once
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <stdint.h>

#define RZK_HD __host__ __device__ __forceinline__

namespace rzk {

// common.h:38-55
constexpr float kPi = 3.14159265358979f;
constexpr float kInvPi = 0.31830988618379f;
constexpr float kTwoPi = 6.28318530717959f;
constexpr float kInv2Pi = 0.15915494309190f;
constexpr float kZeroEps = 1e-6f;
constexpr float kOnbEps = 0.01f;

struct v3 { float x, y, z; };

RZK_HD v3 mk(float x, float y, float z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
RZK_HD v3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }
RZK_HD void st3(float* p, v3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }
RZK_HD v3 operator+(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
RZK_HD v3 operator-(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
RZK_HD v3 operator-(v3 a) { return mk(-a.x, -a.y, -a.z); }
RZK_HD v3 operator*(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }   // componentwise
RZK_HD v3 operator/(v3 a, v3 b) { return mk(a.x / b.x, a.y / b.y, a.z / b.z); }
RZK_HD v3 operator*(v3 a, float t) { return mk(a.x * t, a.y * t, a.z * t); }
RZK_HD v3 operator*(float t, v3 a) { return mk(a.x * t, a.y * t, a.z * t); }      // vec3_cpu.h:257: v.e*t
RZK_HD float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RZK_HD v3 cross(v3 a, v3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
RZK_HD float length(v3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
// vec3_cpu.h:406-413: multiply by the reciprocal of the length
RZK_HD v3 unit(v3 a) { float inv = 1.0f / length(a); return mk(a.x * inv, a.y * inv, a.z * inv); }
RZK_HD bool isnan3(v3 a) { return a.x != a.x || a.y != a.y || a.z != a.z; }
// std::min / std::max semantics (NaN handling and signed zeros follow the operand order)
RZK_HD float smin(float a, float b) { return (b < a) ? b : a; }
RZK_HD float smax(float a, float b) { return (a < b) ? b : a; }
RZK_HD v3 vmin(v3 a, v3 b) { v3 t = a; if (b.x < a.x) t.x = b.x; if (b.y < a.y) t.y = b.y; if (b.z < a.z) t.z = b.z; return t; }
RZK_HD v3 vmax(v3 a, v3 b) { v3 t = a; if (b.x > a.x) t.x = b.x; if (b.y > a.y) t.y = b.y; if (b.z > a.z) t.z = b.z; return t; }

// Ray with the members Renzoku precomputes (ray.h:17-36).  `d` is the
// constructor's normalised direction; inv_d clamps |d_i| < 1e-6 to 1e-6.
struct Ray { v3 o, d, inv_d, org_inv_d; };

RZK_HD Ray ray_from_od(v3 o, v3 d) {
    Ray r;
    r.o = o;
    r.d = d;
    v3 dd = mk(fabsf(d.x) < kZeroEps ? kZeroEps : d.x,
               fabsf(d.y) < kZeroEps ? kZeroEps : d.y,
               fabsf(d.z) < kZeroEps ? kZeroEps : d.z);
    r.inv_d = mk(1.f, 1.f, 1.f) / dd;
    r.org_inv_d = (-o) * r.inv_d;
    return r;
}

// Ray::Ray(o, d): normalise, NaN -> 0 (ray.h:17-21)
RZK_HD Ray ray_ctor(v3 o, v3 d_in) {
    v3 d = unit(d_in);
    if (isnan3(d)) d = mk(0.f, 0.f, 0.f);
    return ray_from_od(o, d);
}

// BoundingBox::hit, boundingbox.h:238-280 (mul then add, never fused).  The
// slab values are never NaN (inv_d is finite and non-zero, boxes and origins
// are finite), so fminf/fmaxf (v_min3/v_max3 on gfx950) pick the same values
// as the reference's std::min/std::max up to the sign of a zero, which no
// comparison below can see: the test's outcome is identical.  The reference
// also requires tmax >= tminbox and tmaxbox >= tmin; with tminbox >= tmin and
// tmaxbox <= tmax (no NaN), both follow from tminbox <= tmaxbox, so the one
// comparison decides.
RZK_HD bool box_hit(v3 bmin, v3 bmax, const Ray& r, float tmin, float tmax) {
    float x0 = bmin.x * r.inv_d.x + r.org_inv_d.x, x1 = bmax.x * r.inv_d.x + r.org_inv_d.x;
    float y0 = bmin.y * r.inv_d.y + r.org_inv_d.y, y1 = bmax.y * r.inv_d.y + r.org_inv_d.y;
    float z0 = bmin.z * r.inv_d.z + r.org_inv_d.z, z1 = bmax.z * r.inv_d.z + r.org_inv_d.z;
    float tminbox = fmaxf(fmaxf(tmin, fminf(x0, x1)), fmaxf(fminf(y0, y1), fminf(z0, z1)));
    float tmaxbox = fminf(fminf(tmax, fmaxf(x0, x1)), fminf(fmaxf(y0, y1), fmaxf(z0, z1)));
    return tminbox <= tmaxbox;
}

// box_hit that also returns the entry distance (max of tmin and the slab entries)
RZK_HD bool box_enter(v3 bmin, v3 bmax, const Ray& r, float tmin, float tmax, float& tenter) {
    float x0 = bmin.x * r.inv_d.x + r.org_inv_d.x, x1 = bmax.x * r.inv_d.x + r.org_inv_d.x;
    float y0 = bmin.y * r.inv_d.y + r.org_inv_d.y, y1 = bmax.y * r.inv_d.y + r.org_inv_d.y;
    float z0 = bmin.z * r.inv_d.z + r.org_inv_d.z, z1 = bmax.z * r.inv_d.z + r.org_inv_d.z;
    float tminbox = fmaxf(fmaxf(tmin, fminf(x0, x1)), fmaxf(fminf(y0, y1), fminf(z0, z1)));
    float tmaxbox = fminf(fminf(tmax, fmaxf(x0, x1)), fminf(fmaxf(y0, y1), fmaxf(z0, z1)));
    tenter = tminbox;
    return tminbox <= tmaxbox;     // the other two conditions follow (box_hit)
}

// Slab test of box_hit when the sign of every direction component is known
// (nx: inv_d.x < 0, ...): the entry plane of each slab is then lo (positive
// component) or hi (negative), so the per-axis min/max drop out.  The values
// are box_hit's bit for bit: with inv_d.x > 0 and lo <= hi, lo * inv + org <=
// hi * inv + org in fp32 (rounded mul and add are monotone), so fminf of the
// two slab values is the lo one; with inv_d.x < 0 it is the hi one.
// The slab test on pre-selected entry corner `en` and exit corner `ex`
// (per axis: lo and hi for a positive component, hi and lo for a negative one).
RZK_HD bool box_hit_planes(v3 en, v3 ex, const Ray& r, float tmin, float tmax) {
    float xn = en.x * r.inv_d.x + r.org_inv_d.x, xf = ex.x * r.inv_d.x + r.org_inv_d.x;
    float yn = en.y * r.inv_d.y + r.org_inv_d.y, yf = ex.y * r.inv_d.y + r.org_inv_d.y;
    float zn = en.z * r.inv_d.z + r.org_inv_d.z, zf = ex.z * r.inv_d.z + r.org_inv_d.z;
    float tminbox = fmaxf(fmaxf(tmin, xn), fmaxf(yn, zn));
    float tmaxbox = fminf(fminf(tmax, xf), fminf(yf, zf));
    return tminbox <= tmaxbox;
}

RZK_HD bool box_hit_signed(v3 bmin, v3 bmax, bool nx, bool ny, bool nz, const Ray& r, float tmin, float tmax) {
    return box_hit_planes(mk(nx ? bmax.x : bmin.x, ny ? bmax.y : bmin.y, nz ? bmax.z : bmin.z),
                          mk(nx ? bmin.x : bmax.x, ny ? bmin.y : bmax.y, nz ? bmin.z : bmax.z), r, tmin, tmax);
}

// Triangle::hit, triangle.cc:158-232, on the precomputed edges p01 = p0 - p1,
// p02 = p0 - p2 (the same subtractions the reference performs per test).
RZK_HD bool tri_hit(v3 p0, v3 p01, v3 p02, const Ray& r, float tmin, float tmax,
                    float& t_out, float& b_out, float& g_out) {
    float A = p01.x, B = p01.y, C = p01.z;
    float D = p02.x, E = p02.y, F = p02.z;
    float G = r.d.x, H = r.d.y, I = r.d.z;
    float J = p0.x - r.o.x, K = p0.y - r.o.y, L = p0.z - r.o.z;
    float EIHF = E * I - H * F;
    float GFDI = G * F - D * I;
    float DHEG = D * H - E * G;
    float denom = (A * EIHF + B * GFDI + C * DHEG);
    if (-kZeroEps < denom && denom < kZeroEps) return false;
    float beta = (J * EIHF + K * GFDI + L * DHEG) / denom;
    if (beta < 0.0f || beta > 1.0f) return false;
    float AKJB = A * K - J * B;
    float JCAL = J * C - A * L;
    float BLKC = B * L - K * C;
    float gamma = (I * AKJB + H * JCAL + G * BLKC) / denom;
    if (gamma < 0.0f || beta + gamma > 1.0f) return false;
    float t = -(F * AKJB + E * JCAL + D * BLKC) / denom;
    if (t >= tmin && t <= tmax) { t_out = t; b_out = beta; g_out = gamma; return true; }
    return false;
}

// Orthonormal basis (onb.cc): u, v, n rows
struct Onb { v3 u, v, n; };

RZK_HD Onb onb_from_n(v3 n) {                       // onb.cc:42-54
    Onb o;
    v3 nn = unit(n);
    v3 uu = cross(nn, mk(1.f, 0.f, 0.f));
    if (length(uu) < kOnbEps) uu = cross(nn, mk(0.f, 1.f, 0.f));
    v3 vv = cross(nn, uu);
    o.u = unit(uu); o.v = unit(vv); o.n = nn;
    return o;
}

RZK_HD Onb onb_from_nu(v3 n, v3 u_in) {             // onb.cc:99-109
    v3 c = unit(cross(n, u_in));
    if (isnan3(c)) return onb_from_n(n);
    Onb o;
    o.n = unit(n); o.v = c; o.u = cross(c, o.n);
    return o;
}

RZK_HD v3 world_to_local(const Onb& b, v3 w) { return mk(dot(b.u, w), dot(b.v, w), dot(b.n, w)); }
RZK_HD v3 local_to_world(const Onb& b, v3 l) { return (l.x * b.u + l.y * b.v) + l.z * b.n; }

// Material record: 12 floats {type, kd[3], ks[3], p0, p1, pad[3]}
enum { kLambert = 0, kPhong = 1, kWard = 2 };

// Lambertian::eval (lambertian.cc:63-68), ModifiedPhong::eval (phong.cc:74-87),
// Ward::eval (ward.cc:104-116)
RZK_HD v3 bsdf_eval(const float* m, const Onb& b, v3 wo, v3 wi) {
    int type = (int)m[0];
    v3 kd = mk(m[1], m[2], m[3]);
    v3 ks = mk(m[4], m[5], m[6]);
    if (type == kWard) {
        float ax = m[7], ay = m[8];
        float ax2 = ax * ax, ay2 = ay * ay;
        float inv4 = 1.0f / (4.0f * kPi * ax * ay);
        v3 lwi = world_to_local(b, wi), lwo = world_to_local(b, wo);
        if (lwi.z <= kZeroEps || lwo.z <= kZeroEps) return mk(0.f, 0.f, 0.f);
        v3 wh = lwi + lwo;
        float expo = expf(-(wh.x * wh.x / ax2 + wh.y * wh.y / ay2) / (wh.z * wh.z));
        float sq = sqrtf(lwi.z * lwo.z);
        v3 spec = mk(ks.x * inv4 / sq, ks.y * inv4 / sq, ks.z * inv4 / sq) * expo;
        return kd * kInvPi + spec;
    }
    if (dot(wi, b.n) <= kZeroEps || dot(wo, b.n) <= kZeroEps) return mk(0.f, 0.f, 0.f);
    v3 brdf = kd * kInvPi;
    if (type == kPhong) {
        float ne = m[7];
        v3 wr = b.n * (2.0f * dot(wi, b.n)) - wi;
        float cos_alpha = dot(wo, wr);
        if (cos_alpha > 0.0f) brdf = brdf + ((ks * (ne + 2.0f)) * kInv2Pi) * powf(cos_alpha, ne);
    }
    return brdf;
}

// FormFactor::form_factor_abs, form_factor.h:11-20
RZK_HD float form_factor_abs(v3 p, v3 pn, v3 q, v3 qn) {
    v3 pq = q - p;
    float d1 = dot(pn, pq);
    float d2 = -dot(qn, pq);
    float dpq = dot(pq, pq);
    return fabsf((d1 * d2) / (dpq * dpq));
}

enum { kClampNone = 0, kClampFormFactor = 1, kClampReflectivity = 2, kClampRadiance = 3 };

RZK_HD float rgb_value(v3 c) { return sqrtf(c.x * c.x + c.y * c.y + c.z * c.z); }   // rgb.h:85

// RadianceTerms::eval_clamp_* (vpl_interface.h:50-123) dispatched as clamp_term (vpl.cc:214-230)
RZK_HD v3 clamp_term(int mode, float thr, v3 power, v3 fs, v3 fl, float G) {
    if (mode == kClampFormFactor) {
        if (G > thr) return ((fl * fs) * thr) * power;
        return ((fl * fs) * G) * power;
    }
    if (mode == kClampReflectivity) {
        v3 refl = (fs * G) * fl;
        float lum = rgb_value(refl);
        if (lum > thr) { float w = thr / lum; return (refl * w) * power; }
        return refl * power;
    }
    if (mode == kClampRadiance) {
        v3 rnc = ((power * G) * fs) * fl;
        float bright = rgb_value(rnc);
        if (bright > thr) { float w = thr / bright; return rnc * w; }
        return rnc;
    }
    return ((power * G) * fs) * fl;
}

// Morton code of a point already normalised to the scene box (morton.h:37-46)
RZK_HD uint32_t spread10(uint32_t v) {
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

RZK_HD uint32_t morton3d(float x, float y, float z) {
    x = smin(smax(x * 1024.0f, 0.0f), 1023.0f);
    y = smin(smax(y * 1024.0f, 0.0f), 1023.0f);
    z = smin(smax(z * 1024.0f, 0.0f), 1023.0f);
    return (spread10((uint32_t)x) << 2) | (spread10((uint32_t)y) << 1) | spread10((uint32_t)z);
}

// ---------------------------------------------------------------------------
// Counter-based random numbers for the path tracer: Philox4x32-10 (Salmon,
// Moraes, Dror, Shaw, SC'11).  Every (pixel, sample, bounce, slot) has its own
// numbers, so a path is the same whichever batch, wave or order runs it.
RZK_HD void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// Random::operator() (random.h:60-63): 24-bit float in [0, 1)
RZK_HD float u24(uint32_t r) { return (float)(r & 0x00ffffffu) * (1.0f / 16777216.0f); }
// Random::rand_exclude_01 (random.h:65-71): in (0, 1)
RZK_HD float u24_open(uint32_t r) { return (float)(r & 0x00ffffffu) * (1.0f / 16777216.0f) + (1.0f / 33554432.0f); }

// 12 numbers for one bounce of one path sample
struct PathRng { uint32_t u[12]; };
RZK_HD PathRng path_rng(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t bounce) {
    PathRng o;
    for (uint32_t b = 0; b < 3; ++b) {
        uint32_t c[4] = {sample, bounce, b, 0x52a7u};
        philox4x32_10(c, seed ^ 0x9a3c5e17u, pixel);
        o.u[4 * b] = c[0]; o.u[4 * b + 1] = c[1]; o.u[4 * b + 2] = c[2]; o.u[4 * b + 3] = c[3];
    }
    return o;
}

// cos and sin of 2*pi*u for u in [0, 1).  The turn is split into octants; the
// angle inside an octant (at most pi/4, reflected in odd octants) goes through
// Taylor polynomials of degree 9 (sin) and 10 (cos), whose truncation error is
// below 2e-9, so the results are within a few ulp of the C library's cos and
// sin of 2*pi*u.  Every operation is a correctly rounded fp32 add or multiply
// in a fixed order, so host and device give the same bits (no contraction).
RZK_HD void sincos_turn(float u, float& s, float& c) {
    const float x = u * 8.0f;
    int o = (int)x;
    if (o > 7) o = 7;
    if (o < 0) o = 0;
    const float r = x - (float)o;
    const float f = (o & 1) ? 1.0f - r : r;
    const float a = f * 0.785398163397448f;
    const float a2 = a * a;
    const float sa =
        a * (1.0f + a2 * (-0.166666667f + a2 * (0.00833333333f + a2 * (-0.000198412698f + a2 * 2.75573192e-6f))));
    const float ca = 1.0f + a2 * (-0.5f + a2 * (0.0416666667f + a2 * (-0.00138888889f +
                                                                     a2 * (2.48015873e-5f + a2 * -2.75573192e-7f))));
    switch (o) {
        case 0: s = sa; c = ca; break;
        case 1: s = ca; c = sa; break;
        case 2: s = ca; c = -sa; break;
        case 3: s = sa; c = -ca; break;
        case 4: s = -sa; c = -ca; break;
        case 5: s = -ca; c = -sa; break;
        case 6: s = -ca; c = sa; break;
        default: s = -sa; c = ca; break;
    }
}

// the 8 numbers of light sample `index` of a VPL generation keyed by `seed`
struct VplRng { uint32_t u[8]; };
RZK_HD VplRng vpl_rng(uint32_t seed, uint64_t index) {
    VplRng o;
    for (uint32_t b = 0; b < 2; ++b) {
        uint32_t c[4] = {(uint32_t)index, (uint32_t)(index >> 32), b, 0x5650u};
        philox4x32_10(c, seed ^ 0x3c6ef372u, 0x4c475650u);
        o.u[4 * b] = c[0]; o.u[4 * b + 1] = c[1]; o.u[4 * b + 2] = c[2]; o.u[4 * b + 3] = c[3];
    }
    return o;
}

// Rgb::max (rgb.h:42-47)
RZK_HD float rgb_max(v3 c) { float t = c.x; if (c.y > t) t = c.y; if (c.z > t) t = c.z; return t; }

// rho_d / rho_s of Phong and Ward (phong.cc:6-16, ward.cc:6-22)
RZK_HD void lobe_weights(const float* m, float& rho_d, float& rho_s) {
    rho_d = rgb_max(mk(m[1], m[2], m[3]));
    rho_s = rgb_max(mk(m[4], m[5], m[6]));
    float total = rho_d + rho_s;
    if (total > 0.0f) { rho_d /= total; rho_s /= total; }
}

// Transcendentals that pick a ray direction, from the mathematics in fp32
// add, multiply, divide and square root only (each correctly rounded, no
// contraction), in a fixed order, so that the host restatement of the same
// steps gives the same bits for every argument and traces the same rays.
// Within 3.5 ulp of the exact functions on the samplers' arguments.
//   xsincos  |x| <= 1e4: k = nearest integer of x * 2/pi, r = x - k pi/2 with
//            pi/2 in three parts of which k times each is exact (Cody-Waite),
//            Taylor polynomials of degree 9 (sin) and 10 (cos) on |r| <= pi/4,
//            the quadrant k mod 4 picking and signing the pair
//   xatan    |x| > 1 through pi/2 - atan(1/|x|), |x| > tan(pi/8) through
//            pi/4 + atan((|x| - 1) / (|x| + 1)), Taylor series to degree 21
//   xatan2   the quadrant from the signs, atan of the smaller over the larger
//            magnitude (the C library's signed-zero cases)
//   xacos    |x| <= 1/2: pi/2 - asin(x); else 2 asin(sqrt((1 - |x|) / 2)),
//            reflected for x < 0; asin's Taylor series to degree 21
//   xlog     x = m 2^e, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh((m - 1) / (m + 1))
//            to s^9, e ln 2 in two parts
//   xexp     k = nearest integer of t / ln 2, r = t - k ln 2 (two parts), Taylor
//            to r^8, times 2^k from the exponent bits
//   xpow     x >= 0: exp(y log x)
constexpr float kXmPio2_1 = 0x1.921ep+0f;      // pi/2 = kXmPio2_1 + kXmPio2_2 + kXmPio2_3 (+ 2^-58)
constexpr float kXmPio2_2 = 0x1.b544p-16f;
constexpr float kXmPio2_3 = 0x1.0b46p-34f;
constexpr float kXmPio2Lo = 0x1.b54442p-16f;   // pi/2 - kXmPio2_1
constexpr float kXmPio4Hi = 0x1.921ep-1f;
constexpr float kXmPio4Lo = 0x1.b54442p-17f;
constexpr float kXmPiHi = 0x1.921ep+1f;
constexpr float kXmPiLo = 0x1.b54442p-15f;
constexpr float kXmLn2Hi = 0x1.62e4p-1f;
constexpr float kXmLn2Lo = 0x1.7f7d1cp-20f;

RZK_HD int xm_round(float v) { return (int)(v + (v >= 0.0f ? 0.5f : -0.5f)); }

RZK_HD void xsincos(float x, float& s, float& c) {
    if (!(fabsf(x) <= 1e4f)) { s = c = (x - x) / (x - x); return; }
    const float kf = x * 0x1.45f306p-1f;
    const int k = xm_round(kf);
    const float fk = (float)k;
    const float r = ((x - fk * kXmPio2_1) - fk * kXmPio2_2) - fk * kXmPio2_3;
    const float r2 = r * r;
    const float sr = r + r * r2 * (-0x1.555556p-3f + r2 * (0x1.111112p-7f + r2 * (-0x1.a01a02p-13f + r2 * 0x1.71de3ap-19f)));
    const float cr = 1.0f + r2 * (-0.5f + r2 * (0x1.555556p-5f + r2 * (-0x1.6c16c2p-10f + r2 * (0x1.a01a02p-16f +
                                                                                            r2 * -0x1.27e4fcp-22f))));
    switch (k & 3) {
    case 0: s = sr; c = cr; break;
    case 1: s = cr; c = -sr; break;
    case 2: s = -sr; c = -cr; break;
    default: s = -cr; c = sr; break;
    }
}
RZK_HD float xsin(float x) { float s, c; xsincos(x, s, c); return s; }
RZK_HD float xcos(float x) { float s, c; xsincos(x, s, c); return c; }
RZK_HD float xtan(float x) { float s, c; xsincos(x, s, c); return s / c; }

RZK_HD float xatan(float x) {
    if (x != x) return x;
    float a = fabsf(x);
    const bool inv = a > 1.0f;
    if (inv) a = 1.0f / a;
    const bool quarter = a > 0x1.a8279ap-2f;
    if (quarter) a = (a - 1.0f) / (a + 1.0f);
    const float z2 = a * a;
    float p = a + a * z2 * (-0x1.555556p-2f + z2 * (0x1.99999ap-3f + z2 * (-0x1.24924ap-3f + z2 * (0x1.c71c72p-4f +
              z2 * (-0x1.745d18p-4f + z2 * (0x1.3b13b2p-4f + z2 * (-0x1.111112p-4f + z2 * (0x1.e1e1e2p-5f +
              z2 * (-0x1.af286cp-5f + z2 * 0x1.861862p-5f)))))))));
    if (quarter) p = kXmPio4Hi + (p + kXmPio4Lo);
    if (inv) p = (kXmPio2_1 - p) + kXmPio2Lo;
    return copysignf(p, x);
}

RZK_HD float xatan2(float y, float x) {
    if (x != x || y != y) return x + y;
    const float ax = fabsf(x), ay = fabsf(y);
    float t;
    if (ay == 0.0f && ax == 0.0f) {
        t = signbit(x) ? kXmPiHi + kXmPiLo : 0.0f;
    } else if (ay <= ax) {
        t = xatan(ay / ax);
        if (signbit(x)) t = (kXmPiHi - t) + kXmPiLo;
    } else {
        const float u = xatan(ax / ay);
        t = signbit(x) ? (kXmPio2_1 + u) + kXmPio2Lo : (kXmPio2_1 - u) + kXmPio2Lo;
    }
    return copysignf(t, y);
}

RZK_HD float xasin_series(float z) {     // |z| <= 1/2
    const float z2 = z * z;
    return z + z * z2 * (0x1.555556p-3f + z2 * (0x1.333334p-4f + z2 * (0x1.6db6dcp-5f + z2 * (0x1.f1c71cp-6f +
           z2 * (0x1.6e8ba2p-6f + z2 * (0x1.1c4ec4p-6f + z2 * (0x1.c9999ap-7f + z2 * (0x1.7a8788p-7f +
           z2 * (0x1.3fde50p-7f + z2 * 0x1.12ef3cp-7f)))))))));
}

RZK_HD float xacos(float x) {
    const float a = fabsf(x);
    if (!(a <= 1.0f)) return (x - x) / (x - x);
    if (a <= 0.5f) return (kXmPio2_1 - xasin_series(x)) + kXmPio2Lo;
    const float t = 2.0f * xasin_series(sqrtf((1.0f - a) * 0.5f));
    return x > 0.0f ? t : (kXmPiHi - t) + kXmPiLo;
}

RZK_HD float xlog(float x) {
    if (x != x || x < 0.0f) return (x - x) / (x - x);
    if (x == 0.0f) return -INFINITY;
    if (x == INFINITY) return x;
    uint32_t ix = __builtin_bit_cast(uint32_t, x);
    int e = (int)(ix >> 23) - 127;
    if (e == -127) {                                   // subnormal
        ix = __builtin_bit_cast(uint32_t, x * 0x1p23f);
        e = (int)(ix >> 23) - 127 - 23;
    }
    float m = __builtin_bit_cast(float, (ix & 0x007fffffu) | 0x3f800000u);
    if (m > 0x1.6a09e6p+0f) { m = m * 0.5f; e += 1; }
    const float sv = (m - 1.0f) / (m + 1.0f);
    const float s2 = sv * sv;
    const float lm = 2.0f * sv + sv * s2 * (0x1.555556p-1f + s2 * (0x1.99999ap-2f + s2 * (0x1.24924ap-2f + s2 * 0x1.c71c72p-3f)));
    const float fe = (float)e;
    return fe * kXmLn2Hi + (lm + fe * kXmLn2Lo);
}

RZK_HD float xexp(float t) {
    if (t != t) return t;
    if (t > 88.7f) return INFINITY;
    if (t < -104.0f) return 0.0f;
    int k = xm_round(t * 0x1.715476p+0f);
    const float fk = (float)k;
    const float r = (t - fk * kXmLn2Hi) - fk * kXmLn2Lo;
    float p = 1.0f + (r + r * r * (0.5f + r * (0x1.555556p-3f + r * (0x1.555556p-5f + r * (0x1.111112p-7f +
              r * (0x1.6c16c2p-10f + r * (0x1.a01a02p-13f + r * 0x1.a01a02p-16f)))))));
    if (k < -126) { p = p * 0x1p-64f; k += 64; }
    return p * __builtin_bit_cast(float, (uint32_t)(k + 127) << 23);
}

RZK_HD float xpow(float x, float y) {     // x >= 0
    if (x == 1.0f || y == 0.0f) return 1.0f;
    if (x == 0.0f) return y > 0.0f ? 0.0f : INFINITY;
    return xexp(y * xlog(x));
}

// Onb::spherical_to_world (onb.cc:155-163)
RZK_HD v3 spherical_to_world(const Onb& b, float theta, float phi) {
    float st = xsin(theta), ct = xcos(theta), sp = xsin(phi), cp = xcos(phi);
    return local_to_world(b, mk(st * cp, st * sp, ct));
}

// Lambertian::pdf (lambertian.cc:70-75), ModifiedPhong::pdf (phong.cc:89-102),
// Ward::pdf (ward.cc:118-134)
RZK_HD float bsdf_pdf(const float* m, const Onb& b, v3 wo, v3 wi) {
    if (dot(wi, b.n) <= kZeroEps || dot(wo, b.n) <= kZeroEps) return 0.0f;
    int type = (int)m[0];
    if (type == kLambert) return dot(b.n, wi) * kInvPi;
    float rho_d, rho_s;
    lobe_weights(m, rho_d, rho_s);
    if (type == kPhong) {
        float ne = m[7];
        float pdf = rho_d * (dot(b.n, wi) * kInvPi);
        v3 wr = b.n * (2.0f * dot(wi, b.n)) - wi;
        float cos_alpha = dot(wo, wr);
        if (cos_alpha > 0.0f) pdf += rho_s * (ne + 1.0f) * kInv2Pi * powf(cos_alpha, ne);
        return pdf;
    }
    float ax = m[7], ay = m[8];
    float ax2 = ax * ax, ay2 = ay * ay;
    float inv4 = 1.0f / (4.0f * kPi * ax * ay);
    float z = dot(b.n, wi);
    v3 wh = unit(world_to_local(b, wi + wo));
    float expo = expf(-(wh.x * wh.x / ax2 + wh.y * wh.y / ay2) / (wh.z * wh.z));
    float ct = wh.z;
    v3 lwo = unit(world_to_local(b, wo));
    float ct3 = ct * ct * ct;
    return rho_d * (z * kInvPi) + rho_s * (inv4 / (dot(wh, lwo) * ct3) * expo);
}

// cosine-weighted direction about b.n from two uniforms (lambertian.cc:13-24)
RZK_HD v3 cosine_dir(const Onb& b, float u0, float u1) {
    float phi = kTwoPi * u0;
    float sin_theta = sqrtf(u1);
    float x = sin_theta * xcos(phi);
    float y = sin_theta * xsin(phi);
    float z = sqrtf(1.0f - sin_theta * sin_theta);
    return unit(local_to_world(b, mk(x, y, z)));
}

// The direction step of BSDF sampling (lambertian.cc:5-27, phong.cc:21-72,
// ward.cc:24-101): u[0] picks the lobe, u[1], u[2] the direction.  Returns 0
// on rejection (wi = 0), 2 for a Lambertian material, whose pdf (lam_pdf, from
// the local z) and BRDF (kd / pi) the sampler gives directly, else 1: the pdf
// and BRDF are bsdf_pdf and bsdf_eval of wi.
RZK_HD int bsdf_sample_dir(const float* m, const Onb& b, v3 wo, const uint32_t* u, v3& wi, float& lam_pdf) {
    wi = mk(0.f, 0.f, 0.f);
    lam_pdf = 0.0f;
    if (dot(wo, b.n) <= kZeroEps) return 0;
    int type = (int)m[0];
    if (type == kLambert) {
        float phi = kTwoPi * u24(u[1]);
        float sin_theta = sqrtf(u24(u[2]));
        float x = sin_theta * xcos(phi), y = sin_theta * xsin(phi);
        float z = sqrtf(1.0f - sin_theta * sin_theta);
        wi = unit(local_to_world(b, mk(x, y, z)));
        lam_pdf = z * kInvPi;
        return 2;
    }
    float rho_d, rho_s;
    lobe_weights(m, rho_d, rho_s);
    if (u24(u[0]) < rho_d) {
        wi = cosine_dir(b, u24(u[1]), u24(u[2]));
    } else if (type == kPhong) {
        float ne = m[7];
        Onb o = onb_from_nu(wo, b.n);
        float phi = kTwoPi * u24(u[1]);
        float alpha = xacos(xpow(u24(u[2]), 1.0f / (ne + 1.0f)));
        v3 wr = spherical_to_world(o, alpha, phi);
        float d = dot(wr, b.n);
        if (d < 0.0f) return 0;
        wi = b.n * (2.0f * d) - wr;
    } else {
        float ax = m[7], ay = m[8];
        float ax2 = ax * ax, ay2 = ay * ay;
        float ay_ax = ay / ax;
        float s0 = u24_open(u[1]), s1 = u24_open(u[2]);
        v3 lwo = unit(world_to_local(b, wo));
        float angle = kTwoPi * s0;
        float phi = xatan(ay_ax * xtan(angle));
        if (xsin(phi) * xsin(angle) < 0 || xcos(phi) * xcos(angle) < 0) phi += kPi;
        float cp = xcos(phi), sp = xsin(phi);
        float theta = xatan(sqrtf(-xlog(s1) / (cp * cp / ax2 + sp * sp / ay2)));
        float st = xsin(theta), ct = xcos(theta);
        v3 wh = mk(st * cp, st * sp, ct);
        v3 lwi = wh * (2.0f * dot(lwo, wh)) - lwo;
        wi = local_to_world(b, lwi);
    }
    return 1;
}

// BSDF sampling: the direction, then its pdf and BRDF.  Returns the BRDF
// value; pdf = 0 on rejection.
RZK_HD v3 bsdf_sample(const float* m, const Onb& b, v3 wo, const uint32_t* u, v3& wi, float& pdf) {
    float lam_pdf;
    const int kind = bsdf_sample_dir(m, b, wo, u, wi, lam_pdf);
    pdf = 0.0f;
    if (kind == 0) return mk(0.f, 0.f, 0.f);
    if (kind == 2) {
        pdf = lam_pdf;
        return mk(m[1], m[2], m[3]) * kInvPi;
    }
    pdf = bsdf_pdf(m, b, wo, wi);
    return bsdf_eval(m, b, wo, wi);
}

// Shirley-Chiu concentric maps (concentric_mapping.h:16-114)
RZK_HD void direction_to_square(v3 wi, float& sx, float& sy) {
    float r2 = 1.0f - wi.z;
    float dx = wi.x / sqrtf(2.0f - r2), dy = wi.y / sqrtf(2.0f - r2);
    float r = sqrtf(dx * dx + dy * dy);
    float phi = xatan2(dy, dx);
    if (phi < -kPi / 4) phi += kTwoPi;
    float a, bb;
    if (phi < kPi / 4) { a = r; bb = phi * a / (kPi / 4); }
    else if (phi < 3 * kPi / 4) { bb = r; a = -(phi - kPi / 2) * bb / (kPi / 4); }
    else if (phi < 5 * kPi / 4) { a = -r; bb = (phi - kPi) * a / (kPi / 4); }
    else { bb = -r; a = -(phi - 3 * kPi / 2) * bb / (kPi / 4); }
    sx = a * 0.5f + 0.5f;
    sy = bb * 0.5f + 0.5f;
}

RZK_HD v3 square_to_direction(float sx, float sy) {
    float phi, r;
    float a = 2.0f * sx - 1.0f, bb = 2.0f * sy - 1.0f;
    if (a > -bb) {
        if (a > bb) { r = a; phi = (kPi / 4) * (bb / a); }
        else { r = bb; phi = (kPi / 4) * (2 - (a / bb)); }
    } else {
        if (a < bb) { r = -a; phi = (kPi / 4) * (4 + (bb / a)); }
        else { r = -bb; if (bb != 0.0f) phi = (kPi / 4) * (6 - (a / bb)); else phi = 0.0f; }
    }
    float dx = r * xcos(phi), dy = r * xsin(phi);
    float r2 = dx * dx + dy * dy;
    return mk(dx * sqrtf(2.0f - r2), dy * sqrtf(2.0f - r2), 1.0f - r2);
}

}  // namespace rzk
