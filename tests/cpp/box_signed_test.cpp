// Host check of box_hit_signed against box_hit (renzoku_amd/csrc/rzk_math.h):
// for rays whose direction component signs are given, the plane-selecting
// slab test must give the same boolean as the min/max slab test
// (boundingbox.h:238-280 restated by box_hit) on every input.  Random boxes and
// rays, plus the cases where rounding could separate the two forms: origins on
// box faces, zero-width boxes, direction components below the 1e-6 clamp, huge
// and tiny coordinates, and segment-style tmax (|q - p| - tmin, possibly
// below tmin).  Host code only: compiled by hipcc, runs without a GPU.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>

#include "rzk_math.h"

using namespace rzk;

static uint64_t checks = 0, hits = 0, fails = 0;

static void check(v3 lo, v3 hi, v3 o, v3 d, float tmin, float tmax) {
    Ray r = ray_ctor(o, d);
    bool nx = r.inv_d.x < 0.f, ny = r.inv_d.y < 0.f, nz = r.inv_d.z < 0.f;
    bool a = box_hit(lo, hi, r, tmin, tmax);
    bool b = box_hit_signed(lo, hi, nx, ny, nz, r, tmin, tmax);
    ++checks;
    hits += a;
    if (a != b && fails++ < 10)
        std::printf("MISMATCH lo (%a %a %a) hi (%a %a %a) o (%a %a %a) d (%a %a %a) tmin %a tmax %a: %d vs %d\n",
                    lo.x, lo.y, lo.z, hi.x, hi.y, hi.z, o.x, o.y, o.z, d.x, d.y, d.z, tmin, tmax, a, b);
}

int main() {
    std::mt19937_64 g(2013);
    std::uniform_real_distribution<float> u(-1.f, 1.f);
    auto rv = [&](float s) { return mk(u(g) * s, u(g) * s, u(g) * s); };
    const float scales[] = {1e-3f, 1.f, 555.f, 1e4f};
    for (float s : scales) {
        for (int i = 0; i < 400000; ++i) {
            v3 a = rv(s), b = rv(s);
            v3 lo = mk(std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z));
            v3 hi = mk(std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z));
            int kind = i % 8;
            if (kind == 1) lo.x = hi.x;                       // zero-width slab
            if (kind == 2) { lo.y = hi.y; lo.z = hi.z; }
            v3 o = rv(2.f * s);
            if (kind == 3) o.x = lo.x;                        // origin on a face
            if (kind == 4) { o.y = hi.y; o.z = lo.z; }
            v3 d = rv(1.f);
            if (kind == 5) d.x = u(g) * 1e-7f;               // below the 1e-6 clamp
            if (kind == 6) { d.y = 0.f; d.z = -0.f; }
            if (kind == 7) d = mk(lo.x + 0.5f * (hi.x - lo.x) - o.x, lo.y + 0.5f * (hi.y - lo.y) - o.y,
                                  lo.z + 0.5f * (hi.z - lo.z) - o.z);   // through the centre
            float tmin = 1e-3f;
            float len = s * (0.5f + 3.f * (u(g) + 1.f));
            float tmax = (i & 1) ? 1e6f : len - tmin;
            if (i % 97 == 0) tmax = tmin * 0.5f;              // segment shorter than 2 tmin
            check(lo, hi, o, d, tmin, tmax);
        }
    }
    // segments between random points of a grid of boxes sharing faces
    for (int i = 0; i < 400000; ++i) {
        float x0 = std::floor(u(g) * 8.f), y0 = std::floor(u(g) * 8.f), z0 = std::floor(u(g) * 8.f);
        v3 lo = mk(x0 * 10.f, y0 * 10.f, z0 * 10.f), hi = mk(x0 * 10.f + 10.f, y0 * 10.f + 10.f, z0 * 10.f + 10.f);
        v3 p = mk(std::floor(u(g) * 80.f), std::floor(u(g) * 80.f), std::floor(u(g) * 80.f));
        v3 q = mk(std::floor(u(g) * 80.f), std::floor(u(g) * 80.f), std::floor(u(g) * 80.f));
        v3 dd = mk(q.x - p.x, q.y - p.y, q.z - p.z);
        float tmin = 1e-3f;
        float tmax = std::sqrt(dd.x * dd.x + dd.y * dd.y + dd.z * dd.z) - tmin;
        check(lo, hi, p, dd, tmin, tmax);
    }
    std::printf("box_hit_signed: %llu checks, %llu hits, %llu mismatches\n", (unsigned long long)checks,
                (unsigned long long)hits, (unsigned long long)fails);
    return fails == 0 ? 0 : 1;
}
