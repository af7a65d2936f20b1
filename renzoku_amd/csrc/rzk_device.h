// rzk_device.h — device-side traversal and shading building blocks.
#pragma once
#pragma clang fp contract(off)

#include "rzk_internal.h"

namespace rzk {

// Traversal statistics: interior nodes visited, triangles tested.
struct TravStats { int nodes = 0, tris = 0; };

// Stack-based BVH2 traversal with Renzoku's visit order (aggregate_bvh.cc:402-478):
// at an interior node both child boxes are tested against the current tmax;
// the right child is taken first, the left one is deferred.  Leaves hold one
// triangle.  The stack lives in LDS, one column per lane (`stk` already offset
// by the lane, `pitch` = block size), depth bounded by the tree depth.
// kAny: return at the first triangle hit in [tmin, tmax].
template <bool kAny>
__device__ __forceinline__ int traverse(const SceneView& sv, const Ray& r, float tmin, float tmax,
                                        int* __restrict__ stk, int pitch, float& t_out,
                                        float& b_out, float& g_out, TravStats& st) {
    int best = -1;
    if (!box_hit(mk(sv.rbmin[0], sv.rbmin[1], sv.rbmin[2]), mk(sv.rbmax[0], sv.rbmax[1], sv.rbmax[2]),
                 r, tmin, tmax))
        return -1;
    int node = sv.root_ref;
    int sp = 0;
    for (;;) {
        if (node < 0) {
            const TriGpu* tp = sv.tris + (~node);
            float4 p0 = tp->p0, e1 = tp->e1, e2 = tp->e2;
            st.tris++;
            float t, b, g;
            if (tri_hit(mk(p0.x, p0.y, p0.z), mk(e1.x, e1.y, e1.z), mk(e2.x, e2.y, e2.z), r, tmin,
                        tmax, t, b, g)) {
                best = __float_as_int(p0.w);
                if (kAny) return best;
                tmax = t; t_out = t; b_out = b; g_out = g;
            }
            if (sp == 0) break;
            node = stk[(--sp) * pitch];
        } else {
            const BvhNodeGpu* np = sv.nodes + node;
            float4 a = np->a, b4 = np->b, c = np->c;
            int2 ch = *reinterpret_cast<const int2*>(&np->d);
            st.nodes++;
            bool hl = box_hit(mk(a.x, a.y, a.z), mk(a.w, b4.x, b4.y), r, tmin, tmax);
            bool hr = box_hit(mk(b4.z, b4.w, c.x), mk(c.y, c.z, c.w), r, tmin, tmax);
            if (hl && hr) {
                stk[(sp++) * pitch] = ch.x;
                node = ch.y;
            } else if (hr) {
                node = ch.y;
            } else if (hl) {
                node = ch.x;
            } else {
                if (sp == 0) break;
                node = stk[(--sp) * pitch];
            }
        }
    }
    return best;
}

// Wave-reduced traversal work, added to 64-bit launch totals {nodes, tris}.
// Every lane of the wave must reach this call.
__device__ __forceinline__ void add_work_totals(unsigned long long* tot, const TravStats& st) {
    unsigned long long a = (unsigned long long)st.nodes, b = (unsigned long long)st.tris;
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(tot, a);
        atomicAdd(tot + 1, b);
    }
}

// Segment visibility p -> q (aggregate.cc:84-90): true when blocked.
__device__ __forceinline__ bool segment_blocked(const SceneView& sv, v3 p, v3 q, float tmin,
                                                int* stk, int pitch, TravStats& st) {
    Ray r = ray_ctor(p, q - p);
    float tmax = length(q - p) - tmin;
    float t, b, g;
    return traverse<true>(sv, r, tmin, tmax, stk, pitch, t, b, g, st) >= 0;
}

// Receiver record from a closest hit (triangle.cc:277-296, surface.cc:62-96,
// local_geometry.h:21-30).  Flat triangles: n0 = n1 = n2 = n.
__device__ __forceinline__ void fill_receiver(const SceneView& sv, const Ray& r, float t, float beta,
                                              float gamma, int tri, float* rec, int& mat,
                                              int& light) {
    if (tri < 0) {
        for (int i = 0; i < RZK_RECV_FLOATS; ++i) rec[i] = 0.f;
        mat = -1; light = -1;
        return;
    }
    float4 nf = sv.tri_n[tri];
    float4 tf = sv.tri_tan[tri];
    v3 n = mk(nf.x, nf.y, nf.z);
    v3 sn = unit((n + (n - n) * beta) + (n - n) * gamma);
    int4 sf = sv.surf[__float_as_int(nf.w)];
    int m = sf.x;
    if (sf.z & RZK_SURF_FLIPPED) { n = -n; sn = -sn; }
    if ((sf.z & RZK_SURF_TWO_SIDED) && dot(sn, r.d) > 0.0f) {
        if (m >= 0) { sn = -sn; n = -n; }
    }
    Onb b = onb_from_nu(sn, mk(tf.x, tf.y, tf.z));
    v3 p = r.o + r.d * t;
    st3(rec + 0, p); st3(rec + 3, n); st3(rec + 6, sn); st3(rec + 9, -r.d);
    st3(rec + 12, b.u); st3(rec + 15, b.v); st3(rec + 18, b.n);
    mat = m;
    light = sf.y;
}

// BrdfPointLight::brdf (brdf_point_light.h:235-240)
__device__ __forceinline__ v3 vpl_brdf(const float* mat, const float* vpl, int vmat, v3 wo) {
    if (vmat < 0) return mk(1.f, 1.f, 1.f);
    Onb b; b.u = ld3(vpl + 6); b.v = ld3(vpl + 9); b.n = ld3(vpl + 12);
    return bsdf_eval(mat + RZK_MAT_FLOATS * vmat, b, ld3(vpl + 15), wo);
}

// VirtualPointLight::gather_point after the visibility test (vpl.cc:674-687)
__device__ __forceinline__ v3 gather_shade(const float* mat, const float* rc, int rmat,
                                           const float* vpl, int vmat, int mode, float thr) {
    v3 p = ld3(rc), pn = ld3(rc + 6), q = ld3(vpl), qn = ld3(vpl + 3);
    float G = form_factor_abs(p, pn, q, qn);
    v3 wi = unit(q - p);
    Onb b; b.u = ld3(rc + 12); b.v = ld3(rc + 15); b.n = ld3(rc + 18);
    v3 fs = bsdf_eval(mat + RZK_MAT_FLOATS * rmat, b, ld3(rc + 9), wi);
    v3 fl = vpl_brdf(mat, vpl, vmat, -wi);
    return clamp_term(mode, thr, ld3(vpl + 18), fs, fl, G);
}

// VirtualPointLight::incident_radiance (vpl.cc:411-440): geometric normal,
// white surface BRDF, no visibility
__device__ __forceinline__ v3 incident_radiance(const float* mat, const float* rc, const float* vpl,
                                                int vmat, int mode, float thr) {
    v3 p = ld3(rc), pn = ld3(rc + 3), q = ld3(vpl), qn = ld3(vpl + 3);
    float G = form_factor_abs(p, pn, q, qn);
    v3 wi = unit(q - p);
    v3 fl = vpl_brdf(mat, vpl, vmat, -wi);
    return clamp_term(mode, thr, ld3(vpl + 18), mk(1.f, 1.f, 1.f), fl, G);
}

}  // namespace rzk
