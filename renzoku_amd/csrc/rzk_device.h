// rzk_device.h — device-side traversal and shading building blocks.
#pragma once
#pragma clang fp contract(off)

#include "rzk_internal.h"

namespace rzk {

// Pointers into read-only scene data for wave-uniform loads (AMDGPU constant
// address space: the compiler may fetch through the scalar cache).  Only the
// device pass sees the address space; the host pass never runs these paths.
#if defined(__HIP_DEVICE_COMPILE__)
#define kconst __attribute__((address_space(4)))
#else
#define kconst
#endif

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
    // one flat loop: a step is a leaf test or a node test, then either the
    // next node is known or the stack is popped (one pop site, so the loop
    // keeps a single level and lanes on leaves and on nodes step together)
    for (;;) {
        bool pop;
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
            pop = true;
        } else {
            const BvhNodeGpu* np = sv.nodes + node;
            float4 a = np->a, b4 = np->b, c = np->c;
            int2 ch = *reinterpret_cast<const int2*>(&np->d);
            st.nodes++;
            bool hl = box_hit(mk(a.x, a.y, a.z), mk(a.w, b4.x, b4.y), r, tmin, tmax);
            bool hr = box_hit(mk(b4.z, b4.w, c.x), mk(c.y, c.z, c.w), r, tmin, tmax);
            if (hl && hr) stk[(sp++) * pitch] = ch.x;
            node = hr ? ch.y : ch.x;
            pop = !(hl || hr);
        }
        if (pop) {
            if (sp == 0) break;
            node = stk[(--sp) * pitch];
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

// Any-hit traversal stack: the top kL entries of a lane live in LDS (one
// column per lane, `pitch` = block size), deeper ones in HBM at
// spill[(entry - kL) * spill_pitch + gid], so a launch's LDS footprint stays
// small whatever the tree depth.  kL is even.
template <int kL>
struct AnyStackT {
    static_assert(kL >= 0 && kL % 2 == 0, "LDS stack depth");
    int* lds;
    int pitch;
    int* spill;
    int64_t spitch;
    int64_t gid;
    __device__ __forceinline__ void push(int& sp, int v) {
        if (sp < kL) lds[sp * pitch] = v;
        else spill[(int64_t)(sp - kL) * spitch + gid] = v;
        ++sp;
    }
    // the two reads name their address spaces, so they cannot be merged into
    // one generic-address read
    __device__ __forceinline__ int pop(int& sp) {
        --sp;
        int v;
        if (sp < kL) v = ((__attribute__((address_space(3))) const int*)lds)[sp * pitch];
        else v = ((__attribute__((address_space(1))) const int*)spill)[(int64_t)(sp - kL) * spitch + gid];
        return v;
    }
    // Two ints as one 8-byte entry, for walks that push only pairs (sp stays
    // even): pair k of a lane sits at 8-byte word k * pitch of its LDS column
    // pair (k < kL / 2), deeper ones at word (k - kL / 2) *
    // spitch + gid of the spill area; the same footprint as 2k single entries.
    __device__ __forceinline__ void push2(int& sp, int a, int b) {
        const unsigned long long v = (unsigned long long)(unsigned)a | ((unsigned long long)(unsigned)b << 32);
        if (sp < kL) lds2()[(sp >> 1) * pitch] = v;
        else reinterpret_cast<unsigned long long*>(spill)[(int64_t)((sp - kL) >> 1) * spitch + gid] = v;
        sp += 2;
    }
    __device__ __forceinline__ int2 pop2(int& sp) {
        sp -= 2;
        unsigned long long v;
        if (sp < kL) {
            v = ((__attribute__((address_space(3))) const unsigned long long*)lds2())[(sp >> 1) * pitch];
        } else {
            v = ((__attribute__((address_space(1))) const unsigned long long*)spill)[(int64_t)((sp - kL) >> 1) *
                                                                                    spitch + gid];
        }
        return make_int2((int)(unsigned)v, (int)(unsigned)(v >> 32));
    }
    __device__ __forceinline__ unsigned long long* lds2() const {
        return reinterpret_cast<unsigned long long*>(lds - threadIdx.x) + threadIdx.x;
    }
};

using AnyStack = AnyStackT<kLdsStack>;

template <int kL = kLdsStack>
__device__ __forceinline__ AnyStackT<kL> any_stack(const SceneView& sv, int* lds_base) {
    AnyStackT<kL> s;
    s.lds = lds_base + threadIdx.x;
    s.pitch = blockDim.x;
    s.spill = sv.spill;
    s.spitch = sv.spill_pitch;
    s.gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    return s;
}

// Any hit over the 4-wide collapse of the reference tree.  The boolean equals
// the binary traversal's (aggregate_bvh.cc:452-478) for every ray: a triangle
// is tested iff its own box and all enclosing boxes pass the slab test, and
// the enclosing ones pass whenever the inner one does.  Visit order is free.
template <class Stk>
__device__ __forceinline__ bool any_hit4(const SceneView& sv, const Ray& r, float tmin, float tmax,
                                         Stk& stk, TravStats& st) {
    if (!box_hit(mk(sv.rbmin[0], sv.rbmin[1], sv.rbmin[2]), mk(sv.rbmax[0], sv.rbmax[1], sv.rbmax[2]),
                 r, tmin, tmax))
        return false;
    int node = sv.root4_ref;
    float t, b, g;
    if (node < 0) {
        const TriGpu* tp = sv.tris4 + (~node);
        float4 p0 = tp->p0, e1 = tp->e1, e2 = tp->e2;
        st.tris++;
        return tri_hit(mk(p0.x, p0.y, p0.z), mk(e1.x, e1.y, e1.z), mk(e2.x, e2.y, e2.z), r, tmin, tmax, t, b, g);
    }
    int sp = 0;
    for (;;) {
        const Bvh4NodeGpu* np = sv.nodes4 + node;
        float4 lx = np->lox, ly = np->loy, lz = np->loz, hx = np->hix, hy = np->hiy, hz = np->hiz;
        int4 rf = np->ref;
        st.nodes++;
        const float lxa[4] = {lx.x, lx.y, lx.z, lx.w}, lya[4] = {ly.x, ly.y, ly.z, ly.w};
        const float lza[4] = {lz.x, lz.y, lz.z, lz.w}, hxa[4] = {hx.x, hx.y, hx.z, hx.w};
        const float hya[4] = {hy.x, hy.y, hy.z, hy.w}, hza[4] = {hz.x, hz.y, hz.z, hz.w};
        const int rfa[4] = {rf.x, rf.y, rf.z, rf.w};
        int next = -1;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            int ref = rfa[c];
            if (ref == kEmptyRef) break;
            if (!box_hit(mk(lxa[c], lya[c], lza[c]), mk(hxa[c], hya[c], hza[c]), r, tmin, tmax)) continue;
            if (ref < 0) {
                const TriGpu* tp = sv.tris4 + (~ref);
                float4 p0 = tp->p0, e1 = tp->e1, e2 = tp->e2;
                st.tris++;
                if (tri_hit(mk(p0.x, p0.y, p0.z), mk(e1.x, e1.y, e1.z), mk(e2.x, e2.y, e2.z), r, tmin, tmax, t, b, g))
                    return true;
            } else if (next < 0) {
                next = ref;
            } else {
                stk.push(sp, ref);
            }
        }
        if (next >= 0) {
            node = next;
        } else {
            if (sp == 0) return false;
            node = stk.pop(sp);
        }
    }
}

// Closest hit over the 4-wide collapse, nearest child first, one node per
// step so that a lane can stop between nodes and resume later (the path
// kernels interleave the walks of a wave's lanes with shading this way).
// A step slab-tests the node's children at the tmax it starts with, then
// tests the triangle children that passed, in slot order, shrinking tmax at
// each hit; interior children that passed are visited in order of entry
// distance, the others pushed as (node, entry) pairs, a child only when its
// entry lies within the shrunk tmax, and dropped at pop time when their entry
// lies beyond the tmax of that time.  The hit is the same triangle, t, beta
// and gamma as the reference-order traversal whenever the closest triangle is
// unique (the triangle test is the same arithmetic and every triangle whose
// box chain passes at the final tmax is tested); on exact ties in t the winner
// may differ.  Testing a node's triangles after all of its boxes lets a wave
// run one triangle test per passing leaf of its busiest lane rather than one
// per child slot that any lane needs.  The stack holds up to 2 * sv.stack4
// ints.  any: stop at the first triangle hit (its index; t, beta, gamma of
// that hit).  tmax then never shrinks, so every node whose box passes is
// visited until a hit: the hit / no-hit answer is any_hit4's.
//
// trav4_begin does the root's part (false when the query is already answered);
// trav4_step visits one node and returns false once the walk is over.  The
// answer is then in best (-1: none) and, after a hit, in t_out, b_out, g_out
// (left alone without one).
struct Trav4 {
    int node, sp, best;
    float tmax;
    bool any;
};

__device__ __forceinline__ bool trav4_begin(const SceneView& sv, const Ray& r, float tmin, float tmax, bool any,
                                            Trav4& T, float& t_out, float& b_out, float& g_out, TravStats& st) {
    T.best = -1;
    T.sp = 0;
    T.tmax = tmax;
    T.any = any;
    if (!box_hit(mk(sv.rbmin[0], sv.rbmin[1], sv.rbmin[2]), mk(sv.rbmax[0], sv.rbmax[1], sv.rbmax[2]),
                 r, tmin, tmax))
        return false;
    T.node = sv.root4_ref;
    if (T.node < 0) {
        const TriGpu* tp = sv.tris4 + (~T.node);
        float4 p0 = tp->p0, e1 = tp->e1, e2 = tp->e2;
        float t, b, g;
        st.tris++;
        if (tri_hit(mk(p0.x, p0.y, p0.z), mk(e1.x, e1.y, e1.z), mk(e2.x, e2.y, e2.z), r, tmin, tmax, t, b, g)) {
            t_out = t; b_out = b; g_out = g;
            T.best = __float_as_int(p0.w);
        }
        return false;
    }
    return true;
}

template <class Stk>
__device__ __forceinline__ bool trav4_step(const SceneView& sv, const Ray& r, float tmin, Trav4& T, Stk& stk,
                                           float& t_out, float& b_out, float& g_out, TravStats& st) {
    const Bvh4NodeGpu* np = sv.nodes4 + T.node;
    float4 lx = np->lox, ly = np->loy, lz = np->loz, hx = np->hix, hy = np->hiy, hz = np->hiz;
    int4 rf = np->ref;
    st.nodes++;
    const float lxa[4] = {lx.x, lx.y, lx.z, lx.w}, lya[4] = {ly.x, ly.y, ly.z, ly.w};
    const float lza[4] = {lz.x, lz.y, lz.z, lz.w}, hxa[4] = {hx.x, hx.y, hx.z, hx.w};
    const float hya[4] = {hy.x, hy.y, hy.z, hy.w}, hza[4] = {hz.x, hz.y, hz.z, hz.w};
    const int rfa[4] = {rf.x, rf.y, rf.z, rf.w};
    // every child slot is tested (an empty slot's box values are never used),
    // so the node's loads all issue at once
    float cd[4];
    int cr[4];
    unsigned leaves = 0;                   // slots whose triangle child passed
    float tmax = T.tmax;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int ref = rfa[c];
        float te;
        const bool pass = box_enter(mk(lxa[c], lya[c], lza[c]), mk(hxa[c], hya[c], hza[c]), r, tmin, tmax, te) &&
                          ref != kEmptyRef;
        const bool inner = pass && ref >= 0;
        leaves |= (pass && ref < 0) ? 1u << c : 0u;
        cd[c] = inner ? te : INFINITY;
        cr[c] = inner ? ref : -1;
    }
    while (leaves) {
        const int c = __builtin_ctz(leaves);
        leaves &= leaves - 1u;
        int ref = rfa[0];
        ref = c == 1 ? rfa[1] : ref;
        ref = c == 2 ? rfa[2] : ref;
        ref = c == 3 ? rfa[3] : ref;
        const TriGpu* tp = sv.tris4 + (~ref);
        float4 p0 = tp->p0, e1 = tp->e1, e2 = tp->e2;
        float t, b, g;
        st.tris++;
        if (tri_hit(mk(p0.x, p0.y, p0.z), mk(e1.x, e1.y, e1.z), mk(e2.x, e2.y, e2.z), r, tmin, tmax, t, b, g)) {
            tmax = t; t_out = t; b_out = b; g_out = g;
            T.best = __float_as_int(p0.w);
            if (T.any) { T.tmax = tmax; return false; }
        }
    }
    T.tmax = tmax;
    // ascending entry distance (5-exchange network); empty slots carry +inf
#define RZK_CX(i, j)                                                              \
    {                                                                             \
        const bool sw = cd[j] < cd[i];                                            \
        const float di = cd[i], dj = cd[j];                                       \
        const int ri = cr[i], rj = cr[j];                                         \
        cd[i] = sw ? dj : di; cd[j] = sw ? di : dj;                               \
        cr[i] = sw ? rj : ri; cr[j] = sw ? ri : rj;                               \
    }
    RZK_CX(0, 1) RZK_CX(2, 3) RZK_CX(0, 2) RZK_CX(1, 3) RZK_CX(1, 2)
#undef RZK_CX
    int node = -1;
#pragma unroll
    for (int c = 3; c >= 1; --c)
        if (cr[c] >= 0 && cd[c] <= tmax) stk.push2(T.sp, cr[c], __float_as_int(cd[c]));
    if (cr[0] >= 0 && cd[0] <= tmax) node = cr[0];
    while (node < 0) {
        if (T.sp == 0) return false;
        const int2 e = stk.pop2(T.sp);
        if (__int_as_float(e.y) <= tmax) node = e.x;
    }
    T.node = node;
    return true;
}

// the whole walk at once
template <class Stk>
__device__ __forceinline__ int closest_hit4(const SceneView& sv, const Ray& r, float tmin, float tmax,
                                            Stk& stk, float& t_out, float& b_out, float& g_out,
                                            TravStats& st, bool any = false) {
    Trav4 T;
    if (trav4_begin(sv, r, tmin, tmax, any, T, t_out, b_out, g_out, st))
        while (trav4_step(sv, r, tmin, T, stk, t_out, b_out, g_out, st)) {
        }
    return T.best;
}

// Wave-coherent any hit: the 64 lanes walk the 4-wide tree together.  The
// node index is wave-uniform (readfirstlane), so a node is fetched once per
// wave through the scalar cache; each lane tests the four child boxes against
// its own ray, a ballot decides which children the wave descends into, and a
// lane leaves the walk once it has found an occluder.  Each lane still tests
// exactly the triangles whose box chain its own ray passes (the result is the
// per-lane any_hit4 result); only the schedule is shared.  kSigned: every
// walking lane's direction has the wave-uniform component signs (nx, ny, nz:
// true for a negative component), so the child tests use box_hit_signed
// (same outcomes, fewer operations).  kCount: st.nodes counts, per lane, the
// nodes its own ray reaches (the parent's test of that child box passed for
// this lane), not every node the wave fetches; own[] keeps that bit for each
// stack entry (entries past 64 are counted as reached).
template <bool kSigned, bool kCount>
__device__ __forceinline__ bool any_hit4_wave_walk(const SceneView& sv, const Ray& r, float tmin, float tmax,
                                                   bool live, bool nx, bool ny, bool nz, int* wstack,
                                                   TravStats& st) {
    // float offsets in a node of the entry (e*) and exit (x*) plane arrays of
    // each axis: lo.x[4] at 0, lo.y at 4, lo.z at 8, hi.x at 12, hi.y at 16,
    // hi.z at 20; a negative component swaps lo and hi.  Picking the planes by
    // load address keeps the selection out of the per-node instructions.
    const int ex = nx ? 12 : 0, xx = nx ? 0 : 12;
    const int ey = ny ? 16 : 4, xy = ny ? 4 : 16;
    const int ez = nz ? 20 : 8, xz = nz ? 8 : 20;
    bool hit = false;
    float t, b, g;
    int node = sv.root4_ref;
    if (node < 0) {
        if (live) {
            const TriGpu* tp = sv.tris4 + (~node);
            float4 p0 = tp->p0, e1 = tp->e1, e2 = tp->e2;
            st.tris++;
            hit = tri_hit(mk(p0.x, p0.y, p0.z), mk(e1.x, e1.y, e1.z), mk(e2.x, e2.y, e2.z), r, tmin, tmax, t, b, g);
        }
        return hit;
    }
    int sp = 0;
    unsigned long long own = 0ull;      // kCount: bit e = this lane's ray passed stack entry e's box
    bool cur_own = true, next_own = false;
    for (;;) {
        node = __builtin_amdgcn_readfirstlane(node);
        // wave-uniform address in the constant address space: scalar loads
        const kconst float* nf = (const kconst float*)(sv.nodes4 + node);
        const kconst int* ni = (const kconst int*)(sv.nodes4 + node);
        // kSigned: lxa..hza hold the entry planes (lxa, lya, lza) and exit planes
        // (hxa, hya, hza) rather than lo and hi
        const int o0 = kSigned ? ex : 0, o1 = kSigned ? ey : 4, o2 = kSigned ? ez : 8;
        const int o3 = kSigned ? xx : 12, o4 = kSigned ? xy : 16, o5 = kSigned ? xz : 20;
        const float lxa[4] = {nf[o0], nf[o0 + 1], nf[o0 + 2], nf[o0 + 3]};
        const float lya[4] = {nf[o1], nf[o1 + 1], nf[o1 + 2], nf[o1 + 3]};
        const float lza[4] = {nf[o2], nf[o2 + 1], nf[o2 + 2], nf[o2 + 3]};
        const float hxa[4] = {nf[o3], nf[o3 + 1], nf[o3 + 2], nf[o3 + 3]};
        const float hya[4] = {nf[o4], nf[o4 + 1], nf[o4 + 2], nf[o4 + 3]};
        const float hza[4] = {nf[o5], nf[o5 + 1], nf[o5 + 2], nf[o5 + 3]};
        const int rfa[4] = {ni[24], ni[25], ni[26], ni[27]};
        if (kCount) {
            if (live && cur_own) st.nodes++;
        } else if (live) {
            st.nodes++;
        }
        // the four child tests first, as straight-line code: every lane evaluates
        // them (idle lanes hold finite rays; empty slots hold finite zero boxes and
        // are masked out below), so the node arrives in one scalar round trip
        bool box[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            if (kSigned) {
                // box_hit_signed, the planes picked by the loads above
                box[c] = box_hit_planes(mk(lxa[c], lya[c], lza[c]), mk(hxa[c], hya[c], hza[c]), r, tmin, tmax);
            } else {
                box[c] = box_hit(mk(lxa[c], lya[c], lza[c]), mk(hxa[c], hya[c], hza[c]), r, tmin, tmax);
            }
        }
        // lanes still walking, as a wave mask (refreshed after a leaf's tests)
        unsigned long long lm = __builtin_amdgcn_ballot_w64(live);
        int next = -1;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            int ref = rfa[c];
            if (ref == kEmptyRef || (__builtin_amdgcn_ballot_w64(box[c]) & lm) == 0ull) continue;
            if (ref < 0) {
                const kconst float* tf = (const kconst float*)(sv.tris4 + (~ref));
                v3 p0 = mk(tf[0], tf[1], tf[2]), e1 = mk(tf[4], tf[5], tf[6]), e2 = mk(tf[8], tf[9], tf[10]);
                if (box[c] & live) {
                    st.tris++;
                    if (tri_hit(p0, e1, e2, r, tmin, tmax, t, b, g)) {
                        hit = true;
                        live = false;
                    }
                }
                lm = __builtin_amdgcn_ballot_w64(live);
            } else if (next < 0) {
                next = ref;
                if (kCount) next_own = box[c];
            } else {
                if (kCount && sp < 64) {
                    const unsigned long long bit = 1ull << sp;
                    own = box[c] ? (own | bit) : (own & ~bit);
                }
                wstack[sp++] = ref;
            }
        }
        if (lm == 0ull) return hit;
        if (next >= 0) {
            node = next;
            if (kCount) cur_own = next_own;
        } else {
            if (sp == 0) return hit;
            node = wstack[--sp];
            if (kCount) cur_own = sp < 64 ? ((own >> sp) & 1ull) != 0ull : true;
        }
    }
}

// Every lane of the wave must call this together; `active` false lanes ride
// along idle.  `wstack` is this wave's LDS stack of sv.stack4 ints.
template <bool kCount = false>
__device__ __forceinline__ bool any_hit4_wave(const SceneView& sv, const Ray& r, float tmin, float tmax,
                                              bool active, int* wstack, TravStats& st) {
    bool live = active && box_hit(mk(sv.rbmin[0], sv.rbmin[1], sv.rbmin[2]),
                                  mk(sv.rbmax[0], sv.rbmax[1], sv.rbmax[2]), r, tmin, tmax);
    const unsigned long long lm = __builtin_amdgcn_ballot_w64(live);
    if (lm == 0ull) return false;
    // direction signs shared by every walking lane?
    const unsigned long long mx = __builtin_amdgcn_ballot_w64(r.inv_d.x < 0.f) & lm;
    const unsigned long long my = __builtin_amdgcn_ballot_w64(r.inv_d.y < 0.f) & lm;
    const unsigned long long mz = __builtin_amdgcn_ballot_w64(r.inv_d.z < 0.f) & lm;
    if ((mx == 0ull || mx == lm) && (my == 0ull || my == lm) && (mz == 0ull || mz == lm))
        return any_hit4_wave_walk<true, kCount>(sv, r, tmin, tmax, live, mx != 0ull, my != 0ull, mz != 0ull, wstack,
                                                st);
    return any_hit4_wave_walk<false, kCount>(sv, r, tmin, tmax, live, false, false, false, wstack, st);
}

// Segment visibility p -> q (aggregate.cc:84-90): true when blocked.
template <class Stk>
__device__ __forceinline__ bool segment_blocked(const SceneView& sv, v3 p, v3 q, float tmin,
                                                Stk& stk, TravStats& st) {
    Ray r = ray_ctor(p, q - p);
    float tmax = length(q - p) - tmin;
    return any_hit4(sv, r, tmin, tmax, stk, st);
}

// Segment visibility for a whole wave of segments (see any_hit4_wave)
template <bool kCount = false>
__device__ __forceinline__ bool segment_blocked_wave(const SceneView& sv, v3 p, v3 q, float tmin, bool active,
                                                     int* wstack, TravStats& st) {
    Ray r = ray_ctor(p, q - p);
    float tmax = length(q - p) - tmin;
    return any_hit4_wave<kCount>(sv, r, tmin, tmax, active, wstack, st);
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
