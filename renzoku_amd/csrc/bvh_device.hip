// bvh_device.hip — the scene's trees built on the device (AggregateBvh::build,
// aggregate_bvh.cc:48-240; morton.h:28-120).
//
// Binary tree: node for node and bit for bit the tree bvh_build_host makes.
// Triangle boxes and the scene box (origin included), 30-bit Morton codes of
// the box centres, a stable radix sort of (code, triangle), then ranges split
// top-down, one depth per launch, after the last code that shares more than
// the range's common prefix with its first (midpoint for equal codes), and
// boxes unioned bottom-up, one depth per launch.  Interior nodes are numbered
// in right-first depth-first preorder straight from their ranges: a range
// [s, e] holds e - s interior nodes, so a node's right child is the next index
// and its left child follows the right subtree.
//
// 4-wide tree: the binary tree collapsed by opening, at each node, the interior
// child with the largest surface area until it has four children (the rule the
// host build applies to its SAH tree), laid out in depth-first preorder from
// bottom-up subtree counts.  Its leaves are the Morton-order triangles.  Any-hit
// and segment answers do not depend on the tree (a triangle is tested iff its
// own box passes); closest hits are the same wherever the nearest triangle is
// unique.
#pragma clang fp contract(off)

#include <rocprim/device/device_radix_sort.hpp>

#include <vector>

#include "rzk_internal.h"

namespace rzk {
namespace {

constexpr int kBB = 256;          // threads per block of every build kernel
constexpr int kMaxLevels = 128;   // binary depth <= 30 prefix bits + 31 midpoint splits

template <typename T>
struct Dev {
    T* p = nullptr;
    explicit Dev(size_t n) { RZK_HIP(hipMalloc((void**)&p, sizeof(T) * (n > 0 ? n : 1))); }
    ~Dev() { if (p) (void)hipFree(p); }
    Dev(const Dev&) = delete;
    Dev& operator=(const Dev&) = delete;
    T* release() { T* q = p; p = nullptr; return q; }
};

inline unsigned blocks(int64_t n) { return (unsigned)((n + kBB - 1) / kBB); }

struct Item { int s, e, idx, pad; };   // an interior node: leaf range [s, e], preorder index
struct Item4 { int b, at; };           // a 4-wide node: its binary node, its 4-wide index

// Per triangle its box; per block the smallest lows and largest highs of its
// triangles (values only: the sign of a zero is settled by k_scene_box).
__global__ void __launch_bounds__(kBB) k_tri_boxes(const float* __restrict__ pos, int n,
                                                   float* __restrict__ tbox, float* __restrict__ part) {
    __shared__ float red[6][kBB];
    int k = blockIdx.x * kBB + threadIdx.x;
    v3 lo = mk(INFINITY, INFINITY, INFINITY), hi = mk(-INFINITY, -INFINITY, -INFINITY);
    if (k < n) {
        const float* p = pos + 9 * (int64_t)k;
        v3 p0 = ld3(p), p1 = ld3(p + 3), p2 = ld3(p + 6);
        lo = vmin(vmin(p0, p1), p2);
        hi = vmax(vmax(p0, p1), p2);
        st3(tbox + 6 * (int64_t)k, lo);
        st3(tbox + 6 * (int64_t)k + 3, hi);
    }
    red[0][threadIdx.x] = lo.x; red[1][threadIdx.x] = lo.y; red[2][threadIdx.x] = lo.z;
    red[3][threadIdx.x] = hi.x; red[4][threadIdx.x] = hi.y; red[5][threadIdx.x] = hi.z;
    __syncthreads();
    for (int w = kBB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            for (int a = 0; a < 3; ++a) red[a][threadIdx.x] = smin(red[a][threadIdx.x], red[a][threadIdx.x + w]);
            for (int a = 3; a < 6; ++a) red[a][threadIdx.x] = smax(red[a][threadIdx.x], red[a][threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// The scene box: the origin and every triangle box (boundingbox.h:100-103).  The
// host build folds the boxes in order from a +0 start and replaces a bound
// only by a strictly smaller (larger) value, so a bound whose value is zero is
// +0 there; the fold here reaches the same values in another order and writes
// a zero bound as +0.  Out: lo xyz, hi xyz, extent xyz.
__global__ void __launch_bounds__(kBB) k_scene_box(const float* __restrict__ part, int nb, float* __restrict__ sbox) {
    __shared__ float red[6][kBB];
    float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int i = threadIdx.x; i < nb; i += kBB) {
        for (int a = 0; a < 3; ++a) v[a] = smin(v[a], part[6 * i + a]);
        for (int a = 3; a < 6; ++a) v[a] = smax(v[a], part[6 * i + a]);
    }
    for (int a = 0; a < 6; ++a) red[a][threadIdx.x] = v[a];
    __syncthreads();
    for (int w = kBB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            for (int a = 0; a < 3; ++a) red[a][threadIdx.x] = smin(red[a][threadIdx.x], red[a][threadIdx.x + w]);
            for (int a = 3; a < 6; ++a) red[a][threadIdx.x] = smax(red[a][threadIdx.x], red[a][threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        float b[6];
        for (int a = 0; a < 6; ++a) {
            float x = a < 3 ? smin(red[a][0], 0.f) : smax(red[a][0], 0.f);
            b[a] = x == 0.f ? 0.f : x;
        }
        for (int a = 0; a < 6; ++a) sbox[a] = b[a];
        for (int a = 0; a < 3; ++a) sbox[6 + a] = b[3 + a] - b[a];
    }
}

__global__ void __launch_bounds__(kBB) k_morton(const float* __restrict__ tbox, const float* __restrict__ sbox, int n,
                                                uint32_t* __restrict__ code, int* __restrict__ tri) {
    int k = blockIdx.x * kBB + threadIdx.x;
    if (k >= n) return;
    v3 lo = ld3(tbox + 6 * (int64_t)k), hi = ld3(tbox + 6 * (int64_t)k + 3);
    v3 c = (lo + hi) * 0.5f;
    c = (c - ld3(sbox)) / ld3(sbox + 6);
    code[k] = morton3d(c.x, c.y, c.z);
    tri[k] = k;
}

// triangle records in sorted (leaf) order: {p0, original id}, {p0 - p1}, {p0 - p2}
__global__ void __launch_bounds__(kBB) k_sorted_tris(const float* __restrict__ pos, const int* __restrict__ sorted_tri,
                                                     int n, TriGpu* __restrict__ tris) {
    int i = blockIdx.x * kBB + threadIdx.x;
    if (i >= n) return;
    int ot = sorted_tri[i];
    const float* p = pos + 9 * (int64_t)ot;
    v3 p0 = ld3(p), p1 = ld3(p + 3), p2 = ld3(p + 6);
    v3 e1 = p0 - p1, e2 = p0 - p2;
    TriGpu t;
    t.p0 = make_float4(p0.x, p0.y, p0.z, __int_as_float(ot));
    t.e1 = make_float4(e1.x, e1.y, e1.z, 0.f);
    t.e2 = make_float4(e2.x, e2.y, e2.z, 0.f);
    tris[i] = t;
}

// per original triangle: geometric normal (triangle.cc:17-18) with the surface id's
// bits, and the tangent (the host build's arithmetic, so the same bits)
__global__ void __launch_bounds__(kBB) k_tri_frames(const float* __restrict__ pos, const float* __restrict__ tan,
                                                    const int* __restrict__ tri_surface, int n,
                                                    float4* __restrict__ tn, float4* __restrict__ tt) {
    int k = blockIdx.x * kBB + threadIdx.x;
    if (k >= n) return;
    const float* p = pos + 9 * (int64_t)k;
    v3 p0 = ld3(p), p1 = ld3(p + 3), p2 = ld3(p + 6);
    v3 nn = unit(cross(p0 - p1, p0 - p2));
    tn[k] = make_float4(nn.x, nn.y, nn.z, __int_as_float(tri_surface[k]));
    const float* t = tan + 3 * (int64_t)k;
    tt[k] = make_float4(t[0], t[1], t[2], 0.f);
}

__device__ int clz32(uint32_t v) { return v ? __clz((int)v) : 32; }

// last index i in [first, last) with clz(code[first] ^ code[i]) above the
// range's common prefix (the predicate holds on a prefix of the sorted range)
__device__ int split_range(const uint32_t* code, int first, int last) {
    uint32_t fc = code[first], lc = code[last];
    if (fc == lc) return (first + last) >> 1;
    int cp = clz32(fc ^ lc);
    int lo = first, hi = last;
    while (hi - lo > 1) {
        int mid = lo + ((hi - lo) >> 1);
        if (clz32(fc ^ code[mid]) > cp) lo = mid; else hi = mid;
    }
    return lo;
}

// One depth of interior nodes: split each range, name its children (preorder
// index, or ~first leaf for a leaf), queue the interior ones for the next depth
// at items[off + n + ...] (their order there does not matter: every output is
// placed by preorder index).
__global__ void __launch_bounds__(kBB) k_split_level(const uint32_t* __restrict__ code, Item* items, int off, int n,
                                                     int* __restrict__ next_n, int2* __restrict__ kid) {
    int t = blockIdx.x * kBB + threadIdx.x;
    if (t >= n) return;
    Item it = items[off + t];
    int sp = split_range(code, it.s, it.e);
    int rref = ~(sp + 1), lref = ~it.s;
    if (sp + 1 < it.e) {
        rref = it.idx + 1;
        int slot = atomicAdd(next_n, 1);
        items[off + n + slot] = Item{sp + 1, it.e, rref, 0};
    }
    if (sp > it.s) {
        lref = it.idx + 1 + (it.e - (sp + 1));
        int slot = atomicAdd(next_n, 1);
        items[off + n + slot] = Item{it.s, sp, lref, 0};
    }
    kid[it.idx] = make_int2(lref, rref);
}

__device__ inline const float* child_box(int ref, const float* nbox, const float* tbox, const int* sorted_tri) {
    return ref >= 0 ? nbox + 6 * (int64_t)ref : tbox + 6 * (int64_t)sorted_tri[~ref];
}

__device__ inline float half_area(const float* b) {
    float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
    return dx * dy + dy * dz + dz * dx;
}

// children of binary node b in the 4-wide tree: open the interior child of
// largest area (the first of equals) in place until there are four
__device__ int open4(const int2* kid, const float* nbox, int b, int c[4]) {
    int2 k = kid[b];
    c[0] = k.x;
    c[1] = k.y;
    c[2] = c[3] = kEmptyRef;
    int cnt = 2;
    while (cnt < 4) {
        int bi = -1;
        float ba = -1.f;
        for (int i = 0; i < 4; ++i)
            if (i < cnt && c[i] >= 0) {
                float a = half_area(nbox + 6 * (int64_t)c[i]);
                if (a > ba) { ba = a; bi = i; }
            }
        if (bi < 0) break;
        int2 kx = kid[c[bi]];
        for (int j = 3; j > 0; --j)
            if (j > bi + 1 && j <= cnt) c[j] = c[j - 1];
        c[bi] = kx.x;
        c[bi + 1] = kx.y;
        ++cnt;
    }
    return cnt;
}

// One depth, deepest first: each node's box is the union of its children's
// (left operand first, as the host build), its record holds both child boxes
// and refs, and its 4-wide subtree size is counted from its 4-wide children.
__global__ void __launch_bounds__(kBB) k_refit_level(const Item* __restrict__ items, int off, int n,
                                                     const int2* __restrict__ kid, const int* __restrict__ sorted_tri,
                                                     const float* __restrict__ tbox, float* nbox,
                                                     BvhNodeGpu* __restrict__ nodes, int* cnt4) {
    int t = blockIdx.x * kBB + threadIdx.x;
    if (t >= n) return;
    int b = items[off + t].idx;
    int2 k = kid[b];
    const float* L = child_box(k.x, nbox, tbox, sorted_tri);
    const float* R = child_box(k.y, nbox, tbox, sorted_tri);
    float l[6], r[6];
    for (int a = 0; a < 6; ++a) { l[a] = L[a]; r[a] = R[a]; }
    v3 lo = vmin(ld3(l), ld3(r)), hi = vmax(ld3(l + 3), ld3(r + 3));
    st3(nbox + 6 * (int64_t)b, lo);
    st3(nbox + 6 * (int64_t)b + 3, hi);
    BvhNodeGpu nd;
    nd.a = make_float4(l[0], l[1], l[2], l[3]);
    nd.b = make_float4(l[4], l[5], r[0], r[1]);
    nd.c = make_float4(r[2], r[3], r[4], r[5]);
    nd.d = make_int4(k.x, k.y, 0, 0);
    nodes[b] = nd;
    int c[4];
    int cnt = open4(kid, nbox, b, c);
    int total = 1;
    for (int i = 0; i < 4; ++i)
        if (i < cnt && c[i] >= 0) total += cnt4[c[i]];
    cnt4[b] = total;
}

// One depth of the 4-wide tree: write each node, place its interior children
// after it in preorder (each after the previous one's subtree), queue them.
__global__ void __launch_bounds__(kBB) k_collapse_level(Item4* items, int off, int n, int* __restrict__ next_n,
                                                        const int2* __restrict__ kid, const float* __restrict__ nbox,
                                                        const float* __restrict__ tbox, const int* __restrict__ sorted_tri,
                                                        const int* __restrict__ cnt4, Bvh4NodeGpu* __restrict__ nodes4) {
    int t = blockIdx.x * kBB + threadIdx.x;
    if (t >= n) return;
    Item4 it = items[off + t];
    int c[4];
    int cnt = open4(kid, nbox, it.b, c);
    float lo[3][4], hi[3][4];
    int rf[4];
    int at = it.at + 1;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (i < cnt) {
            const float* bx = child_box(c[i], nbox, tbox, sorted_tri);
            for (int a = 0; a < 3; ++a) { lo[a][i] = bx[a]; hi[a][i] = bx[3 + a]; }
            if (c[i] >= 0) {
                rf[i] = at;
                int slot = atomicAdd(next_n, 1);
                items[off + n + slot] = Item4{c[i], at};
                at += cnt4[c[i]];
            } else {
                rf[i] = c[i];
            }
        } else {
            for (int a = 0; a < 3; ++a) { lo[a][i] = 0.f; hi[a][i] = 0.f; }
            rf[i] = kEmptyRef;
        }
    }
    Bvh4NodeGpu nd;
    nd.lox = make_float4(lo[0][0], lo[0][1], lo[0][2], lo[0][3]);
    nd.loy = make_float4(lo[1][0], lo[1][1], lo[1][2], lo[1][3]);
    nd.loz = make_float4(lo[2][0], lo[2][1], lo[2][2], lo[2][3]);
    nd.hix = make_float4(hi[0][0], hi[0][1], hi[0][2], hi[0][3]);
    nd.hiy = make_float4(hi[1][0], hi[1][1], hi[1][2], hi[1][3]);
    nd.hiz = make_float4(hi[2][0], hi[2][1], hi[2][2], hi[2][3]);
    nd.ref = make_int4(rf[0], rf[1], rf[2], rf[3]);
    nd.pad = make_int4(0, 0, 0, 0);
    nodes4[it.at] = nd;
}

}  // namespace
}  // namespace rzk

using namespace rzk;

void bvh_build_device(rzk_ctx* ctx, int n, const float* pos, const float* tangents, const int* tri_surface,
                      rzk_scene* s) {
    RZK_CHECK(n >= 2, RZK_E_INVALID, "device build needs at least two triangles");
    hipStream_t st = ctx->stream;
    BvhHost& out = s->bvh;
    out = BvhHost();
    out.ntri = n;
    out.device_built = true;
    PinnedArray<int> hc;
    hc.reserve(2);

    Dev<float> dpos(9 * (size_t)n);
    RZK_HIP(hipMemcpyAsync(dpos.p, pos, sizeof(float) * 9 * (size_t)n, hipMemcpyHostToDevice, st));
    unsigned nb = blocks(n);
    Dev<float> tbox(6 * (size_t)n), part(6 * (size_t)nb), sbox(9);
    k_tri_boxes<<<nb, kBB, 0, st>>>(dpos.p, n, tbox.p, part.p);
    k_scene_box<<<1, kBB, 0, st>>>(part.p, (int)nb, sbox.p);
    RZK_HIP(hipGetLastError());
    {
        Dev<float> dtan(3 * (size_t)n);
        Dev<int> dts(n);
        Dev<float4> tn(n), tt(n);
        RZK_HIP(hipMemcpyAsync(dtan.p, tangents, sizeof(float) * 3 * (size_t)n, hipMemcpyHostToDevice, st));
        RZK_HIP(hipMemcpyAsync(dts.p, tri_surface, sizeof(int) * (size_t)n, hipMemcpyHostToDevice, st));
        k_tri_frames<<<nb, kBB, 0, st>>>(dpos.p, dtan.p, dts.p, n, tn.p, tt.p);
        RZK_HIP(hipGetLastError());
        RZK_HIP(hipStreamSynchronize(st));   // the host arrays may go before the copies ran
        s->d_tri_n = tn.release();
        s->d_tri_tan = tt.release();
    }

    // K1-K2: Morton codes, stable sort (equal codes keep triangle order)
    Dev<uint32_t> code(n), code_s(n);
    Dev<int> tri(n), tri_s(n);
    k_morton<<<nb, kBB, 0, st>>>(tbox.p, sbox.p, n, code.p, tri.p);
    RZK_HIP(hipGetLastError());
    size_t tmp_bytes = 0;
    RZK_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, code.p, code_s.p, tri.p, tri_s.p, n, 0, 30, st));
    Dev<char> tmp(tmp_bytes);
    RZK_HIP(rocprim::radix_sort_pairs(tmp.p, tmp_bytes, code.p, code_s.p, tri.p, tri_s.p, n, 0, 30, st));
    Dev<TriGpu> tris(n);
    k_sorted_tris<<<nb, kBB, 0, st>>>(dpos.p, tri_s.p, n, tris.p);
    RZK_HIP(hipGetLastError());

    // K3: top-down splits, one depth per launch
    Dev<Item> items(n - 1);
    Dev<int2> kid(n - 1);
    Dev<int> ctr(kMaxLevels);
    RZK_HIP(hipMemsetAsync(ctr.p, 0, sizeof(int) * kMaxLevels, st));
    Item root{0, n - 1, 0, 0};
    RZK_HIP(hipMemcpyAsync(items.p, &root, sizeof(Item), hipMemcpyHostToDevice, st));
    std::vector<int> loff, lcnt;
    int off = 0, cnt = 1;
    while (cnt > 0) {
        RZK_CHECK((int)loff.size() < kMaxLevels, RZK_E_STATE, "binary tree deeper than the build allows");
        int lv = (int)loff.size();
        loff.push_back(off);
        lcnt.push_back(cnt);
        k_split_level<<<blocks(cnt), kBB, 0, st>>>(code_s.p, items.p, off, cnt, ctr.p + lv, kid.p);
        RZK_HIP(hipGetLastError());
        RZK_HIP(hipMemcpyAsync(hc.p, ctr.p + lv, sizeof(int), hipMemcpyDeviceToHost, st));
        RZK_HIP(hipStreamSynchronize(st));
        off += cnt;
        cnt = hc.p[0];
    }
    RZK_CHECK(off == n - 1, RZK_E_STATE, "device build lost interior nodes");
    const int levels = (int)loff.size();

    // K4: boxes bottom-up (and 4-wide subtree sizes)
    Dev<float> nbox(6 * (size_t)(n - 1));
    Dev<BvhNodeGpu> nodes(n - 1);
    Dev<int> cnt4(n - 1);
    for (int lv = levels - 1; lv >= 0; --lv) {
        k_refit_level<<<blocks(lcnt[lv]), kBB, 0, st>>>(items.p, loff[lv], lcnt[lv], kid.p, tri_s.p, tbox.p, nbox.p,
                                                        nodes.p, cnt4.p);
        RZK_HIP(hipGetLastError());
    }
    RZK_HIP(hipMemcpyAsync(hc.p, cnt4.p, sizeof(int), hipMemcpyDeviceToHost, st));
    RZK_HIP(hipStreamSynchronize(st));
    const int n4 = hc.p[0];
    RZK_CHECK(n4 >= 1 && n4 <= n - 1, RZK_E_STATE, "device build miscounted 4-wide nodes");

    // K5: the 4-wide layout, one depth per launch
    Dev<Bvh4NodeGpu> nodes4(n4);
    Dev<Item4> items4(n4);
    Dev<int> ctr4(kMaxLevels);
    RZK_HIP(hipMemsetAsync(ctr4.p, 0, sizeof(int) * kMaxLevels, st));
    Item4 root4{0, 0};
    RZK_HIP(hipMemcpyAsync(items4.p, &root4, sizeof(Item4), hipMemcpyHostToDevice, st));
    int depth4 = 0;
    off = 0;
    cnt = 1;
    while (cnt > 0) {
        RZK_CHECK(depth4 < kMaxLevels, RZK_E_STATE, "4-wide tree deeper than the build allows");
        k_collapse_level<<<blocks(cnt), kBB, 0, st>>>(items4.p, off, cnt, ctr4.p + depth4, kid.p, nbox.p, tbox.p,
                                                      tri_s.p, cnt4.p, nodes4.p);
        RZK_HIP(hipGetLastError());
        RZK_HIP(hipMemcpyAsync(hc.p, ctr4.p + depth4, sizeof(int), hipMemcpyDeviceToHost, st));
        RZK_HIP(hipStreamSynchronize(st));
        ++depth4;
        off += cnt;
        cnt = hc.p[0];
    }
    RZK_CHECK(off == n4, RZK_E_STATE, "device build lost 4-wide nodes");
    Dev<TriGpu> tris4(n);
    RZK_HIP(hipMemcpyAsync(tris4.p, tris.p, sizeof(TriGpu) * (size_t)n, hipMemcpyDeviceToDevice, st));

    // what the host side keeps: scalars, the boxes, the sorted order
    out.sorted_tri.resize(n);
    out.sorted_morton.resize(n);
    RZK_HIP(hipMemcpyAsync(out.sorted_tri.data(), tri_s.p, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, st));
    RZK_HIP(hipMemcpyAsync(out.sorted_morton.data(), code_s.p, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToHost, st));
    RZK_HIP(hipMemcpyAsync(out.box, sbox.p, sizeof(float) * 6, hipMemcpyDeviceToHost, st));
    RZK_HIP(hipMemcpyAsync(out.root_box, nbox.p, sizeof(float) * 6, hipMemcpyDeviceToHost, st));
    RZK_HIP(hipStreamSynchronize(st));
    out.num_interior = n - 1;
    out.depth = levels + 1;
    out.root_ref = 0;
    out.root4_ref = 0;
    out.depth4 = depth4;
    out.stack4 = 3 * depth4 + 1;
    out.num_nodes4 = n4;
    s->d_nodes = nodes.release();
    s->d_nodes4 = nodes4.release();
    s->d_tris = tris.release();
    s->d_tris4 = tris4.release();
}
