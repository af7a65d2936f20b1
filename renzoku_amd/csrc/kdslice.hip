// kdslice.hip — LightSlice's slices from a 6-D kd-tree over the gather points
// (LightSlice::generate_slices, lightslice.cc:252-303; KdTreeN's clustering
// build and queries, kdtree_n.h:181-365, :399-429, :540-647; map_vec6,
// lightslice.cc:142-188).
//
// Tree.  Node [s, e] splits at median = (s + e) / 2 along the longest axis of
// its points' box: a quickselect of repeated stable 3-way partitions (less /
// equal / greater than a random pivot) leaves the median in place; the left
// child is [s, median], the right one [median + 1, e]; single points are
// leaves.  The first node on each path with at most max_slice_size points (and
// at least two) is a slice; slices are numbered in depth-first order.  The
// build goes on below the slices, so the points of a slice end in the order
// the full tree leaves them.  Round t of the selection at node [s, e] draws its
// pivot from Philox4x32-10 keyed by the seed with counter (s, e, t) (the
// reference draws from its global MT stream in recursion order), so every node
// of a depth runs at once.  Node ranges depend on n alone, so nodes are
// numbered as a binary heap and the slice table is laid out on the host.
//
// Device layout: points as 6 planes of n floats (plane a = coordinate a);
// `ord` = point index at each kd position.  Depths whose nodes hold more than
// kKdLds points run as rounds of two launches over all of their nodes (count,
// then a block-stable scatter), with ping-pong buffers of (index, key) and the
// positions a round settles written straight to `ord`; smaller nodes run one
// block per node with the node's indices and keys in LDS.
//
// Queries: find_cluster (descend by plane - 1e-3 to the slice node) and
// find_nearest_cluster (k nearest slice boxes, std::priority_queue order,
// furthest first) run one lane per query.
#pragma clang fp contract(off)

#include <algorithm>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "host_sampling.h"
#include "kdslice.h"
#include "rzk_device.h"
#include "std_order.h"

namespace rzk {

constexpr int kKdDim = 6;
constexpr int kKdLds = 4096;           // largest node kept in one block's LDS
constexpr int kKdChunk = 1024;         // points per block in the rounds of larger nodes
constexpr int kKdBigBlock = 256;
constexpr int kKdMaxK = 32;            // largest k of the nearest-slice query
constexpr int kKdStack = 64;           // query stack (tree depth) bound
constexpr float kKdPlaneBias = -1e-3f; // kdtree_n.h:171
constexpr float kKdFltMax = 3.402823466e+38f;

// pivot number of round t of the selection at node [s, e]
RZK_HD float kd_pivot_u(uint32_t seed, int s, int e, int t) {
    uint32_t c[4] = {(uint32_t)s, (uint32_t)e, (uint32_t)t, 0x6b64u};
    philox4x32_10(c, seed ^ 0x2545f491u, 0x6b647472u);
    return u24(c[0]);
}

// pivot position in [lo, hi]: start + (int)(rd() * (end - start + 1))
RZK_HD int kd_pivot(uint32_t seed, int s, int e, int t, int lo, int hi) {
    return lo + (int)(kd_pivot_u(seed, s, e, t) * (float)(hi - lo + 1));
}

// range of node j of depth d (heap index 2^d - 1 + j); false when an ancestor
// is a single point (the node does not exist)
__device__ __forceinline__ bool kd_range(int n, int d, int j, int& s, int& e) {
    s = 0;
    e = n - 1;
    for (int b = d - 1; b >= 0; --b) {
        if (s >= e) return false;
        const int m = (s + e) >> 1;
        if ((j >> b) & 1) s = m + 1;
        else e = m;
    }
    return true;
}

// BoundingBoxN::merge (std::min / std::max) and get_longest_axis
struct KdBox {
    float lo[kKdDim], hi[kKdDim];
    __device__ void clear() {
        for (int a = 0; a < kKdDim; ++a) { lo[a] = kKdFltMax; hi[a] = -kKdFltMax; }
    }
    __device__ void add(int a, float v) {
        lo[a] = (v < lo[a]) ? v : lo[a];
        hi[a] = (hi[a] < v) ? v : hi[a];
    }
    // merge another box's bounds (an empty box is FLT_MAX / -FLT_MAX, which
    // must not meet the other side's bound)
    __device__ void merge(int a, float olo, float ohi) {
        lo[a] = (olo < lo[a]) ? olo : lo[a];
        hi[a] = (hi[a] < ohi) ? ohi : hi[a];
    }
    __device__ int longest() const {
        int index = 0;
        float longest = 0.0f;
        for (int a = 0; a < kKdDim; ++a) {
            const float x = fabsf(hi[a] - lo[a]);
            if (x > longest) { index = a; longest = x; }
        }
        return index;
    }
};

// min / max over the block (order-free up to the sign of zero); every thread
// gets the result.  red: 12 * (NT / 64) floats of LDS.
template <int NT>
__device__ void kd_box_reduce(KdBox& b, float* red) {
    for (int a = 0; a < kKdDim; ++a)
        for (int o = 32; o > 0; o >>= 1) {
            const float olo = __shfl_xor(b.lo[a], o), ohi = __shfl_xor(b.hi[a], o);
            b.merge(a, olo, ohi);
        }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int a = 0; a < kKdDim; ++a) { red[12 * w + a] = b.lo[a]; red[12 * w + 6 + a] = b.hi[a]; }
    __syncthreads();
    b.clear();
    for (int q = 0; q < NT / 64; ++q)
        for (int a = 0; a < kKdDim; ++a) b.merge(a, red[12 * q + a], red[12 * q + 6 + a]);
    __syncthreads();
}

// exclusive block scan of v; *total = the block's sum.  sc: NT / 64 + 1 ints of LDS
template <int NT>
__device__ int kd_block_scan(int v, int* sc, int* total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) sc[w] = x;
    __syncthreads();
    int before = 0, all = 0;
    for (int q = 0; q < NT / 64; ++q) {
        const int t = sc[q];
        before += q < w ? t : 0;
        all += t;
    }
    __syncthreads();
    *total = all;
    return before + x - v;
}

// Selection rounds t0, t0 + 1, ... of node [s, e] on LDS positions [lo, hi]
// (ids and keys ping-pong between A and B) until position k holds the median:
// each round a stable 3-way partition around the round's pivot.
template <int NT>
__device__ void kd_select_lds(int* idA, float* keyA, int* idB, float* keyB, int lo, int hi, int k, uint32_t seed,
                              int s, int e, int t0, int* s_sc) {
    const int tid = (int)threadIdx.x;
    for (int t = t0;; ++t) {
        const int len = hi - lo + 1;
        const float v = keyA[kd_pivot(seed, s, e, t, lo, hi)];
        const int per = (len + NT - 1) / NT;
        const int b0 = lo + tid * per, b1 = min(hi + 1, b0 + per);
        int lt = 0, eq = 0;
        for (int i = b0; i < b1; ++i) {
            const float x = keyA[i];
            lt += x < v;
            eq += x == v;
        }
        int tot;
        const int pre = kd_block_scan<NT>(lt | (eq << 16), s_sc, &tot);
        const int LT = tot & 0xffff, EQ = tot >> 16;
        int pl = lo + (pre & 0xffff), pe = lo + LT + (pre >> 16);
        int pg = lo + LT + EQ + ((b0 - lo) - (pre & 0xffff) - (pre >> 16));
        for (int i = b0; i < b1; ++i) {
            const float x = keyA[i];
            const int pos = x < v ? pl++ : (x == v ? pe++ : pg++);
            idB[pos] = idA[i];
            keyB[pos] = x;
        }
        __syncthreads();
        for (int i = lo + tid; i <= hi; i += NT) { idA[i] = idB[i]; keyA[i] = keyB[i]; }
        __syncthreads();
        if (k < lo + LT) hi = lo + LT - 1;
        else if (k < lo + LT + EQ) break;
        else lo = lo + LT + EQ;
    }
}

// ---- depths with nodes of at most kKdLds points: one block per node --------
// LDS: ids and keys, twice (ping-pong), cap entries each
template <int NT>
__global__ __launch_bounds__(NT) void k_kd_level_lds(const float* __restrict__ P, int n, int d, int cap, uint32_t seed,
                                                     int* __restrict__ ord, int* __restrict__ axis_out,
                                                     float* __restrict__ plane_out, int H, const int* __restrict__ cl,
                                                     float* __restrict__ boxes) {
    extern __shared__ int s_kd[];
    __shared__ float s_red[12 * (NT / 64)];
    __shared__ int s_sc[NT / 64 + 1];
    int* idA = s_kd;
    float* keyA = reinterpret_cast<float*>(s_kd + cap);
    int* idB = s_kd + 2 * cap;
    float* keyB = reinterpret_cast<float*>(s_kd + 3 * cap);
    const int j = (int)blockIdx.x, tid = (int)threadIdx.x;
    int s, e;
    if (!kd_range(n, d, j, s, e) || s >= e) return;
    const int m = e - s + 1;
    const int heap = (1 << d) - 1 + j;
    KdBox b;
    b.clear();
    for (int i = tid; i < m; i += NT) {
        const int id = ord[s + i];
        idA[i] = id;
        for (int a = 0; a < kKdDim; ++a) b.add(a, P[(int64_t)a * n + id]);
    }
    kd_box_reduce<NT>(b, s_red);
    const int ax = b.longest();
    if (heap < H && cl[heap] >= 0 && tid < 12) boxes[12 * cl[heap] + tid] = tid < 6 ? b.lo[tid] : b.hi[tid - 6];
    for (int i = tid; i < m; i += NT) keyA[i] = P[(int64_t)ax * n + idA[i]];
    __syncthreads();
    const int k = (m - 1) >> 1;              // (s + e) / 2 - s
    kd_select_lds<NT>(idA, keyA, idB, keyB, 0, m - 1, k, seed, s, e, 0, s_sc);
    if (tid == 0 && heap < H) { axis_out[heap] = ax; plane_out[heap] = keyA[k]; }
    for (int i = tid; i < m; i += NT) ord[s + i] = idA[i];
}

// ---- depths with larger nodes: rounds over all nodes of the depth ----------
// state per node and round parity: {lo, hi, done, unused}
struct KdBig {
    const float* P;
    int n, d, N, nb;            // nodes at this depth, blocks per node
    uint32_t seed;
    int* ord;
    int *idA, *idB;
    float *keyA, *keyB;
    int4* st;                   // 2 * N
    int2* cnt;                  // N * nb
    float* part;                // N * nb * 12
    int* axis_lvl;              // N
    int* axis_out;
    float* plane_out;
    int H;
    const int* cl;
    float* boxes;
};

__global__ __launch_bounds__(kKdBigBlock) void k_kd_big_part(KdBig g) {
    __shared__ float s_red[12 * (kKdBigBlock / 64)];
    const int j = (int)blockIdx.y, bk = (int)blockIdx.x;
    int s, e;
    if (!kd_range(g.n, g.d, j, s, e) || s >= e) return;
    KdBox b;
    b.clear();
    const int c0 = s + bk * kKdChunk, c1 = min(e, c0 + kKdChunk - 1);
    for (int i = c0 + (int)threadIdx.x; i <= c1; i += kKdBigBlock) {
        const int id = g.ord[i];
        for (int a = 0; a < kKdDim; ++a) b.add(a, g.P[(int64_t)a * g.n + id]);
    }
    kd_box_reduce<kKdBigBlock>(b, s_red);
    if (threadIdx.x < 12)
        g.part[((int64_t)j * g.nb + bk) * 12 + threadIdx.x] = threadIdx.x < 6 ? b.lo[threadIdx.x] : b.hi[threadIdx.x - 6];
}

// one wave per node: the node's box from the block partials, its axis, the
// slice box, the first round's state
__global__ __launch_bounds__(64) void k_kd_big_axis(KdBig g) {
    const int j = (int)blockIdx.x;
    int s, e;
    if (!kd_range(g.n, g.d, j, s, e) || s >= e) {
        if (threadIdx.x == 0) { g.st[j] = make_int4(0, -1, 1, 0); g.st[g.N + j] = make_int4(0, -1, 1, 0); }
        return;
    }
    KdBox b;
    b.clear();
    const int nact = (e - s + kKdChunk) / kKdChunk;
    for (int q = (int)threadIdx.x; q < nact; q += 64)
        for (int a = 0; a < kKdDim; ++a)
            b.merge(a, g.part[((int64_t)j * g.nb + q) * 12 + a], g.part[((int64_t)j * g.nb + q) * 12 + 6 + a]);
    for (int a = 0; a < kKdDim; ++a)
        for (int o = 32; o > 0; o >>= 1) {
            const float olo = __shfl_xor(b.lo[a], o), ohi = __shfl_xor(b.hi[a], o);
            b.merge(a, olo, ohi);
        }
    const int heap = (1 << g.d) - 1 + j;
    if (heap < g.H && g.cl[heap] >= 0 && threadIdx.x < 12)
        g.boxes[12 * g.cl[heap] + threadIdx.x] = threadIdx.x < 6 ? b.lo[threadIdx.x] : b.hi[threadIdx.x - 6];
    if (threadIdx.x == 0) {
        const int ax = b.longest();
        g.axis_lvl[j] = ax;
        if (heap < g.H) g.axis_out[heap] = ax;
        g.st[j] = make_int4(s, e, 0, 0);
    }
}

__global__ __launch_bounds__(kKdBigBlock) void k_kd_big_keys(KdBig g) {
    const int j = (int)blockIdx.y, bk = (int)blockIdx.x;
    int s, e;
    if (!kd_range(g.n, g.d, j, s, e) || s >= e) return;
    const int ax = g.axis_lvl[j];
    const int c0 = s + bk * kKdChunk, c1 = min(e, c0 + kKdChunk - 1);
    for (int i = c0 + (int)threadIdx.x; i <= c1; i += kKdBigBlock) {
        const int id = g.ord[i];
        g.idA[i] = id;
        g.keyA[i] = g.P[(int64_t)ax * g.n + id];
    }
}

// round r, pass 1: (less, equal) counts of each chunk of the active range
__global__ __launch_bounds__(kKdBigBlock) void k_kd_big_count(KdBig g, int r) {
    __shared__ int s_sc[kKdBigBlock / 64 + 1];
    const int j = (int)blockIdx.y, bk = (int)blockIdx.x;
    int s, e;
    if (!kd_range(g.n, g.d, j, s, e) || s >= e) return;
    const int4 S = g.st[(r & 1) * g.N + j];
    if (S.z) return;
    const int c0 = S.x + bk * kKdChunk;
    if (c0 > S.y) return;
    const int c1 = min(S.y, c0 + kKdChunk - 1);
    const float* key = (r & 1) ? g.keyB : g.keyA;
    const float v = key[kd_pivot(g.seed, s, e, r, S.x, S.y)];
    int lt = 0, eq = 0;
    for (int i = c0 + (int)threadIdx.x; i <= c1; i += kKdBigBlock) {
        const float x = key[i];
        lt += x < v;
        eq += x == v;
    }
    int tot;
    kd_block_scan<kKdBigBlock>(lt | (eq << 16), s_sc, &tot);
    if (threadIdx.x == 0) g.cnt[(int64_t)j * g.nb + bk] = make_int2(tot & 0xffff, tot >> 16);
}

// round r, pass 2: stable scatter of each chunk into the other buffer; the
// positions that leave the active range (all of them when the median is
// reached) are final and also go to ord; block 0 writes the next state
__global__ __launch_bounds__(kKdBigBlock) void k_kd_big_scatter(KdBig g, int r) {
    __shared__ int s_sc[kKdBigBlock / 64 + 1];
    __shared__ int s_tot[4][kKdBigBlock / 64];
    const int j = (int)blockIdx.y, bk = (int)blockIdx.x, tid = (int)threadIdx.x;
    int s, e;
    if (!kd_range(g.n, g.d, j, s, e) || s >= e) return;
    const int4 S = g.st[(r & 1) * g.N + j];
    int4* nxt = g.st + ((r + 1) & 1) * g.N + j;
    if (S.z) {
        if (bk == 0 && tid == 0) *nxt = S;
        return;
    }
    const int c0 = S.x + bk * kKdChunk;
    if (c0 > S.y) return;
    const int c1 = min(S.y, c0 + kKdChunk - 1);
    const int nact = (S.y - S.x + kKdChunk) / kKdChunk;
    // chunk counts: before this chunk and in all
    int lb = 0, eb = 0, la = 0, ea = 0;
    for (int q = tid; q < nact; q += kKdBigBlock) {
        const int2 c = g.cnt[(int64_t)j * g.nb + q];
        la += c.x; ea += c.y;
        if (q < bk) { lb += c.x; eb += c.y; }
    }
    for (int o = 32; o > 0; o >>= 1) {
        la += __shfl_xor(la, o); ea += __shfl_xor(ea, o);
        lb += __shfl_xor(lb, o); eb += __shfl_xor(eb, o);
    }
    if ((tid & 63) == 0) { s_tot[0][tid >> 6] = la; s_tot[1][tid >> 6] = ea; s_tot[2][tid >> 6] = lb; s_tot[3][tid >> 6] = eb; }
    __syncthreads();
    int LT = 0, EQ = 0, LB = 0, EB = 0;
    for (int q = 0; q < kKdBigBlock / 64; ++q) { LT += s_tot[0][q]; EQ += s_tot[1][q]; LB += s_tot[2][q]; EB += s_tot[3][q]; }
    const int GB = (c0 - S.x) - LB - EB;
    const int k = (s + e) >> 1;
    int4 ns;
    if (k < S.x + LT) ns = make_int4(S.x, S.x + LT - 1, 0, 0);
    else if (k < S.x + LT + EQ) ns = make_int4(S.x, S.y, 1, 0);
    else ns = make_int4(S.x + LT + EQ, S.y, 0, 0);
    const int* id = (r & 1) ? g.idB : g.idA;
    const float* key = (r & 1) ? g.keyB : g.keyA;
    int* id2 = (r & 1) ? g.idA : g.idB;
    float* key2 = (r & 1) ? g.keyA : g.keyB;
    const float v = key[kd_pivot(g.seed, s, e, r, S.x, S.y)];
    // block-stable order: thread t takes kKdChunk / kKdBigBlock consecutive points
    constexpr int kPer = kKdChunk / kKdBigBlock;
    const int b0 = c0 + tid * kPer, b1 = min(c1 + 1, b0 + kPer);
    int lt = 0, eq = 0;
    for (int i = b0; i < b1; ++i) {
        const float x = key[i];
        lt += x < v;
        eq += x == v;
    }
    int tot;
    const int pre = kd_block_scan<kKdBigBlock>(lt | (eq << 16), s_sc, &tot);
    int pl = S.x + LB + (pre & 0xffff);
    int pe = S.x + LT + EB + (pre >> 16);
    int pg = S.x + LT + EQ + GB + ((b0 - c0) - (pre & 0xffff) - (pre >> 16));
    for (int i = b0; i < b1; ++i) {
        const float x = key[i];
        const int pid = id[i];
        const int pos = x < v ? pl++ : (x == v ? pe++ : pg++);
        id2[pos] = pid;
        key2[pos] = x;
        if (ns.z || pos < ns.x || pos > ns.y) g.ord[pos] = pid;
        if (ns.z && pos == k) {
            const int heap = (1 << g.d) - 1 + j;
            if (heap < g.H) g.plane_out[heap] = x;
        }
    }
    if (bk == 0 && tid == 0) *nxt = ns;
}

// the rest of a depth's selection once every active range fits in LDS: one
// block per node continues from round r in LDS and writes the range's final
// order to ord and the plane
__global__ __launch_bounds__(kKdBigBlock) void k_kd_big_finish(KdBig g, int r) {
    extern __shared__ int s_fin[];
    __shared__ int s_sc[kKdBigBlock / 64 + 1];
    const int j = (int)blockIdx.x, tid = (int)threadIdx.x;
    int s, e;
    if (!kd_range(g.n, g.d, j, s, e) || s >= e) return;
    const int4 S = g.st[(r & 1) * g.N + j];
    if (S.z) return;
    const int len = S.y - S.x + 1;
    int* idA = s_fin;
    float* keyA = reinterpret_cast<float*>(s_fin + kKdLds);
    int* idB = s_fin + 2 * kKdLds;
    float* keyB = reinterpret_cast<float*>(s_fin + 3 * kKdLds);
    const int* id = (r & 1) ? g.idB : g.idA;
    const float* key = (r & 1) ? g.keyB : g.keyA;
    for (int i = tid; i < len; i += kKdBigBlock) { idA[i] = id[S.x + i]; keyA[i] = key[S.x + i]; }
    __syncthreads();
    const int k = ((s + e) >> 1) - S.x;
    kd_select_lds<kKdBigBlock>(idA, keyA, idB, keyB, 0, len - 1, k, g.seed, s, e, r, s_sc);
    for (int i = tid; i < len; i += kKdBigBlock) g.ord[S.x + i] = idA[i];
    const int heap = (1 << g.d) - 1 + j;
    if (tid == 0 && heap < g.H) g.plane_out[heap] = keyA[k];
}

// ---- queries ------------------------------------------------------------------
__device__ __forceinline__ float kd_box_dist2(const float* __restrict__ box, const float* p) {
    float sum = 0.0f;
    for (int a = 0; a < kKdDim; ++a) {
        float c = (p[a] < box[a]) ? box[a] : p[a];
        c = (box[6 + a] < c) ? box[6 + a] : c;
        const float dd = c - p[a];
        sum += dd * dd;
    }
    return sum;
}

struct KdQuery {
    const int* axis;
    const float* plane;
    const int* cl;
    int H;
    const float* boxes;
};

// find_cluster_rec
__device__ __forceinline__ int kd_find(const KdQuery& t, const float* p) {
    int node = 0;
    while (node < t.H) {
        const int c = t.cl[node];
        if (c >= 0) return c;
        const float dist = p[t.axis[node]] - t.plane[node] + kKdPlaneBias;
        node = dist <= 0.0f ? 2 * node + 1 : 2 * node + 2;
    }
    return -1;
}

__global__ void k_kd_find(KdQuery t, const float* __restrict__ q, int m, int* __restrict__ out) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= m) return;
    float p[kKdDim];
    for (int a = 0; a < kKdDim; ++a) p[a] = q[(int64_t)kKdDim * i + a];
    out[i] = kd_find(t, p);
}

// the slice of every point i of the planes, written to out[dst[i]]
__global__ void k_kd_find_points(KdQuery t, const float* __restrict__ P, int n, const int* __restrict__ dst,
                                 int* __restrict__ out) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    float p[kKdDim];
    for (int a = 0; a < kKdDim; ++a) p[a] = P[(int64_t)a * n + i];
    out[dst[i]] = kd_find(t, p);
}

// find_nearest_cluster_rec with an explicit stack of (second child, plane
// distance); the heap is std::priority_queue's (push_heap / pop_heap moves)
// queries: q (m x 6), or when pos is set the points at kd positions pos[i]
// (planes P of n floats, ord = the point at each kd position)
__global__ __launch_bounds__(64) void k_kd_nearest(KdQuery t, const float* __restrict__ q, const float* __restrict__ P,
                                                   int n, const int* __restrict__ ord, const int* __restrict__ pos,
                                                   int m, int k, int* __restrict__ out) {
    __shared__ float s_hk[64 * kKdMaxK];
    __shared__ int s_hv[64 * kKdMaxK];
    __shared__ int s_sn[64 * kKdStack];
    __shared__ float s_sd[64 * kKdStack];
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= m) return;
    float* hk = s_hk + threadIdx.x * kKdMaxK;
    int* hv = s_hv + threadIdx.x * kKdMaxK;
    int* sn = s_sn + threadIdx.x * kKdStack;
    float* sd = s_sd + threadIdx.x * kKdStack;
    float p[kKdDim];
    if (pos) {
        const int id = ord[pos[i]];
        for (int a = 0; a < kKdDim; ++a) p[a] = P[(int64_t)a * n + id];
    } else {
        for (int a = 0; a < kKdDim; ++a) p[a] = q[(int64_t)kKdDim * i + a];
    }
    int hn = 0, sp = 0, node = 0;
    bool ok = t.H > 0;
    while (ok) {
        // descend to a slice node, deferring the far side
        while (node < t.H && t.cl[node] < 0 && sp < kKdStack) {
            const float dist = p[t.axis[node]] - t.plane[node] + kKdPlaneBias;
            const int first = dist <= 0.0f ? 2 * node + 1 : 2 * node + 2;
            sn[sp] = dist <= 0.0f ? 2 * node + 2 : 2 * node + 1;
            sd[sp] = dist;
            ++sp;
            node = first;
        }
        if (node < t.H && t.cl[node] >= 0) {
            const int c = t.cl[node];
            const float dc = kd_box_dist2(t.boxes + 12 * c, p);
            if (hn < k) {
                stdord::heap_push(hk, hv, hn, dc, c);
            } else if (dc < hk[0]) {
                stdord::heap_pop(hk, hv, hn);
                stdord::heap_push(hk, hv, hn, dc, c);
            }
        }
        ok = false;
        while (sp > 0) {
            --sp;
            const float dist = sd[sp];
            if (hn < k || hk[0] > dist * dist) {
                node = sn[sp];
                ok = true;
                break;
            }
        }
    }
    int o = 0;
    while (hn > 0) {
        out[(int64_t)k * i + o++] = hv[0];
        stdord::heap_pop(hk, hv, hn);
    }
    for (; o < k; ++o) out[(int64_t)k * i + o] = -1;
}

// map_vec6 of receiver records (one lane per record), into 6 planes of n floats
__global__ void k_map_vec6(const float* __restrict__ recv, const int* __restrict__ rmat, const int* __restrict__ idx,
                           const float* __restrict__ mat, int64_t n, float diagonal, float* __restrict__ out,
                           int64_t pitch) {
    const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n) return;
    const int64_t i = idx ? (int64_t)idx[o] : o;           // record of output point o
    const float* rc = recv + RZK_RECV_FLOATS * i;
    const float weight = 1.0f / 16.0f * diagonal;
    const float nx = rc[6], ny = rc[7], nz = rc[8], wx = rc[9], wy = rc[10], wz = rc[11];
    float dx = nx, dy = ny, dz = nz;
    const int m = rmat[i];
    const int type = m >= 0 ? (int)mat[RZK_MAT_FLOATS * m] : 0;
    if (type == kPhong || type == kWard) {
        const float* mm = mat + RZK_MAT_FLOATS * m;
        float gexp;
        if (type == kPhong) {
            gexp = mm[7];
        } else {
            const float ix = 1.0f / mm[7], iy = 1.0f / mm[8];
            gexp = (ix < iy) ? iy : ix;
        }
        const float ex = gexp / 1000.0f;
        const float mx = (1.0f < ex) ? ex : 1.0f;
        const float gs = (mx < 4.0f) ? mx : 4.0f;
        const float d2 = 2.0f * (wx * nx + wy * ny + wz * nz);
        dx = (d2 * nx - wx) * gs;
        dy = (d2 * ny - wy) * gs;
        dz = (d2 * nz - wz) * gs;
    }
    out[o] = rc[0];
    out[pitch + o] = rc[1];
    out[2 * pitch + o] = rc[2];
    out[3 * pitch + o] = dx * weight;
    out[4 * pitch + o] = dy * weight;
    out[5 * pitch + o] = dz * weight;
}

// AoS (n x 6) to 6 planes
__global__ void k_kd_planes(const float* __restrict__ in, int n, float* __restrict__ out) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    for (int a = 0; a < kKdDim; ++a) out[(int64_t)a * n + i] = in[(int64_t)kKdDim * i + a];
}

__global__ void k_kd_iota(int* __restrict__ ord, int n) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) ord[i] = i;
}

template <typename T>
__global__ void k_kd_fill(T* __restrict__ p, int n, T v) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) p[i] = v;
}

}  // namespace rzk

using namespace rzk;

namespace {

template <typename F>
int kd_guard(rzk_ctx* ctx, F&& f) {
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

// slice nodes in depth-first order (heap index, first, last); node ranges
// depend on n alone
void kd_topology(int64_t heap, int s, int e, int max_size, std::vector<int64_t>& hs, std::vector<int>& out) {
    if (s >= e) return;
    if (e - s + 1 <= max_size) {
        hs.push_back(heap);
        out.push_back(s);
        out.push_back(e);
        return;
    }
    const int m = (s + e) / 2;
    kd_topology(2 * heap + 1, s, m, max_size, hs, out);
    kd_topology(2 * heap + 2, m + 1, e, max_size, hs, out);
}

unsigned kd_blocks(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

// Build over points already on the device as 6 planes of n floats (kd->P).
void kd_build_device(rzk_kdslices* kd) {
    rzk_ctx* ctx = kd->ctx;
    hipStream_t st = ctx->stream;
    const int n = kd->n;
    std::vector<int64_t> hs;
    kd->h_slices.clear();
    if (n >= 2) kd_topology(0, 0, n - 1, kd->max_size, hs, kd->h_slices);
    kd->S = (int)hs.size();
    int64_t H = 0;
    for (int64_t h : hs) H = std::max(H, h + 1);
    RZK_CHECK(H <= (int64_t)1 << 30, RZK_E_INVALID, "kd-tree too deep for its slice size");
    kd->H = (int)H;
    std::vector<int> h_cl((size_t)std::max<int64_t>(H, 1), -1);
    for (int s = 0; s < kd->S; ++s) h_cl[(size_t)hs[s]] = s;
    kd->ord.alloc(n);
    kd->axis.alloc(H);
    kd->plane.alloc(H);
    kd->cl.alloc(H);
    kd->boxes.alloc((size_t)12 * kd->S);
    kd->slices.alloc((size_t)2 * kd->S);
    if (H > 0) {
        hipLaunchKernelGGL(k_kd_fill<int>, dim3(kd_blocks(H, 256)), dim3(256), 0, st, kd->axis.p, (int)H, 127);
        hipLaunchKernelGGL(k_kd_fill<float>, dim3(kd_blocks(H, 256)), dim3(256), 0, st, kd->plane.p, (int)H, kKdFltMax);
        RZK_HIP(hipMemcpyAsync(kd->cl.p, h_cl.data(), sizeof(int) * (size_t)H, hipMemcpyHostToDevice, st));
    }
    if (kd->S > 0)
        RZK_HIP(hipMemcpyAsync(kd->slices.p, kd->h_slices.data(), sizeof(int) * 2 * (size_t)kd->S,
                               hipMemcpyHostToDevice, st));
    if (n > 0) hipLaunchKernelGGL(k_kd_iota, dim3(kd_blocks(n, 256)), dim3(256), 0, st, kd->ord.p, n);
    RZK_HIP(hipGetLastError());
    // largest node of each depth: the leftmost (sizes of a depth differ by at most one)
    std::vector<int> maxm;
    for (int m = n; m >= 2; m = (m + 1) / 2) maxm.push_back(m);
    // scratch of the large-node rounds
    int big_depths = 0;
    while (big_depths < (int)maxm.size() && maxm[big_depths] > kKdLds) ++big_depths;
    KdBuf<int> idA, idB, axis_lvl;
    KdBuf<float> keyA, keyB, part;
    KdBuf<int4> stbuf;
    KdBuf<int2> cnt;
    if (big_depths > 0) {
        const int Nmax = 1 << (big_depths - 1);
        int64_t cells = 0;                   // most (node, chunk) blocks of a depth
        for (int d = 0; d < big_depths; ++d)
            cells = std::max<int64_t>(cells, ((int64_t)1 << d) * ((maxm[d] + kKdChunk - 1) / kKdChunk));
        idA.alloc(n); idB.alloc(n); keyA.alloc(n); keyB.alloc(n);
        part.alloc((size_t)cells * 12);
        cnt.alloc((size_t)cells);
        stbuf.alloc((size_t)2 * Nmax);
        axis_lvl.alloc(Nmax);
    }
    std::vector<int4> h_st;
    if (big_depths > 0)
        RZK_HIP(hipFuncSetAttribute((const void*)k_kd_big_finish, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)(sizeof(int) * 4 * kKdLds)));
    for (int d = 0; d < (int)maxm.size(); ++d) {
        const int m = maxm[d];
        if (m > kKdLds) {
            KdBig g;
            g.P = kd->P.p; g.n = n; g.d = d; g.N = 1 << d; g.nb = (m + kKdChunk - 1) / kKdChunk;
            g.seed = kd->seed; g.ord = kd->ord.p;
            g.idA = idA.p; g.idB = idB.p; g.keyA = keyA.p; g.keyB = keyB.p;
            g.st = stbuf.p; g.cnt = cnt.p; g.part = part.p; g.axis_lvl = axis_lvl.p;
            g.axis_out = kd->axis.p; g.plane_out = kd->plane.p; g.H = kd->H; g.cl = kd->cl.p; g.boxes = kd->boxes.p;
            const dim3 grid(g.nb, g.N);
            hipLaunchKernelGGL(k_kd_big_part, grid, dim3(kKdBigBlock), 0, st, g);
            hipLaunchKernelGGL(k_kd_big_axis, dim3(g.N), dim3(64), 0, st, g);
            hipLaunchKernelGGL(k_kd_big_keys, grid, dim3(kKdBigBlock), 0, st, g);
            RZK_HIP(hipGetLastError());
            // rounds over HBM until every active range fits in LDS, then one
            // launch finishes the selections there; every round takes the pivot
            // out of the active range, so at most m rounds
            h_st.resize(g.N);
            bool settled = false;
            for (int r = 0; r < m && !settled; ++r) {
                hipLaunchKernelGGL(k_kd_big_count, grid, dim3(kKdBigBlock), 0, st, g, r);
                hipLaunchKernelGGL(k_kd_big_scatter, grid, dim3(kKdBigBlock), 0, st, g, r);
                RZK_HIP(hipGetLastError());
                if ((r & 1) == 1 || r + 1 == m) {
                    RZK_HIP(hipMemcpyAsync(h_st.data(), g.st + ((r + 1) & 1) * g.N, sizeof(int4) * g.N,
                                           hipMemcpyDeviceToHost, st));
                    RZK_HIP(hipStreamSynchronize(st));
                    bool all_done = true, fits = true;
                    for (int q = 0; q < g.N; ++q) {
                        all_done = all_done && h_st[q].z != 0;
                        fits = fits && (h_st[q].z != 0 || h_st[q].y - h_st[q].x + 1 <= kKdLds);
                    }
                    if (!all_done && fits) {
                        hipLaunchKernelGGL(k_kd_big_finish, dim3(g.N), dim3(kKdBigBlock), sizeof(int) * 4 * kKdLds,
                                           st, g, r + 1);
                        RZK_HIP(hipGetLastError());
                    }
                    settled = all_done || fits;
                }
            }
            RZK_CHECK(settled, RZK_E_STATE, "kd-tree selection did not settle");
        } else {
            const size_t lds = sizeof(int) * 4 * (size_t)m;
            const unsigned grid = 1u << d;
            if (m > 256) {
                RZK_HIP(hipFuncSetAttribute((const void*)k_kd_level_lds<256>,
                                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                hipLaunchKernelGGL(k_kd_level_lds<256>, dim3(grid), dim3(256), lds, st, kd->P.p, n, d, m, kd->seed,
                                   kd->ord.p, kd->axis.p, kd->plane.p, kd->H, kd->cl.p, kd->boxes.p);
            } else {
                hipLaunchKernelGGL(k_kd_level_lds<64>, dim3(grid), dim3(64), lds, st, kd->P.p, n, d, m, kd->seed,
                                   kd->ord.p, kd->axis.p, kd->plane.p, kd->H, kd->cl.p, kd->boxes.p);
            }
            RZK_HIP(hipGetLastError());
        }
    }
    RZK_HIP(hipStreamSynchronize(st));
}

static KdQuery kd_query(const rzk_kdslices* kd) {
    return KdQuery{kd->axis.p, kd->plane.p, kd->cl.p, kd->H, kd->boxes.p};
}

void kd_launch_map_vec6(rzk_ctx* ctx, const float* d_mat, const float* d_recv, const int* d_rmat, const int* d_idx,
                        int np, float diagonal, float* d_planes) {
    if (np <= 0) return;
    hipLaunchKernelGGL(k_map_vec6, dim3(kd_blocks(np, 256)), dim3(256), 0, ctx->stream, d_recv, d_rmat, d_idx, d_mat,
                       (int64_t)np, diagonal, d_planes, (int64_t)np);
    RZK_HIP(hipGetLastError());
}

void kd_launch_nearest_at(const rzk_kdslices* kd, const int* d_pos, int m, int k, int* d_out) {
    RZK_CHECK(k >= 1 && k <= kKdMaxK, RZK_E_INVALID, "k must be in 1 .. " + std::to_string(kKdMaxK));
    if (m <= 0) return;
    hipLaunchKernelGGL(k_kd_nearest, dim3(kd_blocks(m, 64)), dim3(64), 0, kd->ctx->stream, kd_query(kd),
                       (const float*)nullptr, kd->P.p, kd->n, kd->ord.p, d_pos, m, k, d_out);
    RZK_HIP(hipGetLastError());
}

void kd_launch_find_points(const rzk_kdslices* kd, const int* d_dst, int* d_out) {
    if (kd->n <= 0) return;
    hipLaunchKernelGGL(k_kd_find_points, dim3(kd_blocks(kd->n, 256)), dim3(256), 0, kd->ctx->stream, kd_query(kd),
                       kd->P.p, kd->n, d_dst, d_out);
    RZK_HIP(hipGetLastError());
}

extern "C" {

int rzk_kdslices_build_mt(rzk_ctx* ctx, const float* points, int32_t n, int32_t max_slice_size, uint32_t seed,
                          int64_t* out_draws, rzk_kdslices** out) {
    if (!ctx || !out) return RZK_E_INVALID;
    *out = nullptr;
    return kd_guard(ctx, [&] {
        RZK_CHECK(n >= 0 && (n == 0 || points), RZK_E_INVALID, "null points");
        RZK_CHECK(max_slice_size >= 3, RZK_E_INVALID, "max_slice_size must be at least 3");
        RZK_CHECK(n < (1 << 30), RZK_E_INVALID, "too many points");
        std::unique_ptr<rzk_kdslices> kd(new rzk_kdslices());
        kd->ctx = ctx;
        kd->n = n;
        kd->max_size = max_slice_size;
        kd->seed = seed;
        std::vector<float> planes((size_t)kKdDim * n);
        for (int i = 0; i < n; ++i)
            for (int a = 0; a < kKdDim; ++a) planes[(size_t)a * n + i] = points[(size_t)kKdDim * i + a];
        kd->P.alloc((size_t)kKdDim * n);
        if (n > 0)
            RZK_HIP(hipMemcpyAsync(kd->P.p, planes.data(), sizeof(float) * planes.size(), hipMemcpyHostToDevice,
                                   ctx->stream));
        Mt19937 rd(seed);
        const int64_t draws = kd_build_host_mt(kd.get(), planes.data(), rd);
        if (out_draws) *out_draws = draws;
        ctx_retain(ctx);
        *out = kd.release();
    });
}

static int kd_build_entry(rzk_ctx* ctx, const float* points, bool on_device, int32_t n, int32_t max_slice_size,
                          uint32_t seed, rzk_kdslices** out) {
    if (!ctx || !out) return RZK_E_INVALID;
    *out = nullptr;
    return kd_guard(ctx, [&] {
        RZK_CHECK(n >= 0 && (n == 0 || points), RZK_E_INVALID, "null points");
        RZK_CHECK(max_slice_size >= 3, RZK_E_INVALID, "max_slice_size must be at least 3");
        RZK_CHECK(n < (1 << 30), RZK_E_INVALID, "too many points");
        std::unique_ptr<rzk_kdslices> kd(new rzk_kdslices());
        kd->ctx = ctx;
        kd->n = n;
        kd->max_size = max_slice_size;
        kd->seed = seed;
        kd->P.alloc((size_t)kKdDim * n);
        if (n > 0) {
            KdBuf<float> aos;
            const float* src = points;
            if (!on_device) {
                aos.alloc((size_t)kKdDim * n);
                RZK_HIP(hipMemcpyAsync(aos.p, points, sizeof(float) * kKdDim * (size_t)n, hipMemcpyHostToDevice,
                                       ctx->stream));
                src = aos.p;
            }
            hipLaunchKernelGGL(k_kd_planes, dim3(kd_blocks(n, 256)), dim3(256), 0, ctx->stream, src, n, kd->P.p);
            RZK_HIP(hipGetLastError());
        }
        kd_build_device(kd.get());
        ctx_retain(ctx);
        *out = kd.release();
    });
}

int rzk_kdslices_build(rzk_ctx* ctx, const float* points, int32_t n, int32_t max_slice_size, uint32_t seed,
                       rzk_kdslices** out) {
    return kd_build_entry(ctx, points, false, n, max_slice_size, seed, out);
}

int rzk_kdslices_build_dev(rzk_ctx* ctx, const float* d_points, int32_t n, int32_t max_slice_size, uint32_t seed,
                           rzk_kdslices** out) {
    return kd_build_entry(ctx, d_points, true, n, max_slice_size, seed, out);
}

int rzk_kdslices_destroy(rzk_kdslices* kd) {
    if (!kd) return RZK_OK;
    rzk_ctx* ctx = kd->ctx;
    (void)hipStreamSynchronize(ctx->stream);
    delete kd;
    ctx_release(ctx);
    return RZK_OK;
}

int rzk_kdslices_info(const rzk_kdslices* kd, int32_t* out_slices, int32_t* out_heap) {
    if (!kd) return RZK_E_INVALID;
    if (out_slices) *out_slices = kd->S;
    if (out_heap) *out_heap = kd->H;
    return RZK_OK;
}

int rzk_kdslices_get(const rzk_kdslices* kd, const char* name, void* dst, int64_t bytes) {
    if (!kd || !name) return RZK_E_INVALID;
    rzk_ctx* ctx = kd->ctx;
    return kd_guard(ctx, [&] {
        const void* src = nullptr;
        int64_t need = 0;
        const std::string nm(name);
        if (nm == "order") { src = kd->ord.p; need = 4LL * kd->n; }
        else if (nm == "slices") { src = kd->slices.p; need = 8LL * kd->S; }
        else if (nm == "boxes") { src = kd->boxes.p; need = 48LL * kd->S; }
        else if (nm == "axis") { src = kd->axis.p; need = 4LL * kd->H; }
        else if (nm == "plane") { src = kd->plane.p; need = 4LL * kd->H; }
        else if (nm == "cl") { src = kd->cl.p; need = 4LL * kd->H; }
        else RZK_CHECK(false, RZK_E_INVALID, "unknown kd-slice array '" + nm + "'");
        RZK_CHECK(bytes == need && (need == 0 || dst), RZK_E_INVALID,
                  "kd-slice array '" + nm + "' holds " + std::to_string(need) + " bytes");
        if (need > 0) {
            RZK_HIP(hipMemcpyAsync(dst, src, (size_t)need, hipMemcpyDeviceToHost, ctx->stream));
            RZK_HIP(hipStreamSynchronize(ctx->stream));
        }
    });
}

int rzk_kdslices_find(const rzk_kdslices* kd, const float* queries, int32_t m, int32_t* out) {
    if (!kd) return RZK_E_INVALID;
    rzk_ctx* ctx = kd->ctx;
    return kd_guard(ctx, [&] {
        RZK_CHECK(m >= 0 && (m == 0 || (queries && out)), RZK_E_INVALID, "null argument");
        if (m == 0) return;
        KdBuf<float> q;
        KdBuf<int> o;
        q.alloc((size_t)kKdDim * m);
        o.alloc(m);
        RZK_HIP(hipMemcpyAsync(q.p, queries, sizeof(float) * kKdDim * (size_t)m, hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(k_kd_find, dim3(kd_blocks(m, 256)), dim3(256), 0, ctx->stream, kd_query(kd), q.p, m, o.p);
        RZK_HIP(hipGetLastError());
        RZK_HIP(hipMemcpyAsync(out, o.p, sizeof(int) * (size_t)m, hipMemcpyDeviceToHost, ctx->stream));
        RZK_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int rzk_kdslices_nearest(const rzk_kdslices* kd, const float* queries, int32_t m, int32_t k, int32_t* out) {
    if (!kd) return RZK_E_INVALID;
    rzk_ctx* ctx = kd->ctx;
    return kd_guard(ctx, [&] {
        RZK_CHECK(m >= 0 && (m == 0 || (queries && out)), RZK_E_INVALID, "null argument");
        RZK_CHECK(k >= 1 && k <= kKdMaxK, RZK_E_INVALID, "k must be in 1 .. " + std::to_string(kKdMaxK));
        if (m == 0) return;
        KdBuf<float> q;
        KdBuf<int> o;
        q.alloc((size_t)kKdDim * m);
        o.alloc((size_t)k * m);
        RZK_HIP(hipMemcpyAsync(q.p, queries, sizeof(float) * kKdDim * (size_t)m, hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(k_kd_nearest, dim3(kd_blocks(m, 64)), dim3(64), 0, ctx->stream, kd_query(kd), q.p,
                           (const float*)nullptr, 0, (const int*)nullptr, (const int*)nullptr, m, k, o.p);
        RZK_HIP(hipGetLastError());
        RZK_HIP(hipMemcpyAsync(out, o.p, sizeof(int) * (size_t)k * m, hipMemcpyDeviceToHost, ctx->stream));
        RZK_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int rzk_map_vec6(rzk_ctx* ctx, const rzk_scene* scene, const float* recv, const int32_t* mat, int64_t n,
                 float diagonal, float* out) {
    if (!ctx) return RZK_E_INVALID;
    return kd_guard(ctx, [&] {
        RZK_CHECK(scene && n >= 0 && (n == 0 || (recv && mat && out)), RZK_E_INVALID, "null argument");
        if (n == 0) return;
        for (int64_t i = 0; i < n; ++i)
            RZK_CHECK(mat[i] < scene->nmat, RZK_E_INVALID, "material id out of range");
        SceneView sv = scene_view(scene);
        KdBuf<float> r, o;
        KdBuf<int> mm;
        r.alloc((size_t)RZK_RECV_FLOATS * n);
        mm.alloc(n);
        o.alloc((size_t)kKdDim * n);
        RZK_HIP(hipMemcpyAsync(r.p, recv, sizeof(float) * RZK_RECV_FLOATS * (size_t)n, hipMemcpyHostToDevice,
                               ctx->stream));
        RZK_HIP(hipMemcpyAsync(mm.p, mat, sizeof(int) * (size_t)n, hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(k_map_vec6, dim3(kd_blocks(n, 256)), dim3(256), 0, ctx->stream, r.p, mm.p, (const int*)nullptr,
                           sv.mat, n, diagonal, o.p, n);
        RZK_HIP(hipGetLastError());
        std::vector<float> planes((size_t)kKdDim * n);
        RZK_HIP(hipMemcpyAsync(planes.data(), o.p, sizeof(float) * kKdDim * (size_t)n, hipMemcpyDeviceToHost,
                               ctx->stream));
        RZK_HIP(hipStreamSynchronize(ctx->stream));
        for (int64_t i = 0; i < n; ++i)
            for (int a = 0; a < kKdDim; ++a) out[kKdDim * i + a] = planes[(size_t)a * n + i];
    });
}

}  // extern "C"
