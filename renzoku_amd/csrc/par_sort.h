// par_sort.h — std::sort of (key, value) pairs by key (the libstdc++ introsort
// restated move for move in std_order.h) computed by a block of lanes, with
// the same final order, equal keys included.
//
// Why the parallel form gives the same permutation:
//  * Every recursion level of introsort partitions disjoint ranges, so all
//    ranges of one level can be partitioned at once (in any order), and every
//    range of level r carries the depth limit 2*floor(log2 n) - r.
//  * The unguarded Hoare partition of [a+1, b) around the pivot P at a swaps
//    the k-th position from the left whose key is not < P (L_k) with the k-th
//    position from the right whose key is not > P (R_k), for k = 1..m where m
//    is the largest k with L_k < R_k, and returns min(L_{m+1}, R_m) (R_0 = b,
//    L_{m+1} absent = +inf).  Ranks of the candidates come from prefix counts;
//    the condition L_k < R_k holds exactly for k <= m, so each L_k decides its
//    own swap and the one at the boundary writes the cut.
//  * After the partitions every key left of a cut is <= every key right of it,
//    so the final insertion sort (16 elements guarded, then unguarded) never
//    moves an element across a cut: it is a stable sort of each final range
//    (at most 16 elements, or already heap-sorted).
//
// The phases are written once for two executors: on the device every lane of
// the block runs a phase for its own lane index and phases are separated by
// barriers; on the host (tests) a phase is a loop over the lane indices in a
// shuffled order.  A value that crosses a barrier lives in a per-lane,
// per-position or per-range array, never in a lane's local variable.
#pragma once

#include "std_order.h"

namespace rzk {
namespace psort {

constexpr unsigned char kFlagL = 1;       // key not < pivot
constexpr unsigned char kFlagR = 2;       // key not > pivot
constexpr unsigned char kFlagCut = 4;     // a final range starts here
constexpr unsigned char kFlagHeap = 8;    // ... and was heap-sorted
constexpr unsigned short kNoRange = 0xffff;

// Buffers of one sort.  Ranges being partitioned have more than 16 elements,
// so at most n / 17 of them exist at once: lanes >= n / 17 + 1 is required.
template <typename K, typename V>
struct Work {
    K* key;
    V* val;
    unsigned short* posL;         // n: positions of the L candidates by global rank
    unsigned short* posR;         // n
    unsigned char* fl;            // n
    unsigned short* rid;          // n: range of each position (kNoRange: finished)
    int* seg;                     // 18 * lanes: per-range data, two generations
    int* lane_a;                  // lanes: per-lane counts / exclusive offsets
    int* lane_b;                  // lanes
    int* wsum;                    // 2 * lanes / 64 (device scan) + 1 counter
};

// a block of NL lanes (blockDim.x == NL, NL a multiple of 64)
template <int NL>
struct DevLanes {
    static constexpr int kLanes = NL;
    __device__ static int count() { return 1; }
    __device__ static int lane(int) { return (int)threadIdx.x; }
    __device__ static void sync() { __syncthreads(); }
    __device__ static int atomic_add(int* p, int v) { return atomicAdd(p, v); }
    // a[t], b[t] -> exclusive prefix sums over t; totals to ta, tb (every lane)
    __device__ static void scan2(int* a, int* b, int* wsum, int& ta, int& tb) {
        const int t = (int)threadIdx.x, ln = t & 63, w = t >> 6;
        const int va = a[t], vb = b[t];
        int ia = va, ib = vb;
        for (int d = 1; d < 64; d <<= 1) {
            int ya = __shfl_up(ia, d, 64), yb = __shfl_up(ib, d, 64);
            if (ln >= d) { ia += ya; ib += yb; }
        }
        if (ln == 63) { wsum[2 * w] = ia; wsum[2 * w + 1] = ib; }
        __syncthreads();
        int oa = 0, ob = 0, sa = 0, sb = 0;
        for (int k = 0; k < NL / 64; ++k) {
            const int xa = wsum[2 * k], xb = wsum[2 * k + 1];
            if (k < w) { oa += xa; ob += xb; }
            sa += xa;
            sb += xb;
        }
        a[t] = oa + ia - va;
        b[t] = ob + ib - vb;
        ta = sa;
        tb = sb;
        __syncthreads();
    }
};

// NL lanes run one after another, in an order shuffled per phase (host tests)
template <int NL>
struct HostLanes {
    static constexpr int kLanes = NL;
    static unsigned& state() { static unsigned s = 12345u; return s; }
    static int* order() { static int o[NL]; return o; }
    static int count() {              // starts a phase: a new lane order
        int* o = order();
        for (int t = 0; t < NL; ++t) o[t] = t;
        for (int t = NL - 1; t > 0; --t) {
            state() = state() * 1664525u + 1013904223u;
            const int u = (int)((state() >> 8) % (unsigned)(t + 1));
            const int x = o[t]; o[t] = o[u]; o[u] = x;
        }
        return NL;
    }
    static int lane(int i) { return order()[i]; }
    static void sync() {}
    static int atomic_add(int* p, int v) { const int old = *p; *p += v; return old; }
    static void scan2(int* a, int* b, int*, int& ta, int& tb) {
        int sa = 0, sb = 0;
        for (int t = 0; t < NL; ++t) {
            const int x = a[t], y = b[t];
            a[t] = sa;
            b[t] = sb;
            sa += x;
            sb += y;
        }
        ta = sa;
        tb = sb;
    }
};

// one phase: every lane index t of the executor, once
#define RZK_PS_LANES(X, t) for (int i_##t = 0, n_##t = X::count(), t = X::lane(0); i_##t < n_##t; \
                                ++i_##t, t = i_##t < n_##t ? X::lane(i_##t) : 0)

// stable sort of key/val[p, p + len), len <= 16, in registers (by rank)
template <typename K, typename V>
RZK_HD void small_stable_sort(K* key, V* val, int p, int len) {
    constexpr int T = stdord::kSortThreshold;
    K k[T];
    V v[T];
#pragma unroll
    for (int q = 0; q < T; ++q) {
        k[q] = q < len ? key[p + q] : key[p];
        v[q] = q < len ? val[p + q] : val[p];
    }
    int r[T];
#pragma unroll
    for (int q = 0; q < T; ++q) {
        int c = 0;
#pragma unroll
        for (int o = 0; o < T; ++o)
            if (o != q && o < len && (k[o] < k[q] || (!(k[q] < k[o]) && o < q))) ++c;
        r[q] = c;
    }
#pragma unroll
    for (int q = 0; q < T; ++q)
        if (q < len) {
            key[p + r[q]] = k[q];
            val[p + r[q]] = v[q];
        }
}

// std::sort(key/val[0, n)) by key; n <= 65535 and n / 17 < X::kLanes.
// Every lane of the executor calls it with the same arguments.  depth0 < 0:
// the introsort depth limit 2*floor(log2 n); tests pass a smaller one to reach
// the heap-sort fallback.
template <typename X, typename K, typename V>
RZK_HD void sort_pairs(Work<K, V> w, int n, int depth0 = -1) {
    constexpr int NL = X::kLanes;
    constexpr int T = stdord::kSortThreshold;
    if (n <= 1) return;
    if (n <= T) {
        RZK_PS_LANES(X, t) if (t == 0) stdord::insertion_sort(w.key, w.val, 0, n);
        X::sync();
        return;
    }
    const int E = (n + NL - 1) / NL;              // positions per lane
    // two generations of per-range data: ranges [A, B), rank bases, ends, cut, children
    int* G0 = w.seg;
    int* G1 = w.seg + 9 * NL;
    int* cnt = w.wsum + 2 * NL / 64;              // next generation's range count
    RZK_PS_LANES(X, t) {
        const int p0 = t * E, p1 = p0 + E < n ? p0 + E : n;
        for (int p = p0; p < p1; ++p) {
            w.fl[p] = p == 0 ? kFlagCut : 0;
            w.rid[p] = 0;
        }
        if (t == 0) {
            G0[0] = 0;                             // A
            G0[NL] = n;                            // B
        }
    }
    X::sync();
    int nr = 1;
    bool relabel = false;                         // rid still holds the previous generation's ranges
    int depth = depth0 >= 0 ? depth0 : 2 * stdord::floor_log2(n);
    for (;;) {
        int *A = G0, *B = G0 + NL, *BL = G0 + 2 * NL, *BR = G0 + 3 * NL, *EL = G0 + 4 * NL, *ER = G0 + 5 * NL;
        int *CUT = G0 + 6 * NL, *LID = G0 + 7 * NL, *RID = G0 + 8 * NL;
        int *PCUT = G1 + 6 * NL, *PLID = G1 + 7 * NL, *PRID = G1 + 8 * NL;   // previous generation
        if (nr == 0) break;
        if (depth == 0) {                         // std::__partial_sort(first, last, last)
            RZK_PS_LANES(X, t)
            for (int j = t; j < nr; j += NL) {
                stdord::heap_sort_range(w.key, w.val, A[j], B[j]);
                w.fl[A[j]] |= kFlagHeap;
            }
            X::sync();
            break;
        }
        --depth;
        // median of (a+1, mid, b-1) to a; the next generation's count
        RZK_PS_LANES(X, t) {
            for (int j = t; j < nr; j += NL) {
                const int a = A[j], b = B[j];
                stdord::median_to_first(w.key, w.val, a, a + 1, a + (b - a) / 2, b - 1);
            }
            if (t == 0) *cnt = 0;
        }
        X::sync();
        // ranges of the positions (from the previous generation's cuts), candidate flags, per-lane counts
        RZK_PS_LANES(X, t) {
            const int p0 = t * E, p1 = p0 + E < n ? p0 + E : n;
            int cl = 0, cr = 0;
            for (int p = p0; p < p1; ++p) {
                int j = w.rid[p];
                if (relabel && j != kNoRange) {
                    j = p < PCUT[j] ? PLID[j] : PRID[j];
                    w.rid[p] = (unsigned short)j;
                }
                unsigned char f = w.fl[p] & (kFlagCut | kFlagHeap);
                if (j != kNoRange && p > A[j]) {
                    const K piv = w.key[A[j]], kp = w.key[p];
                    if (!(kp < piv)) { f |= kFlagL; ++cl; }
                    if (!(piv < kp)) { f |= kFlagR; ++cr; }
                }
                w.fl[p] = f;
            }
            w.lane_a[t] = cl;
            w.lane_b[t] = cr;
        }
        X::sync();
        int totl, totr;
        X::scan2(w.lane_a, w.lane_b, w.wsum, totl, totr);
        // candidate positions by rank; each range's rank bases and R end
        RZK_PS_LANES(X, t) {
            const int p0 = t * E, p1 = p0 + E < n ? p0 + E : n;
            int gl = w.lane_a[t], gr = w.lane_b[t];
            for (int p = p0; p < p1; ++p) {
                const int j = w.rid[p];
                if (j != kNoRange && p == A[j]) { BL[j] = gl; BR[j] = gr; }
                const unsigned char f = w.fl[p];
                if (f & kFlagL) w.posL[gl++] = (unsigned short)p;
                if (f & kFlagR) w.posR[gr++] = (unsigned short)p;
                if (j != kNoRange && p == B[j] - 1) { EL[j] = gl; ER[j] = gr; }
            }
        }
        X::sync();
        // L_k decides its swap (L_k < R_k holds exactly for k <= m); the cut
        // min(L_{m+1}, R_m) is written by L_{m+1} (the first that does not
        // swap) or, when every L candidate swaps, by the last one (cut = R_m)
        RZK_PS_LANES(X, t) {
            const int p0 = t * E, p1 = p0 + E < n ? p0 + E : n;
            int gl = w.lane_a[t];
            for (int p = p0; p < p1; ++p) {
                if (!(w.fl[p] & kFlagL)) continue;
                const int j = w.rid[p];
                const int k = gl - BL[j] + 1, nrr = ER[j] - BR[j], nll = EL[j] - BL[j];
                const int rk = k <= nrr ? (int)w.posR[ER[j] - k] : -1;
                if (p < rk) {
                    stdord::swap_at(w.key, w.val, p, rk);
                    if (k == nll) CUT[j] = rk;
                } else {
                    bool first = k == 1;
                    if (!first) {
                        const int rp = k - 1 <= nrr ? (int)w.posR[ER[j] - (k - 1)] : -1;
                        first = (int)w.posL[gl - 1] < rp;
                    }
                    if (first) {
                        const int rm = k >= 2 ? (int)w.posR[ER[j] - (k - 1)] : B[j];
                        CUT[j] = p < rm ? p : rm;
                    }
                }
                ++gl;
            }
        }
        X::sync();
        // children: [a, cut) and [cut, b); those over the threshold go on
        RZK_PS_LANES(X, t)
        for (int j = t; j < nr; j += NL) {
            const int a = A[j], b = B[j], c = CUT[j];
            w.fl[c] |= kFlagCut;
            int lid = kNoRange, rid = kNoRange;
            if (c - a > T) { lid = X::atomic_add(cnt, 1); G1[lid] = a; G1[NL + lid] = c; }
            if (b - c > T) { rid = X::atomic_add(cnt, 1); G1[rid] = c; G1[NL + rid] = b; }
            LID[j] = lid;
            RID[j] = rid;
        }
        X::sync();
        nr = *cnt;
        X::sync();
        int* tg = G0; G0 = G1; G1 = tg;
        relabel = true;
    }
    // the final insertion sort, range by range (a stable sort of each)
    RZK_PS_LANES(X, t) {
        const int p0 = t * E, p1 = p0 + E < n ? p0 + E : n;
        for (int p = p0; p < p1; ++p) {
            const unsigned char f = w.fl[p];
            if (!(f & kFlagCut) || (f & kFlagHeap)) continue;
            int e = p + 1;
            while (e < n && !(w.fl[e] & kFlagCut)) ++e;
            small_stable_sort(w.key, w.val, p, e - p);
        }
    }
    X::sync();
}

}  // namespace psort
}  // namespace rzk
