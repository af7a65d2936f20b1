// refine.hip — LightSlice's per-slice cluster refinement (lightslice.cc:837-902):
// for every slice, the global clusters are split, most expensive first, until
// the slice holds max_slice_clusters clusters; each split projects the
// cluster's columns of the slice matrix (the reduced matrix restricted to the
// slice's neighbour rows) on the line through two sampled columns and cuts at
// the largest gap (split_cluster, :405-488); the two halves get new
// representatives sampled by column norm (choose_representative, :385-400).
//
// One 256-lane workgroup refines one slice.  The numeric steps run across the
// lanes with every floating-point sum kept in the reference's order: columns
// are gathered into LDS up to 256 at a time, one lane per column; per-column
// sums (norms, projections, products) run one lane per column, per-row sums
// one lane per row, and the sums over the cluster (the cost, the norm total,
// the pdf total and cdf) run as one identical chain in every lane over values
// held in LDS.  The sort of the projections follows std::sort's moves (the
// block-parallel form in par_sort.h, up to 4,096 members; the one-lane form in
// std_order.h beyond) and the cost heap std::priority_queue's (one lane, on a
// table in LDS when it fits), because the reference's result depends on their
// tie orders.
//
// Draws: the reference takes them from one MT stream, slice after slice.  A
// slice takes 4 per split as a rule (2 to pick the line, 1 per new
// representative) but fewer when a split cluster has two members or the
// refinement stops early, and how many depends on the numbers it reads, so a
// slice's offset in the pre-generated stream is known only once every earlier
// slice has run at its own exact offset.  Passes run (slice, offset) jobs, one
// workgroup each: the first pass every slice at the rule's offset; each later
// pass the first unresolved slice at its exact offset and the slices after it
// at every offset near their estimate (the draws each slice took on its runs
// so far, averaged), the window widening with the distance.  After a pass the
// host follows the chain of exact offsets through the jobs that guessed right
// and keeps their results.  When the slice matrices do not all fit in memory,
// slices run in batches at offsets that assume the rule after the first wrong
// one, re-running from there.
#pragma clang fp contract(off)

#include <algorithm>
#include <climits>
#include <cstring>
#include <memory>

#include "host_sampling.h"
#include "rzk_device.h"
#include "par_sort.h"
#include "std_order.h"

namespace rzk {

constexpr int kRefBlock = 256;
constexpr int kRefSortLds = 4096;      // clusters of up to this many members are worked on in LDS
constexpr int kRefMaxNb = 32;          // neighbour rows per slice (3 * 32 flattened rows)
constexpr unsigned kRefLdsMax = 160u * 1024u;   // LDS of one gfx950 compute unit

struct RefineArgs {
    int S;
    const int* nb;           // S x nnb neighbour rows
    int nnb;
    int ncl;                 // kept global clusters
    const int* off;          // ncl + 1 member offsets
    const int* mem;          // members, cluster by cluster, increasing VPL order
    int nmem;
    const int* grep;
    const float* gw;
    const float* mags;       // per member position: flattened column magnitude (pdfs of the global clusters)
    const float* norms_R;    // per member position: column norm (pdfs of split clusters)
    const float* SM;         // per scratch row: the slice matrix, nmem x 3*nnb, member by member
    int maxC;
    int lds_clusters;        // cluster table and cost heap in LDS (else in the scratch rows)
    const float* draws;      // pre-generated MT numbers
    const int64_t* draw_off; // per slice
    int s0;                  // first slice of the launch
    int batch0;              // slice index of slice-data row 0 (SM, the initial heap)
    // speculative jobs: block j refines slice job_slice[j] from draw offset
    // job_off[j] in scratch row j and writes result row j; null: block j is
    // slice s0 + j at draw_off[s], in scratch row s - batch0, result row s
    const int* job_slice;
    const int64_t* job_off;
    // results (rows x maxC, rows)
    int* rep;
    float* w;
    int* count;
    int* used;               // draws taken
    // per-slice state (scratch rows of this batch, or of the jobs)
    int* perm;               // nmem member positions (index into mem), cluster by cluster
    int* perm_tmp;           // nmem (clusters over kRefSortLds members)
    float* nrm;              // nmem (idem)
    float* line_t;           // nmem (idem)
    int* line_k;             // nmem
    int* cst;                // maxC
    int* csz;
    int* cflag;              // 0: pdf over mags (global cluster), 1: over norms_R
    float* hkey;             // maxC
    int* hval;
    // computed once per slice row (they do not depend on the draws)
    const int* iota;         // 0 .. nmem - 1: the member order before any split
    float* cost0;            // ncl costs of the kept clusters
    const float* hkey0;      // the cost heap after pushing them in order (ncl entries)
    const int* hval0;
};

// LDS layout of one refinement block, shared by the kernel and its launcher
struct RefLayout {
    int ldx, kc;
    unsigned oX, oU, oA, oT, oPm, oMap, oNb, oInt, oKey, oVal, oSeg, oLa, oLb, oWs, oCl, oPosL, oPosR, oRid, oFl, bytes;
};

RZK_HD unsigned ref_take(unsigned& o, unsigned bytes) {
    const unsigned r = o;
    o += (bytes + 15u) & ~15u;
    return r;
}

RZK_HD RefLayout ref_layout(int nnb, int maxC, bool lds_clusters) {
    RefLayout L;
    L.ldx = 3 * nnb + 1;
    L.kc = 4 * 256 * L.ldx <= 64 * 1024 ? 256 : 128;     // columns staged per pass
    unsigned o = 0;
    L.oX = ref_take(o, 4u * L.kc * L.ldx);
    L.oU = ref_take(o, 4u * 3 * kRefMaxNb);
    L.oA = ref_take(o, 4u * 3 * kRefMaxNb);
    L.oT = ref_take(o, 4u * kRefBlock);
    L.oPm = ref_take(o, 4u * kRefBlock);
    L.oMap = ref_take(o, 4u * kRefBlock);
    L.oNb = ref_take(o, 4u * kRefMaxNb);
    L.oInt = ref_take(o, 4u * 16);
    L.oKey = ref_take(o, 4u * kRefSortLds);
    L.oVal = ref_take(o, 4u * kRefSortLds);
    L.oSeg = ref_take(o, 4u * 18 * kRefBlock);
    L.oLa = ref_take(o, 4u * kRefBlock);
    L.oLb = ref_take(o, 4u * kRefBlock);
    L.oWs = ref_take(o, 4u * (2 * kRefBlock / 64 + 1));
    L.oCl = lds_clusters ? ref_take(o, 20u * (unsigned)maxC) : 0u;
    L.oPosL = ref_take(o, 2u * kRefSortLds);
    L.oPosR = ref_take(o, 2u * kRefSortLds);
    L.oRid = ref_take(o, 2u * kRefSortLds);
    L.oFl = ref_take(o, kRefSortLds);
    L.bytes = o;
    return L;
}

struct RefineShared {
    float* sX;      // kc x ldx gathered columns
    int ldx, kc;
    float* sU;      // flattened rows: row sums / the line direction
    float* sA;      // column a
    float* sT;      // kRefBlock per-column values of a pass
    int* sPm;       // kRefBlock member positions of the staged columns
    int* sMap;      // kRefBlock: after a split staged whole, the staged column of each new position
    const float* sm;   // this slice's matrix copy (nmem x 3*nnb)
    int* sNb;       // the slice's neighbour rows
    int* sInt;      // small broadcast area
    float* sKey;    // kRefSortLds: projections / pdf values / column norms / new member order
    int* sVal;
    psort::Work<float, int> pw;   // the sort's buffers (sKey, sVal and the rest)
};

// Ordered fp32 chains over LDS values.  The loads of 16 elements are issued
// before their adds, so a chain waits for LDS once per 16 elements instead of
// once per element; the adds keep their order.
__device__ __forceinline__ float chain_sum(const float* v, int n, float acc) {
    int k = 0;
    for (; k + 16 <= n; k += 16) {
        float x[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) x[q] = v[k + q];
#pragma unroll
        for (int q = 0; q < 16; ++q) acc += x[q];
    }
    for (; k < n; ++k) acc += v[k];
    return acc;
}

// sum_k v[k * stride] (a column of a row-major LDS tile)
__device__ __forceinline__ float chain_sum_strided(const float* v, int n, int stride, float acc) {
    int k = 0;
    for (; k + 16 <= n; k += 16) {
        float x[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) x[q] = v[(k + q) * stride];
#pragma unroll
        for (int q = 0; q < 16; ++q) acc += x[q];
    }
    for (; k < n; ++k) acc += v[k * stride];
    return acc;
}

// sum_k v[map[k] * stride] (a row of the staged columns, in a new column order)
__device__ __forceinline__ float chain_sum_mapped(const float* v, const int* map, int n, int stride, float acc) {
    int k = 0;
    for (; k + 16 <= n; k += 16) {
        int m[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) m[q] = map[k + q];
        float x[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) x[q] = v[m[q] * stride];
#pragma unroll
        for (int q = 0; q < 16; ++q) acc += x[q];
    }
    for (; k < n; ++k) acc += v[map[k] * stride];
    return acc;
}

// sum_r x[r] * x[r]
__device__ __forceinline__ float chain_sumsq(const float* x, int n) {
    float acc = 0.f;
    int k = 0;
    for (; k + 16 <= n; k += 16) {
        float y[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) y[q] = x[k + q];
#pragma unroll
        for (int q = 0; q < 16; ++q) acc += y[q] * y[q];
    }
    for (; k < n; ++k) acc += x[k] * x[k];
    return acc;
}

// sum_r x[r] * y[r]
__device__ __forceinline__ float chain_dot(const float* x, const float* y, int n) {
    float acc = 0.f;
    int k = 0;
    for (; k + 16 <= n; k += 16) {
        float p[16], q2[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) { p[q] = x[k + q]; q2[q] = y[k + q]; }
#pragma unroll
        for (int q = 0; q < 16; ++q) acc += p[q] * q2[q];
    }
    for (; k < n; ++k) acc += x[k] * y[k];
    return acc;
}

// sum_r (a[r] - x[r]) * d[r]
__device__ __forceinline__ float chain_proj(const float* a, const float* x, const float* d, int n) {
    float acc = 0.f;
    int k = 0;
    for (; k + 16 <= n; k += 16) {
        float p[16], q2[16], r2[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) { p[q] = a[k + q]; q2[q] = x[k + q]; r2[q] = d[k + q]; }
#pragma unroll
        for (int q = 0; q < 16; ++q) acc += (p[q] - q2[q]) * r2[q];
    }
    for (; k < n; ++k) acc += (a[k] - x[k]) * d[k];
    return acc;
}

// columns perm[st + c0 .. st + c0 + cnt) of the slice matrix into sX (cnt <= kc),
// copied from the slice's contiguous copy: consecutive lanes read consecutive
// values of a column, eight loads in flight per lane
__device__ __forceinline__ void ref_stage(const RefineArgs& a, const RefineShared& sh, const int* perm, int st, int c0,
                                          int cnt) {
    const int M = 3 * a.nnb, t = threadIdx.x;
    if (t < cnt) sh.sPm[t] = perm[st + c0 + t];
    __syncthreads();
    const int total = cnt * M, dcol = kRefBlock / M, dr = kRefBlock - dcol * M;
    int col = t / M, r = t - col * M;
    for (int e0 = 0; e0 < total; e0 += 8 * kRefBlock) {
        float v[8];
        int dst[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const bool ok = e0 + q * kRefBlock + t < total;
            dst[q] = ok ? col * sh.ldx + r : -1;
            v[q] = ok ? sh.sm[(int64_t)sh.sPm[col] * M + r] : 0.f;
            col += dcol;
            r += dr;
            if (r >= M) { r -= M; ++col; }
        }
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if (dst[q] >= 0) sh.sX[dst[q]] = v[q];
    }
}

// compute_distance (lightslice.cc:336-357) of columns perm[st .. st + n)
__device__ float ref_cost(const RefineArgs& a, const RefineShared& sh, const int* perm, float* nrm_g, int st, int n) {
    const int M = 3 * a.nnb, t = threadIdx.x, kc = sh.kc;
    float* nrm = n <= kRefSortLds ? sh.sKey : nrm_g + st;     // column norms between the passes
    float u = 0.f, c = 0.f;
    // pass 1: column norms (get_column_norm), row sums (row_sum), their total (sum)
    for (int c0 = 0; c0 < n; c0 += kc) {
        const int cnt = min(kc, n - c0);
        ref_stage(a, sh, perm, st, c0, cnt);
        __syncthreads();
        if (t < cnt) {
            const float nv = sqrtf(chain_sumsq(sh.sX + t * sh.ldx, M));
            nrm[c0 + t] = nv;
            sh.sT[t] = nv;
        }
        if (t < M) u = chain_sum_strided(sh.sX + t, cnt, sh.ldx, u);
        __syncthreads();
        c = chain_sum(sh.sT, cnt, c);
        __syncthreads();
    }
    if (t < M) sh.sU[t] = u;
    __syncthreads();
    // pass 2: v_j = sum_r x_rj u_r (Rt.mul), cost = sum_j (norm_j * c - v_j)
    float cost = 0.f;
    for (int c0 = 0; c0 < n; c0 += kc) {
        const int cnt = min(kc, n - c0);
        ref_stage(a, sh, perm, st, c0, cnt);
        __syncthreads();
        if (t < cnt) sh.sT[t] = nrm[c0 + t] * c - chain_dot(sh.sX + t * sh.ldx, sh.sU, M);
        __syncthreads();
        cost = chain_sum(sh.sT, cnt, cost);
        __syncthreads();
    }
    return cost;
}

// compute_distance of the staged columns sX[sMap[off + j]], j < n <= kc: the
// same sums in the same order as ref_cost (one chunk), without staging again
__device__ float ref_cost_local(const RefineArgs& a, const RefineShared& sh, int off, int n) {
    const int M = 3 * a.nnb, t = threadIdx.x;
    if (t < n) sh.sT[t] = sqrtf(chain_sumsq(sh.sX + sh.sMap[off + t] * sh.ldx, M));
    float u = 0.f;
    if (t < M) u = chain_sum_mapped(sh.sX + t, sh.sMap + off, n, sh.ldx, 0.f);
    __syncthreads();
    const float c = chain_sum(sh.sT, n, 0.f);
    if (t < M) sh.sU[t] = u;
    __syncthreads();
    if (t < n) sh.sT[t] = sh.sT[t] * c - chain_dot(sh.sX + sh.sMap[off + t] * sh.ldx, sh.sU, M);
    __syncthreads();
    const float cost = chain_sum(sh.sT, n, 0.f);
    __syncthreads();
    return cost;
}

// DiscretePdf over src[perm[st + k]] (k < n) sampled at nu (1 or 2) numbers:
// out_k[i] = first k with u[i] < cdf[k] (-1 if none), out_p[i] = pdf[k]
// (pdf.cc:18-77: ordered total, pdf = a / total, cdf running sum, last = 1).
// Every lane runs the same ordered chains over the values staged in LDS.
__device__ void ref_pdf_sample(const RefineShared& sh, const float* src, const int* perm, int st, int n, int nu,
                               const float* u, int* out_k, float* out_p) {
    const int t = threadIdx.x;
    if (n <= kRefSortLds) {
        float* v = sh.sKey;
        for (int k0 = t; k0 < n; k0 += 4 * kRefBlock) {
            int pk[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) pk[q] = k0 + q * kRefBlock < n ? perm[st + k0 + q * kRefBlock] : 0;
            float x[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) x[q] = k0 + q * kRefBlock < n ? src[pk[q]] : 0.f;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (k0 + q * kRefBlock < n) v[k0 + q * kRefBlock] = x[q];
        }
        __syncthreads();
        const float sum = chain_sum(v, n, 0.f);
        __syncthreads();
        if (sum == 0.0f) {
            for (int i = 0; i < nu; ++i) { out_k[i] = -1; out_p[i] = 0.f; }
            return;
        }
        for (int k = t; k < n; k += kRefBlock) v[k] = v[k] / sum;
        __syncthreads();
        // the cdf, as one ordered chain (lane 0), then every lane looks for the first k with u < cdf[k]
        float* cdf = reinterpret_cast<float*>(sh.sVal);
        if (t == 0) {
            float run = 0.f;
            int k = 0;
            for (; k + 16 <= n; k += 16) {
                float x[16];
#pragma unroll
                for (int q = 0; q < 16; ++q) x[q] = v[k + q];
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    run = k + q == 0 ? x[q] : run + x[q];
                    cdf[k + q] = run;
                }
            }
            for (; k < n; ++k) {
                run = k == 0 ? v[0] : run + v[k];
                cdf[k] = run;
            }
            cdf[n - 1] = 1.0f;
        }
        __syncthreads();
        int k0 = INT_MAX, k1 = INT_MAX;
        for (int k = t; k < n; k += kRefBlock) {
            const float ck = cdf[k];
            if (k0 == INT_MAX && u[0] < ck) k0 = k;
            if (nu > 1 && k1 == INT_MAX && u[1] < ck) k1 = k;
        }
        for (int o = 32; o >= 1; o >>= 1) {
            k0 = min(k0, __shfl_xor(k0, o, 64));
            k1 = min(k1, __shfl_xor(k1, o, 64));
        }
        if ((t & 63) == 0) { sh.sInt[8 + 2 * (t >> 6)] = k0; sh.sInt[9 + 2 * (t >> 6)] = k1; }
        __syncthreads();
        k0 = INT_MAX;
        k1 = INT_MAX;
        for (int w2 = 0; w2 < kRefBlock / 64; ++w2) {
            k0 = min(k0, sh.sInt[8 + 2 * w2]);
            k1 = min(k1, sh.sInt[9 + 2 * w2]);
        }
        if (k0 == INT_MAX) k0 = -1;
        if (k1 == INT_MAX) k1 = -1;
        out_k[0] = k0;
        out_p[0] = k0 >= 0 ? v[k0] : 0.f;
        if (nu > 1) {
            out_k[1] = k1;
            out_p[1] = k1 >= 0 ? v[k1] : 0.f;
        }
        __syncthreads();
        return;
    }
    // long clusters: 256 values at a time through sT, first hits by LDS atomicMin
    float sum = 0.f;
    for (int c0 = 0; c0 < n; c0 += kRefBlock) {
        const int cnt = min(kRefBlock, n - c0);
        sh.sT[t] = t < cnt ? src[perm[st + c0 + t]] : 0.f;
        __syncthreads();
        for (int k = 0; k < cnt; ++k) sum += sh.sT[k];
        __syncthreads();
    }
    if (t == 0) { sh.sInt[0] = INT_MAX; sh.sInt[1] = INT_MAX; }
    __syncthreads();
    if (sum == 0.0f) {
        for (int i = 0; i < nu; ++i) { out_k[i] = -1; out_p[i] = 0.f; }
        return;
    }
    float run = 0.f;
    for (int c0 = 0; c0 < n; c0 += kRefBlock) {
        const int cnt = min(kRefBlock, n - c0);
        const float p = t < cnt ? src[perm[st + c0 + t]] / sum : 0.f;
        sh.sT[t] = p;
        __syncthreads();
        float mine = 0.f;
        for (int k = 0; k < cnt; ++k) {
            run = (c0 + k == 0) ? sh.sT[0] : run + sh.sT[k];
            if (k == t) mine = run;
        }
        if (c0 + t == n - 1) mine = 1.0f;
        if (t < cnt)
            for (int i = 0; i < nu; ++i)
                if (u[i] < mine) atomicMin(&sh.sInt[i], c0 + t);
        __syncthreads();
        if (sh.sInt[0] != INT_MAX && (nu < 2 || sh.sInt[1] != INT_MAX)) break;    // uniform
        __syncthreads();
    }
    __syncthreads();
    for (int i = 0; i < nu; ++i) {
        const int k = sh.sInt[i];
        if (k == INT_MAX) { out_k[i] = -1; out_p[i] = 0.f; }
        else { out_k[i] = k; out_p[i] = src[perm[st + k]] / sum; }
    }
    __syncthreads();
}

// the first k in [1, nl) with the largest gap key[k] - key[k-1] > 0, or -1
// (the sequential scan at lightslice.cc:466-473: strict >, starting from 0)
__device__ int ref_max_gap(const RefineShared& sh, const float* key, int nl) {
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    float best = 0.f;
    int bk = -1;
    for (int k = 1 + t; k < nl; k += kRefBlock) {
        const float d = key[k] - key[k - 1];
        if (d > best) { best = d; bk = k; }
    }
    for (int o = 32; o >= 1; o >>= 1) {
        const float ob = __shfl_xor(best, o, 64);
        const int ok = __shfl_xor(bk, o, 64);
        if (ok >= 0 && (bk < 0 || ob > best || (ob == best && ok < bk))) { best = ob; bk = ok; }
    }
    if (lane == 0) {
        sh.sInt[8 + 2 * wv] = __float_as_int(best);
        sh.sInt[9 + 2 * wv] = bk;
    }
    __syncthreads();
    best = 0.f;
    bk = -1;
    for (int w2 = 0; w2 < kRefBlock / 64; ++w2) {
        const float ob = __int_as_float(sh.sInt[8 + 2 * w2]);
        const int ok = sh.sInt[9 + 2 * w2];
        if (ok >= 0 && (bk < 0 || ob > best || (ob == best && ok < bk))) { best = ob; bk = ok; }
    }
    __syncthreads();
    return bk;
}

// split_cluster (lightslice.cc:405-488) of perm[st .. st + n), n >= 2: the
// range is rewritten as indices1 then indices2; returns |indices1|.
// local: on return, sX holds the cluster's columns in their old order and
// sMap[p] gives the staged column of every new position p (clusters of at
// most kc members that reached the projection)
__device__ int ref_split(const RefineArgs& a, const RefineShared& sh, int* perm, int* perm_tmp, float* line_t,
                         int* line_k, int st, int n, int flag, int64_t draw0, int& drawn, bool& local) {
    const int M = 3 * a.nnb, t = threadIdx.x, kc = sh.kc;
    local = false;
    if (n == 2) return 1;
    const float u[2] = {a.draws[draw0 + drawn], a.draws[draw0 + drawn + 1]};
    drawn += 2;
    int ij[2];
    float pij[2];
    ref_pdf_sample(sh, flag ? a.norms_R : a.mags, perm, st, n, 2, u, ij, pij);
    const int i = ij[0], j = ij[1];
    if (i < 0 || j < 0) return n / 2;                                    // split by half, order kept
    // a, a - b
    if (t < M) {
        const float av = sh.sm[(int64_t)perm[st + i] * M + t];
        const float bv = sh.sm[(int64_t)perm[st + j] * M + t];
        sh.sA[t] = av;
        sh.sU[t] = av - bv;
    }
    __syncthreads();
    // t_k = (a - c_k) . (a - b) for k != i, j, at line position k - [k > i] - [k > j]
    const int nl = n - 2;
    const bool in_lds = n <= kRefSortLds;
    float* key = in_lds ? sh.sKey : line_t;
    int* val = in_lds ? sh.sVal : line_k;
    for (int c0 = 0; c0 < n; c0 += kc) {
        const int cnt = min(kc, n - c0);
        ref_stage(a, sh, perm, st, c0, cnt);
        __syncthreads();
        if (t < cnt) {
            const int k = c0 + t;
            if (k != i && k != j) {
                const float tk = chain_proj(sh.sA, sh.sX + t * sh.ldx, sh.sU, M);
                const int p = k - (k > i ? 1 : 0) - (k > j ? 1 : 0);
                key[p] = tk;
                val[p] = k;
            }
        }
        __syncthreads();
    }
    if (in_lds) {
        psort::sort_pairs<psort::DevLanes<kRefBlock>>(sh.pw, nl);
    } else {
        if (t == 0) stdord::sort_pairs(line_t, line_k, nl);
        __syncthreads();
    }
    const int cut = ref_max_gap(sh, key, nl);
    local = n <= kc;
    if (cut < 0) {
        if (local) {
            for (int p = t; p < n; p += kRefBlock) sh.sMap[p] = p;
            __syncthreads();
        }
        return n / 2;
    }
    // indices1 = [i, line[0 .. cut)], indices2 = [j, line[cut ..)]
    int* tmp = in_lds ? reinterpret_cast<int*>(sh.sKey) : perm_tmp;
    for (int p0 = t; p0 < n; p0 += 4 * kRefBlock) {
        int x[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int p = p0 + q * kRefBlock;
            int src = 0;
            if (p < n) {
                if (p == 0) src = i;
                else if (p <= cut) src = val[p - 1];
                else if (p == cut + 1) src = j;
                else src = val[p - 2];
            }
            x[q] = p < n ? perm[st + src] : 0;
            if (local && p < n) sh.sMap[p] = src;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (p0 + q * kRefBlock < n) tmp[p0 + q * kRefBlock] = x[q];
    }
    __syncthreads();
    for (int p = t; p < n; p += kRefBlock) perm[st + p] = tmp[p];
    __syncthreads();
    return cut + 1;
}

// the slice matrix of slices s0 .. s0 + gridDim.y - 1, member by member:
// SM[row][p][3 q + c] = R[(mem[p] * S + nb[s][q]) * 3 + c], row = s - s0
__global__ __launch_bounds__(256) void k_slice_columns(const float* __restrict__ R, int S,
                                                      const int* __restrict__ nb, int nnb,
                                                      const int* __restrict__ mem, int nmem, int s0,
                                                      float* __restrict__ SM) {
    const int M = 3 * nnb, s = s0 + (int)blockIdx.y;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)nmem * M) return;
    const int p = (int)(e / M), r = (int)(e - (int64_t)p * M), q = r / 3, ch = r - 3 * q;
    SM[(int64_t)blockIdx.y * nmem * M + e] = R[((int64_t)mem[p] * S + nb[(int64_t)s * nnb + q]) * 3 + ch];
}

// src[mem[p]] for every member position p
__global__ void k_gather_members(const float* __restrict__ src, const int* __restrict__ mem, int nmem,
                                 float* __restrict__ out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < nmem) out[p] = src[mem[p]];
}

// the LDS views of one block (the layout is the same for every kernel here;
// k_initial_costs allocates only up to oVal)
__device__ __forceinline__ RefineShared ref_shared(unsigned char* s_ref, const RefLayout& L, const RefineArgs& a,
                                                   int row) {
    RefineShared sh;
    sh.ldx = L.ldx;
    sh.kc = L.kc;
    sh.sX = reinterpret_cast<float*>(s_ref + L.oX);
    sh.sU = reinterpret_cast<float*>(s_ref + L.oU);
    sh.sA = reinterpret_cast<float*>(s_ref + L.oA);
    sh.sT = reinterpret_cast<float*>(s_ref + L.oT);
    sh.sPm = reinterpret_cast<int*>(s_ref + L.oPm);
    sh.sMap = reinterpret_cast<int*>(s_ref + L.oMap);
    sh.sm = a.SM + (int64_t)row * a.nmem * (3 * a.nnb);
    sh.sNb = reinterpret_cast<int*>(s_ref + L.oNb);
    sh.sInt = reinterpret_cast<int*>(s_ref + L.oInt);
    sh.sKey = reinterpret_cast<float*>(s_ref + L.oKey);
    sh.sVal = reinterpret_cast<int*>(s_ref + L.oVal);
    return sh;
}

// compute_distance of every kept cluster in every slice of the launch: block
// (c, y) costs cluster c of slice s0 + y (scratch row s0 + y - batch0)
__global__ __launch_bounds__(kRefBlock) void k_initial_costs(RefineArgs a) {
    extern __shared__ __align__(16) unsigned char s_ref[];
    const RefLayout L = ref_layout(a.nnb, a.maxC, false);
    const int c = (int)blockIdx.x, row = a.s0 + (int)blockIdx.y - a.batch0;
    const RefineShared sh = ref_shared(s_ref, L, a, row);
    const int st = a.off[c], n = a.off[c + 1] - a.off[c];
    const float cost = ref_cost(a, sh, a.iota, a.nrm + (int64_t)row * a.nmem, st, n);
    if (threadIdx.x == 0) a.cost0[(int64_t)row * a.ncl + c] = cost;
}

// the cost heap of each slice after pushing its kept clusters' costs in
// order (lightslice.cc:862-866), built by one lane in LDS
__global__ __launch_bounds__(64) void k_initial_heap(const float* __restrict__ cost0, int ncl, float* __restrict__ hkey0,
                                                     int* __restrict__ hval0) {
    extern __shared__ __align__(16) unsigned char s_heap[];
    float* hk = reinterpret_cast<float*>(s_heap);
    int* hv = reinterpret_cast<int*>(hk + ncl);
    const int64_t row = blockIdx.x;
    if (threadIdx.x == 0) {
        int hn = 0;
        for (int c = 0; c < ncl; ++c) stdord::heap_push(hk, hv, hn, cost0[row * ncl + c], c);
    }
    __syncthreads();
    for (int c = threadIdx.x; c < ncl; c += 64) {
        hkey0[row * ncl + c] = hk[c];
        hval0[row * ncl + c] = hv[c];
    }
}

__global__ __launch_bounds__(kRefBlock) void k_refine_slices(RefineArgs a) {
    extern __shared__ __align__(16) unsigned char s_ref[];
    const RefLayout L = ref_layout(a.nnb, a.maxC, a.lds_clusters != 0);
    const int t = threadIdx.x;
    const int j = (int)blockIdx.x;
    const int s = a.job_slice ? a.job_slice[j] : a.s0 + j;
    const int row = s - a.batch0;                  // slice data
    const int jrow = a.job_slice ? j : row;        // scratch
    const int orow = a.job_slice ? j : s;          // results
    RefineShared sh = ref_shared(s_ref, L, a, row);
    sh.pw.key = sh.sKey;
    sh.pw.val = sh.sVal;
    sh.pw.posL = reinterpret_cast<unsigned short*>(s_ref + L.oPosL);
    sh.pw.posR = reinterpret_cast<unsigned short*>(s_ref + L.oPosR);
    sh.pw.rid = reinterpret_cast<unsigned short*>(s_ref + L.oRid);
    sh.pw.fl = s_ref + L.oFl;
    sh.pw.seg = reinterpret_cast<int*>(s_ref + L.oSeg);
    sh.pw.lane_a = reinterpret_cast<int*>(s_ref + L.oLa);
    sh.pw.lane_b = reinterpret_cast<int*>(s_ref + L.oLb);
    sh.pw.wsum = reinterpret_cast<int*>(s_ref + L.oWs);

    if (s >= a.S) return;
    int* perm = a.perm + (int64_t)jrow * a.nmem;
    int* perm_tmp = a.perm_tmp + (int64_t)jrow * a.nmem;
    float* nrm = a.nrm + (int64_t)jrow * a.nmem;
    float* line_t = a.line_t + (int64_t)jrow * a.nmem;
    int* line_k = a.line_k + (int64_t)jrow * a.nmem;
    int *cst, *csz, *cflag, *hval;
    float* hkey;
    if (a.lds_clusters) {
        cst = reinterpret_cast<int*>(s_ref + L.oCl);
        csz = cst + a.maxC;
        cflag = csz + a.maxC;
        hkey = reinterpret_cast<float*>(cflag + a.maxC);
        hval = reinterpret_cast<int*>(hkey + a.maxC);
    } else {
        cst = a.cst + (int64_t)jrow * a.maxC;
        csz = a.csz + (int64_t)jrow * a.maxC;
        cflag = a.cflag + (int64_t)jrow * a.maxC;
        hkey = a.hkey + (int64_t)jrow * a.maxC;
        hval = a.hval + (int64_t)jrow * a.maxC;
    }
    int* rep = a.rep + (int64_t)orow * a.maxC;
    float* w = a.w + (int64_t)orow * a.maxC;
    const int64_t draw0 = a.job_slice ? a.job_off[j] : a.draw_off[s];

    if (t < a.nnb) sh.sNb[t] = a.nb[(int64_t)s * a.nnb + t];
    for (int e = t; e < a.nmem; e += kRefBlock) perm[e] = e;
    for (int c = t; c < a.ncl; c += kRefBlock) {
        cst[c] = a.off[c];
        csz[c] = a.off[c + 1] - a.off[c];
        cflag[c] = 0;
        rep[c] = a.grep[c];
        w[c] = a.gw[c];
        hkey[c] = a.hkey0[(int64_t)row * a.ncl + c];      // the kept clusters' costs, pushed in order
        hval[c] = a.hval0[(int64_t)row * a.ncl + c];
    }
    __syncthreads();
    int hn = a.ncl;                // heap size (lane 0's copy is the one used)
    int cnt = a.ncl, drawn = 0;
    for (;;) {
        __syncthreads();
        if (t == 0) {
            int stop = (hn == 0 || cnt >= a.maxC) ? 1 : 0;
            if (!stop && csz[hval[0]] <= 1) stop = 1;
            sh.sInt[3] = stop;
            sh.sInt[4] = stop ? -1 : hval[0];
        }
        __syncthreads();
        if (sh.sInt[3]) break;
        const int top = sh.sInt[4];
        const int st = cst[top], n = csz[top], flag = cflag[top];
        __syncthreads();
        bool local;
        const int n1 = ref_split(a, sh, perm, perm_tmp, line_t, line_k, st, n, flag, draw0, drawn, local);
        const float cost1 = local ? ref_cost_local(a, sh, 0, n1) : ref_cost(a, sh, perm, nrm, st, n1);
        const float cost2 = local ? ref_cost_local(a, sh, n1, n - n1) : ref_cost(a, sh, perm, nrm, st + n1, n - n1);
        int k1, k2;
        float p1, p2;
        const float u1 = a.draws[draw0 + drawn], u2 = a.draws[draw0 + drawn + 1];
        drawn += 2;
        ref_pdf_sample(sh, a.norms_R, perm, st, n1, 1, &u1, &k1, &p1);
        ref_pdf_sample(sh, a.norms_R, perm, st + n1, n - n1, 1, &u2, &k2, &p2);
        if (t == 0) {
            const int n2 = n - n1;
            csz[top] = n1;
            cflag[top] = 1;
            rep[top] = a.mem[k1 < 0 ? perm[st + n1 / 2] : perm[st + k1]];
            w[top] = k1 < 0 ? 0.0f : 1.0f / p1;
            cst[cnt] = st + n1;
            csz[cnt] = n2;
            cflag[cnt] = 1;
            rep[cnt] = a.mem[k2 < 0 ? perm[st + n1 + n2 / 2] : perm[st + n1 + k2]];
            w[cnt] = k2 < 0 ? 0.0f : 1.0f / p2;
            stdord::heap_pop(hkey, hval, hn);
            stdord::heap_push(hkey, hval, hn, cost1, top);
            stdord::heap_push(hkey, hval, hn, cost2, cnt);
        }
        ++cnt;
    }
    if (t == 0) {
        a.count[orow] = cnt;
        a.used[orow] = drawn;
    }
}

// the chosen jobs' results into their slices' rows: pick[k] = (slice, job)
__global__ void k_take_jobs(const int2* __restrict__ pick, int maxC, const int* __restrict__ jrep,
                            const float* __restrict__ jw, const int* __restrict__ jcount, int* __restrict__ rep,
                            float* __restrict__ w, int* __restrict__ count) {
    const int2 pk = pick[blockIdx.x];
    for (int c = threadIdx.x; c < maxC; c += blockDim.x) {
        rep[(int64_t)pk.x * maxC + c] = jrep[(int64_t)pk.y * maxC + c];
        w[(int64_t)pk.x * maxC + c] = jw[(int64_t)pk.y * maxC + c];
    }
    if (threadIdx.x == 0) count[pk.x] = jcount[pk.y];
}

// The chain of draw offsets: slice s starts where slice s - 1 stopped.  A
// speculative pass runs jobs (slice, offset): the first unresolved slice at
// its exact offset, and each later slice at every offset within z standard
// deviations of its estimate (the exact offset plus the mean draws each slice
// in between took on its runs so far; the error's deviation grows with the
// square root of the distance).  Counts are even (2 or 4 numbers a split), so
// offsets step by 2.  The host then follows the chain through the jobs whose
// offset turned out right.
struct RefineChain {
    int64_t per;
    std::vector<double> sum, sq;
    std::vector<int> runs;
    explicit RefineChain(int S, int64_t per_) : per(per_), sum(S, 0.0), sq(S, 0.0), runs(S, 0) {}
    void observe(int s, int used) {
        sum[s] += used;
        sq[s] += (double)used * used;
        ++runs[s];
    }
    double mean(int s) const {
        if (runs[s] > 0) return sum[s] / runs[s];
        double a = 0;
        int n = 0;
        for (size_t k = 0; k < runs.size(); ++k) { a += sum[k]; n += runs[k]; }
        return n > 0 ? a / n : (double)per;
    }
    // deviation of one slice's count about its mean: pooled over slices run
    // more than once, else the spread between slices
    double sigma() const {
        double within = 0, all = 0, all2 = 0;
        int dof = 0, n = 0;
        for (size_t k = 0; k < runs.size(); ++k) {
            if (runs[k] == 0) continue;
            const double m = sum[k] / runs[k];
            within += sq[k] - runs[k] * m * m;
            dof += runs[k] - 1;
            all += sum[k];
            all2 += sq[k];
            n += runs[k];
        }
        if (dof >= 16) return std::sqrt(std::max(0.0, within / dof));
        if (n < 2) return 0.0;
        const double m = all / n;
        return std::sqrt(std::max(0.0, all2 / n - m * m));
    }
};

constexpr double kRefSpecZ = 2.0;

}  // namespace rzk

using namespace rzk;

int64_t refine_draws_needed(int S, int ncl, int maxC) { return 4 * (int64_t)(maxC - ncl) * S; }

// Per-slice refinement of the ncl kept clusters (members d_mem by d_off,
// representatives/weights d_grep/d_gw) into at most maxC clusters per slice.
// h_draws: the MT numbers that follow the clustering's, at least
// refine_draws_needed(S, ncl, maxC) of them; outputs d_rep/d_w (S x maxC) and
// d_count (S).  Returns the number of passes (1 when every slice took 4
// numbers per split).
int launch_refine_slices(rzk_ctx* ctx, const float* d_R, int S, const int* d_nb, int nnb, int ncl, const int* d_off,
                          const int* d_mem, int nmem, const int* d_grep, const float* d_gw, const float* d_mags,
                          const float* d_norms_R, int maxC, const std::vector<float>& h_draws_in, int* d_rep,
                          float* d_w, int* d_count) {
    RZK_CHECK(nnb >= 1 && nnb <= kRefMaxNb, RZK_E_INVALID, "refinement supports 1 to 32 neighbour slices");
    RZK_CHECK(ncl >= 1 && maxC >= ncl && nmem >= ncl, RZK_E_INVALID, "bad refinement sizes");
    RZK_CHECK((int64_t)h_draws_in.size() >= refine_draws_needed(S, ncl, maxC), RZK_E_INVALID,
              "too few random numbers for the refinement");
    hipStream_t st = ctx->stream;
    const int M = 3 * nnb;
    const int64_t per = 4 * (int64_t)(maxC - ncl);         // draws per slice under the rule
    const std::vector<float> h_draws = h_draws_in.empty() ? std::vector<float>(1, 0.f) : h_draws_in;
    std::vector<int64_t> h_off(S);
    for (int s = 0; s < S; ++s) h_off[s] = per * s;
    // per slice in flight: its slice-matrix copy and state rows; as many slices
    // as half the free memory (at most 16 GiB) holds
    const size_t row_bytes = (size_t)nmem * (20 + 4 * (size_t)M) + (size_t)maxC * 20 + (size_t)ncl * 12 + 20 * 256;
    size_t free_b = 0, total_b = 0;
    RZK_HIP(hipMemGetInfo(&free_b, &total_b));
    const size_t budget = std::min(free_b / 2, (size_t)16 << 30);
    const int batch = (int)std::max<size_t>(1, std::min<size_t>((size_t)S, budget / row_bytes));
    const bool resident = batch >= S;                      // every slice's rows at once: copies built once
    // 20 regions below, each rounded up to 256 bytes
    // speculative jobs per pass (resident only): one block per compute unit, at
    // least one per slice; their results, slices, offsets and the picks
    const int B = resident ? std::max(S, std::max(ctx->num_cus, 1)) : 0;
    const size_t job_bytes = (size_t)B * ((size_t)maxC * 8 + 4 + 4 + 4 + 8) + sizeof(int2) * (size_t)S + 7 * 256;
    // scratch rows: a batch of slices, or the jobs of a pass
    const int rows = resident ? B : batch;
    const size_t extra_rows = (size_t)(rows - batch) * ((size_t)nmem * 20 + (size_t)maxC * 20 + (size_t)ncl * 12 + 20 * 256);
    const size_t shared_bytes = sizeof(float) * h_draws.size() + sizeof(int64_t) * S + sizeof(int) * S +
                                sizeof(float) * 3 * (size_t)nmem + 8 * 256 + job_bytes + extra_rows;
    void* raw = nullptr;
    RZK_HIP(hipMalloc(&raw, shared_bytes + row_bytes * batch));
    std::unique_ptr<void, void (*)(void*)> hold(raw, [](void* q) { (void)hipFree(q); });
    char* p = static_cast<char*>(raw);
    auto take = [&](size_t bytes) { char* q = p; p += (bytes + 255) & ~(size_t)255; return (void*)q; };
    RefineArgs a{};
    a.S = S; a.nb = d_nb; a.nnb = nnb; a.ncl = ncl; a.off = d_off; a.mem = d_mem; a.nmem = nmem;
    a.grep = d_grep; a.gw = d_gw; a.maxC = maxC;
    float* d_draws = (float*)take(sizeof(float) * h_draws.size());
    int64_t* d_doff = (int64_t*)take(sizeof(int64_t) * S);
    int* d_used = (int*)take(sizeof(int) * S);
    float* d_mm = (float*)take(sizeof(float) * (size_t)nmem);
    float* d_nn = (float*)take(sizeof(float) * (size_t)nmem);
    int* d_iota = (int*)take(sizeof(int) * (size_t)nmem);
    a.draws = d_draws; a.draw_off = d_doff; a.rep = d_rep; a.w = d_w; a.count = d_count; a.used = d_used;
    a.mags = d_mm; a.norms_R = d_nn;
    float* d_SM = (float*)take(sizeof(float) * (size_t)nmem * M * batch);
    a.SM = d_SM;
    a.perm = (int*)take(sizeof(int) * (size_t)nmem * rows);
    a.perm_tmp = (int*)take(sizeof(int) * (size_t)nmem * rows);
    a.nrm = (float*)take(sizeof(float) * (size_t)nmem * rows);
    a.line_t = (float*)take(sizeof(float) * (size_t)nmem * rows);
    a.line_k = (int*)take(sizeof(int) * (size_t)nmem * rows);
    a.cst = (int*)take(sizeof(int) * (size_t)maxC * rows);
    a.csz = (int*)take(sizeof(int) * (size_t)maxC * rows);
    a.cflag = (int*)take(sizeof(int) * (size_t)maxC * rows);
    a.hkey = (float*)take(sizeof(float) * (size_t)maxC * rows);
    a.hval = (int*)take(sizeof(int) * (size_t)maxC * rows);
    float* d_cost0 = (float*)take(sizeof(float) * (size_t)ncl * rows);
    float* d_hkey0 = (float*)take(sizeof(float) * (size_t)ncl * rows);
    int* d_hval0 = (int*)take(sizeof(int) * (size_t)ncl * rows);
    int* d_jrep = B ? (int*)take(sizeof(int) * (size_t)maxC * B) : nullptr;
    float* d_jw = B ? (float*)take(sizeof(float) * (size_t)maxC * B) : nullptr;
    int* d_jcount = B ? (int*)take(sizeof(int) * (size_t)B) : nullptr;
    int* d_jused = B ? (int*)take(sizeof(int) * (size_t)B) : nullptr;
    int* d_jslice = B ? (int*)take(sizeof(int) * (size_t)B) : nullptr;
    int64_t* d_joff = B ? (int64_t*)take(sizeof(int64_t) * (size_t)B) : nullptr;
    int2* d_pick = B ? (int2*)take(sizeof(int2) * (size_t)S) : nullptr;
    a.iota = d_iota; a.cost0 = d_cost0; a.hkey0 = d_hkey0; a.hval0 = d_hval0;
    RZK_CHECK((size_t)(p - static_cast<char*>(raw)) <= shared_bytes + row_bytes * batch, RZK_E_STATE,
              "refinement scratch layout overflow");
    RZK_HIP(hipMemcpyAsync(d_draws, h_draws.data(), sizeof(float) * h_draws.size(), hipMemcpyHostToDevice, st));
    std::vector<int> h_iota(nmem);
    for (int e = 0; e < nmem; ++e) h_iota[e] = e;
    RZK_HIP(hipMemcpyAsync(d_iota, h_iota.data(), sizeof(int) * (size_t)nmem, hipMemcpyHostToDevice, st));
    const unsigned gm = (unsigned)((nmem + 255) / 256);
    hipLaunchKernelGGL(k_gather_members, dim3(gm), dim3(256), 0, st, d_mags, d_mem, nmem, d_mm);
    hipLaunchKernelGGL(k_gather_members, dim3(gm), dim3(256), 0, st, d_norms_R, d_mem, nmem, d_nn);
    RZK_HIP(hipGetLastError());
    // the cluster table and cost heap go to LDS when they fit beside the rest
    const bool lds_cl = ref_layout(nnb, maxC, true).bytes <= kRefLdsMax;
    const RefLayout lay = ref_layout(nnb, maxC, lds_cl);
    RZK_CHECK(lay.bytes <= kRefLdsMax, RZK_E_INVALID, "refinement LDS layout exceeds a compute unit");
    a.lds_clusters = lds_cl ? 1 : 0;
    const size_t lds = lay.bytes, lds_cost = lay.oVal;
    RZK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_refine_slices),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    RZK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_initial_costs),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_cost));
    // per-row data that does not depend on the draws, for slices s0 .. s0 + n - 1
    // in rows s0 - batch0 ..: the slice-matrix copies, the kept clusters'
    // costs and the heap they make
    auto prepare_rows = [&](int s0, int n, int batch0) {
        const unsigned gx = (unsigned)(((int64_t)nmem * M + 255) / 256);
        for (int y0 = 0; y0 < n; y0 += 65535) {
            const int ny = std::min(65535, n - y0);
            hipLaunchKernelGGL(k_slice_columns, dim3(gx, (unsigned)ny), dim3(256), 0, st, d_R, S, d_nb, nnb, d_mem, nmem,
                               s0 + y0, d_SM + (size_t)(s0 + y0 - batch0) * nmem * M);
            RefineArgs b = a;
            b.s0 = s0 + y0;
            b.batch0 = batch0;
            hipLaunchKernelGGL(k_initial_costs, dim3((unsigned)ncl, (unsigned)ny), dim3(kRefBlock), lds_cost, st, b);
            hipLaunchKernelGGL(k_initial_heap, dim3((unsigned)ny), dim3(64), sizeof(float) * 2 * (size_t)ncl, st,
                               d_cost0 + (size_t)(s0 + y0 - batch0) * ncl, ncl, d_hkey0 + (size_t)(s0 + y0 - batch0) * ncl,
                               d_hval0 + (size_t)(s0 + y0 - batch0) * ncl);
        }
        RZK_HIP(hipGetLastError());
    };
    if (resident) prepare_rows(0, S, 0);
    if (resident) {
        // speculative passes over (slice, offset) jobs
        RefineChain chain(S, per);
        RefineArgs b = a;
        b.s0 = 0;
        b.batch0 = 0;
        b.job_slice = d_jslice;
        b.job_off = d_joff;
        b.rep = d_jrep;
        b.w = d_jw;
        b.count = d_jcount;
        b.used = d_jused;
        std::vector<int> h_js, h_jused;
        std::vector<int64_t> h_jo;
        std::vector<int2> h_pick;
        int first = 0;
        int64_t o_first = 0;
        for (int pass = 0; first < S; ++pass) {
            RZK_CHECK(pass <= S, RZK_E_STATE, "refinement draw offsets did not settle");
            h_js.clear();
            h_jo.clear();
            const double sig = chain.sigma();
            double center = (double)o_first;
            for (int i = 0; first + i < S && (int)h_js.size() < B; ++i) {
                const int sl = first + i;
                if (i > 0) center += chain.mean(sl - 1);
                // offsets lie in [o_first, per * sl] and differ from o_first by an even count
                const int64_t lo = o_first, hi = per * sl;
                int64_t c = o_first + 2 * (int64_t)std::llround((center - (double)o_first) / 2.0);
                c = std::min(std::max(c, lo), hi - ((hi - lo) & 1));
                const int64_t h = i == 0 ? 0 : (int64_t)std::ceil(kRefSpecZ * sig * std::sqrt((double)i) / 2.0);
                // c, c - 2, c + 2, c - 4, ... while they fit
                for (int64_t d = 0; d <= h && (int)h_js.size() < B; ++d) {
                    const int64_t cand[2] = {c - 2 * d, c + 2 * d};
                    for (int q = 0; q < (d == 0 ? 1 : 2) && (int)h_js.size() < B; ++q) {
                        if (cand[q] < lo || cand[q] > hi) continue;
                        h_js.push_back(sl);
                        h_jo.push_back(cand[q]);
                    }
                }
            }
            const int nj = (int)h_js.size();
            RZK_HIP(hipMemcpyAsync(d_jslice, h_js.data(), sizeof(int) * nj, hipMemcpyHostToDevice, st));
            RZK_HIP(hipMemcpyAsync(d_joff, h_jo.data(), sizeof(int64_t) * nj, hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_refine_slices, dim3((unsigned)nj), dim3(kRefBlock), lds, st, b);
            RZK_HIP(hipGetLastError());
            h_jused.resize(nj);
            RZK_HIP(hipMemcpyAsync(h_jused.data(), d_jused, sizeof(int) * nj, hipMemcpyDeviceToHost, st));
            RZK_HIP(hipStreamSynchronize(st));
            // follow the chain from the exact slice through the jobs at the right offsets
            h_pick.clear();
            int k = 0;
            int64_t o = o_first;
            for (int sl = first; sl < S; ++sl) {
                while (k < nj && h_js[k] < sl) ++k;
                int hit = -1;
                for (int q = k; q < nj && h_js[q] == sl; ++q)
                    if (h_jo[q] == o) { hit = q; break; }
                if (hit < 0) break;
                h_pick.push_back(make_int2(sl, hit));
                o += h_jused[hit];
            }
            for (int q = 0; q < nj; ++q) chain.observe(h_js[q], h_jused[q]);
            RZK_CHECK(!h_pick.empty(), RZK_E_STATE, "refinement pass resolved no slice");
            const int np_ = (int)h_pick.size();
            RZK_HIP(hipMemcpyAsync(d_pick, h_pick.data(), sizeof(int2) * np_, hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_take_jobs, dim3((unsigned)np_), dim3(256), 0, st, d_pick, maxC, d_jrep, d_jw, d_jcount,
                               d_rep, d_w, d_count);
            RZK_HIP(hipGetLastError());
            // the picks are read before the host vectors change again
            RZK_HIP(hipStreamSynchronize(st));
            first += np_;
            o_first = o;
            if (first == S) return pass + 1;
        }
        return 0;
    }
    std::vector<int> h_used(S);
    int first = 0;                                         // slices before `first` are final
    for (int pass = 0; first < S; ++pass) {
        RZK_CHECK(pass <= S, RZK_E_STATE, "refinement draw offsets did not settle");
        RZK_HIP(hipMemcpyAsync(d_doff + first, h_off.data() + first, sizeof(int64_t) * (S - first),
                               hipMemcpyHostToDevice, st));
        for (int b0 = first; b0 < S; b0 += batch) {
            const int nb0 = std::min(batch, S - b0);
            a.s0 = b0;
            a.batch0 = resident ? 0 : b0;
            if (!resident) prepare_rows(b0, nb0, b0);
            hipLaunchKernelGGL(k_refine_slices, dim3(nb0), dim3(kRefBlock), lds, st, a);
            RZK_HIP(hipGetLastError());
        }
        RZK_HIP(hipMemcpyAsync(h_used.data() + first, d_used + first, sizeof(int) * (S - first),
                               hipMemcpyDeviceToHost, st));
        RZK_HIP(hipStreamSynchronize(st));
        // the first slice whose successor read from a wrong offset
        int wrong = S;
        for (int s = first; s + 1 < S; ++s)
            if (h_off[s] + h_used[s] != h_off[s + 1]) { wrong = s + 1; break; }
        if (wrong == S) return pass + 1;
        h_off[wrong] = h_off[wrong - 1] + h_used[wrong - 1];
        for (int s = wrong + 1; s < S; ++s) h_off[s] = h_off[s - 1] + per;
        first = wrong;
    }
    return 0;
}
