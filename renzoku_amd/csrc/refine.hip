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
// are gathered into LDS 64 at a time; per-column sums (norms, projections,
// products) run one lane per column, per-row sums one lane per row, and the
// sums over the cluster (the cost, the norm total, the pdf total and cdf) run
// as one identical chain in every lane.  The sort of the projections and the
// cost heap follow std::sort and std::priority_queue move for move (one lane,
// std_order.h), because the reference's result depends on their tie orders.
//
// Draws: the reference takes them from one MT stream, slice after slice.  A
// slice takes 4 per split as a rule (2 to pick the line, 1 per new
// representative) but fewer when a split cluster has two members or the
// refinement stops early, so slice s reads the pre-generated stream from an
// offset that assumes the rule for every earlier slice; the host checks the
// counts afterwards and re-runs, with corrected offsets, the slices after the
// first one that took a different number.
#pragma clang fp contract(off)

#include <algorithm>
#include <climits>
#include <cstring>

#include "host_sampling.h"
#include "rzk_device.h"
#include "std_order.h"

namespace rzk {

constexpr int kRefBlock = 256;
constexpr int kRefChunk = 64;          // columns gathered per pass
constexpr int kRefSortLds = 4096;      // projections sorted in LDS (longer lines sort in HBM scratch)
constexpr int kRefMaxNb = 32;          // neighbour rows per slice (3 * 32 flattened rows)

struct RefineArgs {
    const float* R;          // column-major Rgb reduced matrix, R[(k * S + r) * 3 + c]
    int S;
    const int* nb;           // S x nnb neighbour rows
    int nnb;
    int ncl;                 // kept global clusters
    const int* off;          // ncl + 1 member offsets
    const int* mem;          // members, cluster by cluster, increasing VPL order
    int nmem;
    const int* grep;
    const float* gw;
    const float* mags;       // flattened column magnitudes (pdfs of the global clusters)
    const float* norms_R;    // column norms (pdfs of split clusters)
    int maxC;
    const float* draws;      // pre-generated MT numbers
    const int64_t* draw_off; // per slice
    int s0;                  // first slice of the launch
    int batch0;              // slice index of scratch row 0
    // per-slice results (S x maxC, S)
    int* rep;
    float* w;
    int* count;
    int* used;               // draws taken
    // per-slice state (scratch rows of this batch)
    int* perm;               // nmem
    int* perm_tmp;           // nmem
    float* nrm;              // nmem
    float* line_t;           // nmem (long lines)
    int* line_k;             // nmem
    int* cst;                // maxC
    int* csz;
    int* cflag;              // 0: pdf over mags (global cluster), 1: over norms_R
    float* hkey;             // maxC
    int* hval;
};

struct RefineShared {
    float* sX;      // kRefChunk x ldx gathered columns
    int ldx;
    float* sU;      // flattened rows: row sums / the line direction
    float* sA;      // column a
    float* sT;      // kRefBlock chain staging
    float* sKey;    // kRefSortLds
    int* sVal;
    int* sInt;      // small broadcast area
};

// gather columns perm[st + c0 .. st + c0 + cnt) of the slice matrix into sX
__device__ __forceinline__ void ref_stage(const RefineArgs& a, const int* nbs, const int* perm, int st, int c0,
                                          int cnt, const RefineShared& sh) {
    const int M = 3 * a.nnb;
    for (int e = threadIdx.x; e < cnt * M; e += kRefBlock) {
        int col = e / M, r = e - col * M, q = r / 3, ch = r - 3 * q;
        int k = perm[st + c0 + col];
        sh.sX[col * sh.ldx + r] = a.R[((int64_t)k * a.S + nbs[q]) * 3 + ch];
    }
}

// compute_distance (lightslice.cc:336-357) of columns perm[st .. st + n)
__device__ float ref_cost(const RefineArgs& a, const int* nbs, const int* perm, float* nrm, int st, int n,
                          const RefineShared& sh) {
    const int M = 3 * a.nnb, tid = threadIdx.x;
    float u = 0.f, c = 0.f;
    // pass 1: column norms (get_column_norm), row sums (row_sum), their total (sum)
    for (int c0 = 0; c0 < n; c0 += kRefChunk) {
        const int cnt = min(kRefChunk, n - c0);
        ref_stage(a, nbs, perm, st, c0, cnt, sh);
        __syncthreads();
        if (tid < cnt) {
            float mag = 0.f;
            for (int r = 0; r < M; ++r) {
                float x = sh.sX[tid * sh.ldx + r];
                mag += x * x;
            }
            float nv = sqrtf(mag);
            nrm[c0 + tid] = nv;
            sh.sT[tid] = nv;
        }
        if (tid < M)
            for (int col = 0; col < cnt; ++col) u += sh.sX[col * sh.ldx + tid];
        __syncthreads();
        for (int k = 0; k < cnt; ++k) c += sh.sT[k];
        __syncthreads();
    }
    if (tid < M) sh.sU[tid] = u;
    __syncthreads();
    // pass 2: v_j = sum_r x_rj u_r (Rt.mul), cost = sum_j (norm_j * c - v_j)
    float cost = 0.f;
    for (int c0 = 0; c0 < n; c0 += kRefChunk) {
        const int cnt = min(kRefChunk, n - c0);
        ref_stage(a, nbs, perm, st, c0, cnt, sh);
        __syncthreads();
        if (tid < cnt) {
            float v = 0.f;
            for (int r = 0; r < M; ++r) v += sh.sX[tid * sh.ldx + r] * sh.sU[r];
            sh.sT[tid] = nrm[c0 + tid] * c - v;
        }
        __syncthreads();
        for (int k = 0; k < cnt; ++k) cost += sh.sT[k];
        __syncthreads();
    }
    return cost;
}

// DiscretePdf over src[perm[st + k]] (k < n) sampled at nu (1 or 2) numbers:
// out_k[i] = first k with u[i] < cdf[k] (-1 if none), out_p[i] = pdf[k]
// (pdf.cc:18-77: ordered total, pdf = a / total, cdf running sum, last = 1).
__device__ void ref_pdf_sample(const float* src, const int* perm, int st, int n, int nu, const float* u,
                               int* out_k, float* out_p, const RefineShared& sh) {
    const int tid = threadIdx.x;
    float sum = 0.f;
    for (int c0 = 0; c0 < n; c0 += kRefBlock) {
        const int cnt = min(kRefBlock, n - c0);
        sh.sT[tid] = tid < cnt ? src[perm[st + c0 + tid]] : 0.f;
        __syncthreads();
        for (int k = 0; k < cnt; ++k) sum += sh.sT[k];
        __syncthreads();
    }
    if (tid == 0) { sh.sInt[0] = INT_MAX; sh.sInt[1] = INT_MAX; }
    __syncthreads();
    if (sum == 0.0f) {
        for (int i = 0; i < nu; ++i) { out_k[i] = -1; out_p[i] = 0.f; }
        return;
    }
    float run = 0.f;
    for (int c0 = 0; c0 < n; c0 += kRefBlock) {
        const int cnt = min(kRefBlock, n - c0);
        const float p = tid < cnt ? src[perm[st + c0 + tid]] / sum : 0.f;
        sh.sT[tid] = p;
        __syncthreads();
        float mine = 0.f;
        for (int k = 0; k < cnt; ++k) {
            run = (c0 + k == 0) ? sh.sT[0] : run + sh.sT[k];
            if (k == tid) mine = run;
        }
        if (c0 + tid == n - 1) mine = 1.0f;
        if (tid < cnt)
            for (int i = 0; i < nu; ++i)
                if (u[i] < mine) atomicMin(&sh.sInt[i], c0 + tid);
        __syncthreads();
        if (sh.sInt[0] != INT_MAX && (nu < 2 || sh.sInt[1] != INT_MAX)) break;    // uniform
        __syncthreads();
    }
    __syncthreads();
    for (int i = 0; i < nu; ++i) {
        int k = sh.sInt[i];
        if (k == INT_MAX) { out_k[i] = -1; out_p[i] = 0.f; }
        else { out_k[i] = k; out_p[i] = src[perm[st + k]] / sum; }
    }
    __syncthreads();
}

// split_cluster (lightslice.cc:405-488) of perm[st .. st + n), n >= 2: the
// range is rewritten as indices1 then indices2; returns |indices1|.
__device__ int ref_split(const RefineArgs& a, const int* nbs, int* perm, int* perm_tmp, float* line_t,
                         int* line_k, int st, int n, int flag, int64_t draw0, int& drawn, const RefineShared& sh) {
    const int M = 3 * a.nnb, tid = threadIdx.x;
    if (n == 2) return 1;
    float u[2] = {a.draws[draw0 + drawn], a.draws[draw0 + drawn + 1]};
    drawn += 2;
    int ij[2];
    float pij[2];
    ref_pdf_sample(flag ? a.norms_R : a.mags, perm, st, n, 2, u, ij, pij, sh);
    const int i = ij[0], j = ij[1];
    if (i < 0 || j < 0) return n / 2;                                    // split by half, order kept
    // a, a - b
    for (int r = tid; r < M; r += kRefBlock) {
        int q = r / 3, ch = r - 3 * q;
        float av = a.R[((int64_t)perm[st + i] * a.S + nbs[q]) * 3 + ch];
        float bv = a.R[((int64_t)perm[st + j] * a.S + nbs[q]) * 3 + ch];
        sh.sA[r] = av;
        sh.sU[r] = av - bv;
    }
    __syncthreads();
    // t_k = (a - c_k) . (a - b) for k != i, j, at line position k - [k > i] - [k > j]
    const int nl = n - 2;
    float* key = nl <= kRefSortLds ? sh.sKey : line_t;
    int* val = nl <= kRefSortLds ? sh.sVal : line_k;
    for (int c0 = 0; c0 < n; c0 += kRefChunk) {
        const int cnt = min(kRefChunk, n - c0);
        ref_stage(a, nbs, perm, st, c0, cnt, sh);
        __syncthreads();
        if (tid < cnt) {
            const int k = c0 + tid;
            if (k != i && k != j) {
                float t = 0.0f;
                for (int r = 0; r < M; ++r) t += (sh.sA[r] - sh.sX[tid * sh.ldx + r]) * sh.sU[r];
                const int p = k - (k > i ? 1 : 0) - (k > j ? 1 : 0);
                key[p] = t;
                val[p] = k;
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        stdord::sort_pairs(key, val, nl);
        // largest gap, first on ties, only if positive
        float max_dist = 0.0f;
        int max_k = -1;
        for (int k = 1; k < nl; ++k) {
            float d = key[k] - key[k - 1];
            if (d > max_dist) { max_dist = d; max_k = k; }
        }
        sh.sInt[2] = max_k;
    }
    __syncthreads();
    const int cut = sh.sInt[2];
    if (cut < 0) return n / 2;
    // indices1 = [i, line[0 .. cut)], indices2 = [j, line[cut ..)]
    for (int p = tid; p < n; p += kRefBlock) {
        int src;
        if (p == 0) src = i;
        else if (p <= cut) src = val[p - 1];
        else if (p == cut + 1) src = j;
        else src = val[p - 2];
        perm_tmp[p] = perm[st + src];
    }
    __syncthreads();
    for (int p = tid; p < n; p += kRefBlock) perm[st + p] = perm_tmp[p];
    __syncthreads();
    return cut + 1;
}

__global__ __launch_bounds__(kRefBlock) void k_refine_slices(RefineArgs a) {
    extern __shared__ float s_ref[];
    const int M = 3 * a.nnb, tid = threadIdx.x;
    RefineShared sh;
    sh.ldx = M + 1;
    sh.sX = s_ref;
    sh.sU = sh.sX + kRefChunk * sh.ldx;
    sh.sA = sh.sU + 3 * kRefMaxNb;
    sh.sT = sh.sA + 3 * kRefMaxNb;
    sh.sKey = sh.sT + kRefBlock;
    sh.sVal = reinterpret_cast<int*>(sh.sKey + kRefSortLds);
    sh.sInt = sh.sVal + kRefSortLds;

    const int s = a.s0 + (int)blockIdx.x;
    if (s >= a.S) return;
    const int row = s - a.batch0;
    const int* nbs = a.nb + (int64_t)s * a.nnb;
    int* perm = a.perm + (int64_t)row * a.nmem;
    int* perm_tmp = a.perm_tmp + (int64_t)row * a.nmem;
    float* nrm = a.nrm + (int64_t)row * a.nmem;
    float* line_t = a.line_t + (int64_t)row * a.nmem;
    int* line_k = a.line_k + (int64_t)row * a.nmem;
    int* cst = a.cst + (int64_t)row * a.maxC;
    int* csz = a.csz + (int64_t)row * a.maxC;
    int* cflag = a.cflag + (int64_t)row * a.maxC;
    float* hkey = a.hkey + (int64_t)row * a.maxC;
    int* hval = a.hval + (int64_t)row * a.maxC;
    int* rep = a.rep + (int64_t)s * a.maxC;
    float* w = a.w + (int64_t)s * a.maxC;
    const int64_t draw0 = a.draw_off[s];

    for (int e = tid; e < a.nmem; e += kRefBlock) perm[e] = a.mem[e];
    for (int c = tid; c < a.ncl; c += kRefBlock) {
        cst[c] = a.off[c];
        csz[c] = a.off[c + 1] - a.off[c];
        cflag[c] = 0;
        rep[c] = a.grep[c];
        w[c] = a.gw[c];
    }
    __syncthreads();
    int hn = 0;                    // heap size (lane 0's copy is the one used)
    for (int c = 0; c < a.ncl; ++c) {
        float cost = ref_cost(a, nbs, perm, nrm, cst[c], csz[c], sh);
        if (tid == 0) stdord::heap_push(hkey, hval, hn, cost, c);
    }
    int cnt = a.ncl, drawn = 0;
    for (;;) {
        __syncthreads();
        if (tid == 0) {
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
        const int n1 = ref_split(a, nbs, perm, perm_tmp, line_t, line_k, st, n, flag, draw0, drawn, sh);
        const float cost1 = ref_cost(a, nbs, perm, nrm, st, n1, sh);
        const float cost2 = ref_cost(a, nbs, perm, nrm, st + n1, n - n1, sh);
        int k1, k2;
        float p1, p2;
        float u1 = a.draws[draw0 + drawn], u2 = a.draws[draw0 + drawn + 1];
        drawn += 2;
        ref_pdf_sample(a.norms_R, perm, st, n1, 1, &u1, &k1, &p1, sh);
        ref_pdf_sample(a.norms_R, perm, st + n1, n - n1, 1, &u2, &k2, &p2, sh);
        if (tid == 0) {
            const int n2 = n - n1;
            csz[top] = n1;
            cflag[top] = 1;
            rep[top] = k1 < 0 ? perm[st + n1 / 2] : perm[st + k1];
            w[top] = k1 < 0 ? 0.0f : 1.0f / p1;
            cst[cnt] = st + n1;
            csz[cnt] = n2;
            cflag[cnt] = 1;
            rep[cnt] = k2 < 0 ? perm[st + n1 + n2 / 2] : perm[st + n1 + k2];
            w[cnt] = k2 < 0 ? 0.0f : 1.0f / p2;
            stdord::heap_pop(hkey, hval, hn);
            stdord::heap_push(hkey, hval, hn, cost1, top);
            stdord::heap_push(hkey, hval, hn, cost2, cnt);
        }
        ++cnt;
    }
    if (tid == 0) {
        a.count[s] = cnt;
        a.used[s] = drawn;
    }
}

}  // namespace rzk

using namespace rzk;

size_t refine_lds_bytes(int nnb) {
    const int M = 3 * nnb;
    return sizeof(float) * ((size_t)kRefChunk * (M + 1) + 2 * 3 * kRefMaxNb + kRefBlock + kRefSortLds) +
           sizeof(int) * ((size_t)kRefSortLds + 8);
}

int64_t refine_draws_needed(int S, int ncl, int maxC) { return 4 * (int64_t)(maxC - ncl) * S; }

// Per-slice refinement of the ncl kept clusters (members d_mem by d_off,
// representatives/weights d_grep/d_gw) into at most maxC clusters per slice.
// h_draws: the MT numbers that follow the clustering's, at least
// refine_draws_needed(S, ncl, maxC) of them; outputs d_rep/d_w (S x maxC) and
// d_count (S).
void launch_refine_slices(rzk_ctx* ctx, const float* d_R, int S, const int* d_nb, int nnb, int ncl, const int* d_off,
                          const int* d_mem, int nmem, const int* d_grep, const float* d_gw, const float* d_mags,
                          const float* d_norms_R, int maxC, const std::vector<float>& h_draws_in, int* d_rep,
                          float* d_w, int* d_count) {
    RZK_CHECK(nnb >= 1 && nnb <= kRefMaxNb, RZK_E_INVALID, "refinement supports 1 to 32 neighbour slices");
    RZK_CHECK(ncl >= 1 && maxC >= ncl && nmem >= ncl, RZK_E_INVALID, "bad refinement sizes");
    RZK_CHECK((int64_t)h_draws_in.size() >= refine_draws_needed(S, ncl, maxC), RZK_E_INVALID,
              "too few random numbers for the refinement");
    hipStream_t st = ctx->stream;
    const int64_t per = 4 * (int64_t)(maxC - ncl);         // draws per slice under the rule
    const std::vector<float> h_draws = h_draws_in.empty() ? std::vector<float>(1, 0.f) : h_draws_in;
    std::vector<int64_t> h_off(S);
    for (int s = 0; s < S; ++s) h_off[s] = per * s;
    // slice batches bound the scratch to about 1 GiB
    const size_t row_bytes = (size_t)nmem * 20 + (size_t)maxC * 20;
    const int batch = (int)std::max<size_t>(1, std::min<size_t>((size_t)S, ((size_t)1 << 30) / row_bytes));
    // 13 regions below, each rounded up to 256 bytes
    char* scratch = (char*)ctx_scratch(ctx, row_bytes * batch + sizeof(float) * h_draws.size() +
                                               sizeof(int64_t) * S + sizeof(int) * S + 16 * 256);
    char* p = scratch;
    auto take = [&](size_t bytes) { char* q = p; p += (bytes + 255) & ~(size_t)255; return (void*)q; };
    RefineArgs a{};
    a.R = d_R; a.S = S; a.nb = d_nb; a.nnb = nnb; a.ncl = ncl; a.off = d_off; a.mem = d_mem; a.nmem = nmem;
    a.grep = d_grep; a.gw = d_gw; a.mags = d_mags; a.norms_R = d_norms_R; a.maxC = maxC;
    float* d_draws = (float*)take(sizeof(float) * h_draws.size());
    int64_t* d_doff = (int64_t*)take(sizeof(int64_t) * S);
    int* d_used = (int*)take(sizeof(int) * S);
    a.draws = d_draws; a.draw_off = d_doff; a.rep = d_rep; a.w = d_w; a.count = d_count; a.used = d_used;
    a.perm = (int*)take(sizeof(int) * (size_t)nmem * batch);
    a.perm_tmp = (int*)take(sizeof(int) * (size_t)nmem * batch);
    a.nrm = (float*)take(sizeof(float) * (size_t)nmem * batch);
    a.line_t = (float*)take(sizeof(float) * (size_t)nmem * batch);
    a.line_k = (int*)take(sizeof(int) * (size_t)nmem * batch);
    a.cst = (int*)take(sizeof(int) * (size_t)maxC * batch);
    a.csz = (int*)take(sizeof(int) * (size_t)maxC * batch);
    a.cflag = (int*)take(sizeof(int) * (size_t)maxC * batch);
    a.hkey = (float*)take(sizeof(float) * (size_t)maxC * batch);
    a.hval = (int*)take(sizeof(int) * (size_t)maxC * batch);
    RZK_HIP(hipMemcpyAsync(d_draws, h_draws.data(), sizeof(float) * h_draws.size(), hipMemcpyHostToDevice, st));
    const size_t lds = refine_lds_bytes(nnb);
    std::vector<int> h_used(S);
    int first = 0;                                         // slices before `first` are final
    for (int pass = 0; first < S; ++pass) {
        RZK_CHECK(pass <= S, RZK_E_STATE, "refinement draw offsets did not settle");
        RZK_HIP(hipMemcpyAsync(d_doff + first, h_off.data() + first, sizeof(int64_t) * (S - first),
                               hipMemcpyHostToDevice, st));
        for (int b0 = first; b0 < S; b0 += batch) {
            a.s0 = b0;
            a.batch0 = b0;
            hipLaunchKernelGGL(k_refine_slices, dim3(std::min(batch, S - b0)), dim3(kRefBlock), lds, st, a);
            RZK_HIP(hipGetLastError());
        }
        RZK_HIP(hipMemcpyAsync(h_used.data() + first, d_used + first, sizeof(int) * (S - first),
                               hipMemcpyDeviceToHost, st));
        RZK_HIP(hipStreamSynchronize(st));
        // the first slice whose successor read from a wrong offset
        int wrong = S;
        for (int s = first; s + 1 < S; ++s)
            if (h_off[s] + h_used[s] != h_off[s + 1]) { wrong = s + 1; break; }
        if (wrong == S) break;
        h_off[wrong] = h_off[wrong - 1] + h_used[wrong - 1];
        for (int s = wrong + 1; s < S; ++s) h_off[s] = h_off[s - 1] + per;
        first = wrong;
    }
}
