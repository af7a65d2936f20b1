// cluster.hip — reductions over the reduced matrix: column alphas
// (light_transport_matrix.cc:72-106), flattened-column magnitudes
// (cluster_builder.cc:14-19) and the nearest-centre assignment
// (cluster_builder.cc:158-225).
//
// Bit-exactness: every sum keeps the reference's sequential order, so the
// parallelism is across independent sums (columns, rows, (column, centre)
// pairs), never inside one.  R is column-major Rgb: R[(k*rows + r)*3 + c].
#pragma clang fp contract(off)

#include <cfloat>

#include "rzk_device.h"

namespace rzk {

// flatten_magnitude entry (matrix_c.cc:48-59)
__device__ __forceinline__ float entry_mag(const float* e) {
    return sqrtf(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
}

// per column: norm_j = sqrt(Σ_i m_ij^2) and mags_j = sqrt(Σ_t x_t^2) over the
// 3*rows flattened column
__global__ void k_col_norms(int rows, int64_t cols, const float* __restrict__ R, float* __restrict__ norms,
                            float* __restrict__ mags) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cols) return;
    const float* c = R + 3 * j * rows;
    float s = 0.f, s2 = 0.f;
    for (int i = 0; i < rows; ++i) {
        float m = entry_mag(c + 3 * i);
        s += m * m;
        s2 += c[3 * i] * c[3 * i];
        s2 += c[3 * i + 1] * c[3 * i + 1];
        s2 += c[3 * i + 2] * c[3 * i + 2];
    }
    norms[j] = sqrtf(s);
    mags[j] = sqrtf(s2);
}

// per column: sqrt(sum_i (r^2 + g^2 + b^2)), the per-row sum grouped
// (LightTransportMatrix::get_column_norms, light_transport_matrix.cc:117-130);
// the refinement samples representatives by it
__global__ void k_col_norms_R(int rows, int64_t cols, const float* __restrict__ R, float* __restrict__ out) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cols) return;
    const float* c = R + 3 * j * rows;
    float mag = 0.f;
    for (int i = 0; i < rows; ++i) mag += c[3 * i] * c[3 * i] + c[3 * i + 1] * c[3 * i + 1] + c[3 * i + 2] * c[3 * i + 2];
    out[j] = sqrtf(mag);
}

// m^T[i][j] = |R(i, j)|: 64 x 64 tiles through LDS so that both the column
// reads of R and the row writes of m^T are coalesced
__global__ __launch_bounds__(256) void k_mag_transpose(int rows, int64_t cols, const float* __restrict__ R,
                                                       float* __restrict__ mT) {
    __shared__ float tile[64][65];
    int64_t j0 = (int64_t)blockIdx.x * 64;
    int i0 = blockIdx.y * 64;
    for (int e = threadIdx.x; e < 64 * 64; e += 256) {
        int jj = e / 64, ii = e % 64;
        int64_t j = j0 + jj;
        int i = i0 + ii;
        tile[jj][ii] = (j < cols && i < rows) ? entry_mag(R + 3 * (j * rows + i)) : 0.f;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * 64; e += 256) {
        int ii = e / 64, jj = e % 64;
        int64_t j = j0 + jj;
        int i = i0 + ii;
        if (j < cols && i < rows) mT[(int64_t)i * cols + j] = tile[jj][ii];
    }
}

// Ordered sums (matrix_c.h:236-245 / light_transport_matrix.cc:99-100): one
// float add after another, j = 0, 1, ..., for the rows of x (out[i], blocks
// 0 .. rows-1) and, in block `rows`, for the vector x1 (out1[0]); every sum
// has `cols` terms.  One wave per sum.  The chain of dependent adds is the
// bound, so every lane runs the same chain (no lane-0 bottleneck) on values
// it reads from LDS with wave-uniform addresses (broadcast reads, 16 float4
// ahead of the adds), while the global loads of the next 1,024-value chunk
// are in flight.
__global__ __launch_bounds__(64) void k_ordered_sums(int rows, int64_t cols, const float* __restrict__ x,
                                                     float* __restrict__ out, const float* __restrict__ x1,
                                                     float* __restrict__ out1) {
    __shared__ float4 buf[2][256];
    const int i = blockIdx.x;
    if (i > rows || (i == rows && !x1)) return;
    const float* row = i < rows ? x + (int64_t)i * cols : x1;
    const int lane = threadIdx.x;
    const int64_t nchunk = (cols + 1023) / 1024;
    float4 pre[4];
    auto fetch = [&](int64_t c) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            int64_t e = c * 1024 + (int64_t)(q * 64 + lane) * 4;
            pre[q].x = e < cols ? row[e] : 0.f;
            pre[q].y = e + 1 < cols ? row[e + 1] : 0.f;
            pre[q].z = e + 2 < cols ? row[e + 2] : 0.f;
            pre[q].w = e + 3 < cols ? row[e + 3] : 0.f;
        }
    };
    auto put = [&](int slot) {
#pragma unroll
        for (int q = 0; q < 4; ++q) buf[slot][q * 64 + lane] = pre[q];
    };
    float acc = 0.f;
    if (nchunk > 0) {
        fetch(0);
        put(0);
        __syncthreads();
    }
    for (int64_t c = 0; c < nchunk; ++c) {
        const int slot = (int)(c & 1);
        if (c + 1 < nchunk) fetch(c + 1);
        const int64_t n = cols - c * 1024 < 1024 ? cols - c * 1024 : 1024;
        if (n == 1024) {
            for (int q = 0; q < 256; q += 16) {
                float4 v[16];
#pragma unroll
                for (int k = 0; k < 16; ++k) v[k] = buf[slot][q + k];
#pragma unroll
                for (int k = 0; k < 16; ++k) {
                    acc += v[k].x;
                    acc += v[k].y;
                    acc += v[k].z;
                    acc += v[k].w;
                }
            }
        } else {
            for (int64_t e = 0; e < n; ++e) {
                float4 v = buf[slot][e >> 2];
                int k = (int)(e & 3);
                acc += k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w;
            }
        }
        if (c + 1 < nchunk) put(slot ^ 1);
        __syncthreads();
    }
    if (lane == 0) {
        if (i < rows) out[i] = acc;
        else out1[0] = acc;
    }
}

// per column: v_j = Σ_i m_ij * u_i, i sequential (matrix_c.h:224-233 on Rt)
__global__ void k_col_dot_u(int rows, int64_t cols, const float* __restrict__ R, const float* __restrict__ u,
                            float* __restrict__ v) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cols) return;
    const float* c = R + 3 * j * rows;
    float acc = 0.f;
    for (int i = 0; i < rows; ++i) acc += entry_mag(c + 3 * i) * u[i];
    v[j] = acc;
}

// alpha_j = norm_j * c0 - v_j, c0 = Σ_j norm_j in order (k_ordered_sums, block `rows`)
__global__ void k_alphas(int64_t cols, const float* __restrict__ norms, const float* __restrict__ c0,
                         const float* __restrict__ v, float* __restrict__ alphas) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cols) return;
    alphas[j] = norms[j] * c0[0] - v[j];
}

// K_j = R column(center_j) * weight_j; Knorm_j = sqrt(Σ_r (r^2 + g^2 + b^2))
// (light_transport_matrix.cc:117-130)
__global__ void k_build_K(int rows, int ncl, const float* __restrict__ R, const int* __restrict__ centers,
                          const float* __restrict__ weights, float* __restrict__ K, float* __restrict__ Knorm) {
    int j = blockIdx.x;
    if (j >= ncl) return;
    const float* c = R + 3 * (int64_t)centers[j] * rows;
    float w = weights[j];
    float* kj = K + 3 * (int64_t)j * rows;
    for (int t = threadIdx.x; t < 3 * rows; t += blockDim.x) kj[t] = c[t] * w;
    __syncthreads();
    if (threadIdx.x == 0) {
        float mag = 0.f;
        for (int r = 0; r < rows; ++r) {
            float a = kj[3 * r], b = kj[3 * r + 1], d = kj[3 * r + 2];
            mag += a * a + b * b + d * d;
        }
        Knorm[j] = sqrtf(mag);
    }
}

// Assignment: block = 64 columns x one tile of 64 centres (grid.y = the tile).
// Each of the 128 lanes accumulates an 8 x 4 block of (column, centre) dot
// products over the 3*rows flattened entries in order, staged through LDS in
// chunks of 32 with 16-byte aligned rows, so a step of the chain reads its 8
// column and 4 centre values with three 16-byte LDS loads; a chunk's global
// loads are issued together into registers while the chunk before it is being
// accumulated.  A column's minimum over the tile goes to key[column] by a
// 64-bit atomic min of (distance bits << 32 | centre): distances are finite
// and non-negative, so the smallest key is the smallest distance and, among
// equal ones, the smallest centre, which is what the sequential scan with
// strict < keeps.
constexpr int kAsCols = 64, kAsCen = 64, kAsChunk = 32, kAsPitch = 68, kAsBlock = 128;
constexpr int kAsCw = 8, kAsKw = 4;                       // columns, centres per lane
constexpr int kAsGroups = kAsCen / kAsKw;                 // centre groups (16)
constexpr int kAsPer = kAsChunk * kAsCols / kAsBlock;     // staged values per lane and array (16)
constexpr unsigned long long kAsNone = ~0ull;

__global__ void k_assign_init(int64_t cols, unsigned long long* __restrict__ key) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < cols) key[c] = kAsNone;
}

__global__ void k_assign_final(int64_t cols, const unsigned long long* __restrict__ key, int* __restrict__ assign) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < cols) assign[c] = key[c] == kAsNone ? -1 : (int)(unsigned)(key[c] & 0xffffffffull);
}

__global__ __launch_bounds__(kAsBlock) void k_assign(int rows, int64_t cols, int ncl, const float* __restrict__ R,
                                                     const float* __restrict__ K, const float* __restrict__ Knorm,
                                                     const float* __restrict__ mags,
                                                     unsigned long long* __restrict__ key) {
    // the staged chunks, then (after the last chunk) the groups' minima
    __shared__ __align__(16) float s_lds[2 * kAsChunk * kAsPitch];
    static_assert(2 * kAsGroups * kAsCols <= 2 * kAsChunk * kAsPitch, "minima fit the staging area");
    float (*s_m)[kAsPitch] = reinterpret_cast<float (*)[kAsPitch]>(s_lds);
    float (*s_k)[kAsPitch] = reinterpret_cast<float (*)[kAsPitch]>(s_lds + kAsChunk * kAsPitch);
    float (*s_bd)[kAsCols] = reinterpret_cast<float (*)[kAsCols]>(s_lds);
    int (*s_bj)[kAsCols] = reinterpret_cast<int (*)[kAsCols]>(s_lds + kAsGroups * kAsCols);
    const int n3 = 3 * rows;
    const int64_t col0 = (int64_t)blockIdx.x * kAsCols;
    const int j0 = blockIdx.y * kAsCen;
    const int tc = threadIdx.x & 7;           // column group: columns tc*8 .. tc*8+7
    const int tk = threadIdx.x >> 3;          // centre group: centres tk*4 .. tk*4+3
    // staged element k of this lane: column / centre cc0 + 4k of the tile,
    // entry tt0 of the chunk (e = threadIdx.x + k * kAsBlock, cc = e / kAsChunk)
    static_assert(kAsBlock % kAsChunk == 0, "a lane keeps one entry of the chunk");
    constexpr int kAsStep = kAsBlock / kAsChunk;           // columns between a lane's staged values (4)
    const int cc0 = threadIdx.x / kAsChunk, tt0 = threadIdx.x % kAsChunk;
    const int64_t cmax = cols - (col0 + cc0);             // staged columns k with k * kAsStep < cmax exist
    const int jmax = ncl - (j0 + cc0);
    float vm[kAsPer], vk[kAsPer];
    auto load_chunk = [&](int t0) {
        const int t = t0 + tt0;
        const bool tok = t < n3;
        const float* rb = R + (col0 + cc0) * n3 + (tok ? t : 0);
        const float* kb = K + (int64_t)(j0 + cc0) * n3 + (tok ? t : 0);
#pragma unroll
        for (int k = 0; k < kAsPer; ++k) {
            const bool okm = tok && (int64_t)k * kAsStep < cmax, okk = tok && k * kAsStep < jmax;
            const float xm = okm ? rb[(int64_t)k * kAsStep * n3] : 0.f;
            const float xk = okk ? kb[(int64_t)k * kAsStep * n3] : 0.f;
            vm[k] = xm;
            vk[k] = xk;
        }
    };
    float acc[kAsCw][kAsKw];
#pragma unroll
    for (int a = 0; a < kAsCw; ++a)
#pragma unroll
        for (int b = 0; b < kAsKw; ++b) acc[a][b] = 0.f;
    load_chunk(0);
    for (int t0 = 0; t0 < n3; t0 += kAsChunk) {
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kAsPer; ++k) {
            s_m[tt0][cc0 + k * kAsStep] = vm[k];
            s_k[tt0][cc0 + k * kAsStep] = vk[k];
        }
        __syncthreads();
        if (t0 + kAsChunk < n3) load_chunk(t0 + kAsChunk);
        const int tmax = (n3 - t0) < kAsChunk ? (n3 - t0) : kAsChunk;
#pragma unroll 2
        for (int tt = 0; tt < tmax; ++tt) {
            const float4 m0 = *reinterpret_cast<const float4*>(&s_m[tt][tc * kAsCw]);
            const float4 m1 = *reinterpret_cast<const float4*>(&s_m[tt][tc * kAsCw + 4]);
            const float4 kq = *reinterpret_cast<const float4*>(&s_k[tt][tk * kAsKw]);
            const float mv[kAsCw] = {m0.x, m0.y, m0.z, m0.w, m1.x, m1.y, m1.z, m1.w};
            const float kv[kAsKw] = {kq.x, kq.y, kq.z, kq.w};
#pragma unroll
            for (int a = 0; a < kAsCw; ++a)
#pragma unroll
                for (int b = 0; b < kAsKw; ++b) acc[a][b] += mv[a] * kv[b];
        }
    }
    // dist = |Knorm_j * norm_mi - dot|; strict < keeps the first minimum
    float best_d[kAsCw];
    int best_j[kAsCw];
#pragma unroll
    for (int a = 0; a < kAsCw; ++a) {
        best_d[a] = FLT_MAX;
        best_j[a] = -1;
    }
#pragma unroll
    for (int b = 0; b < kAsKw; ++b) {
        int j = j0 + tk * kAsKw + b;
        if (j >= ncl) continue;
        float kn = Knorm[j];
#pragma unroll
        for (int a = 0; a < kAsCw; ++a) {
            int64_t c = col0 + tc * kAsCw + a;
            float nm = c < cols ? mags[c] : 0.f;
            float dist = fabsf(kn * nm - acc[a][b]);
            if (dist < best_d[a]) { best_d[a] = dist; best_j[a] = j; }
        }
    }
    // combine the centre groups: smallest distance, ties -> smallest centre index
    __syncthreads();
#pragma unroll
    for (int a = 0; a < kAsCw; ++a) { s_bd[tk][tc * kAsCw + a] = best_d[a]; s_bj[tk][tc * kAsCw + a] = best_j[a]; }
    __syncthreads();
    if (threadIdx.x < kAsCols) {
        int cc = threadIdx.x;
        float bd = FLT_MAX;
        int bj = -1;
        for (int g = 0; g < kAsGroups; ++g) {
            float d = s_bd[g][cc];
            int j = s_bj[g][cc];
            if (j < 0) continue;
            if (d < bd || (d == bd && (bj < 0 || j < bj))) { bd = d; bj = j; }
        }
        int64_t c = col0 + cc;
        if (c < cols && bj >= 0)
            atomicMin(key + c, ((unsigned long long)__float_as_uint(bd) << 32) | (unsigned long long)(unsigned)bj);
    }
}

}  // namespace rzk

using namespace rzk;

static unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// alphas and magnitudes of all columns; scratch: u (rows), v (cols), norms
// (cols), c0 (1), mT (rows * cols)
void launch_alphas_full(rzk_ctx* ctx, int rows, int64_t cols, const float* d_R, float* d_alphas,
                        float* d_mags, float* d_norms, float* d_u, float* d_v, float* d_c0, float* d_mT) {
    if (cols <= 0 || rows <= 0) return;
    hipLaunchKernelGGL(k_col_norms, dim3(nblk(cols, 256)), dim3(256), 0, ctx->stream, rows, cols, d_R, d_norms,
                       d_mags);
    hipLaunchKernelGGL(k_mag_transpose, dim3(nblk(cols, 64), nblk(rows, 64)), dim3(256), 0, ctx->stream, rows, cols,
                       d_R, d_mT);
    // u_i = Σ_j m_ij (rows of m^T) and c0 = Σ_j norm_j: rows + 1 independent ordered sums in one launch
    hipLaunchKernelGGL(k_ordered_sums, dim3(rows + 1), dim3(64), 0, ctx->stream, rows, cols, d_mT, d_u, d_norms,
                       d_c0);
    hipLaunchKernelGGL(k_col_dot_u, dim3(nblk(cols, 256)), dim3(256), 0, ctx->stream, rows, cols, d_R, d_u, d_v);
    hipLaunchKernelGGL(k_alphas, dim3(nblk(cols, 256)), dim3(256), 0, ctx->stream, cols, d_norms, d_c0, d_v,
                       d_alphas);
    RZK_HIP(hipGetLastError());
}

void launch_column_norms_R(rzk_ctx* ctx, int rows, int64_t cols, const float* d_R, float* d_out) {
    if (cols <= 0 || rows <= 0) return;
    hipLaunchKernelGGL(k_col_norms_R, dim3(nblk(cols, 256)), dim3(256), 0, ctx->stream, rows, cols, d_R, d_out);
    RZK_HIP(hipGetLastError());
}

// d_key: cols 64-bit words of scratch
void launch_assign_full(rzk_ctx* ctx, int rows, int64_t cols, const float* d_R, int ncl, const int* d_centers,
                        const float* d_weights, const float* d_mags, float* d_K, float* d_Knorm,
                        unsigned long long* d_key, int* d_assign) {
    if (cols <= 0 || ncl <= 0) return;
    hipLaunchKernelGGL(k_build_K, dim3(ncl), dim3(256), 0, ctx->stream, rows, ncl, d_R, d_centers, d_weights,
                       d_K, d_Knorm);
    hipLaunchKernelGGL(k_assign_init, dim3(nblk(cols, 256)), dim3(256), 0, ctx->stream, cols, d_key);
    hipLaunchKernelGGL(k_assign, dim3(nblk(cols, kAsCols), (unsigned)((ncl + kAsCen - 1) / kAsCen)), dim3(kAsBlock), 0,
                       ctx->stream, rows, cols, ncl, d_R, d_K, d_Knorm, d_mags, d_key);
    hipLaunchKernelGGL(k_assign_final, dim3(nblk(cols, 256)), dim3(256), 0, ctx->stream, cols, d_key, d_assign);
    RZK_HIP(hipGetLastError());
}
