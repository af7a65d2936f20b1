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

// Assignment: block = 64 columns x all centres in tiles of 64.  Each of the 256
// lanes accumulates a 4x4 block of (column, centre) dot products over the
// 3*rows flattened entries in order, staged through LDS in chunks of 32.
constexpr int kAsCols = 64, kAsCen = 64, kAsChunk = 32;

__global__ __launch_bounds__(256) void k_assign(int rows, int64_t cols, int ncl, const float* __restrict__ R,
                                                const float* __restrict__ K, const float* __restrict__ Knorm,
                                                const float* __restrict__ mags, int* __restrict__ assign) {
    __shared__ float s_m[kAsChunk][kAsCols + 1];
    __shared__ float s_k[kAsChunk][kAsCen + 1];
    __shared__ float s_bd[16][kAsCols];
    __shared__ int s_bj[16][kAsCols];
    const int n3 = 3 * rows;
    const int64_t col0 = (int64_t)blockIdx.x * kAsCols;
    const int tc = threadIdx.x & 15;          // column group: columns tc*4 .. tc*4+3
    const int tk = threadIdx.x >> 4;          // centre group: centres tk*4 .. tk*4+3
    float best_d[4];
    int best_j[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) { best_d[a] = FLT_MAX; best_j[a] = -1; }
    float nm[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        int64_t c = col0 + tc * 4 + a;
        nm[a] = c < cols ? mags[c] : 0.f;
    }
    for (int j0 = 0; j0 < ncl; j0 += kAsCen) {
        float acc[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = 0.f;
        for (int t0 = 0; t0 < n3; t0 += kAsChunk) {
            __syncthreads();
            for (int e = threadIdx.x; e < kAsChunk * kAsCols; e += 256) {
                int cc = e / kAsChunk, tt = e - cc * kAsChunk;
                int64_t c = col0 + cc;
                s_m[tt][cc] = (c < cols && t0 + tt < n3) ? R[c * n3 + t0 + tt] : 0.f;
                int jj = j0 + cc;
                s_k[tt][cc] = (jj < ncl && t0 + tt < n3) ? K[(int64_t)jj * n3 + t0 + tt] : 0.f;
            }
            __syncthreads();
            int tmax = (n3 - t0) < kAsChunk ? (n3 - t0) : kAsChunk;
            for (int tt = 0; tt < tmax; ++tt) {
                float mv[4], kv[4];
#pragma unroll
                for (int a = 0; a < 4; ++a) mv[a] = s_m[tt][tc * 4 + a];
#pragma unroll
                for (int b = 0; b < 4; ++b) kv[b] = s_k[tt][tk * 4 + b];
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b) acc[a][b] += mv[a] * kv[b];
            }
        }
        // dist = |Knorm_j * norm_mi - dot|; strict < keeps the first minimum
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            int j = j0 + tk * 4 + b;
            if (j >= ncl) continue;
            float kn = Knorm[j];
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                float dist = fabsf(kn * nm[a] - acc[a][b]);
                if (dist < best_d[a]) { best_d[a] = dist; best_j[a] = j; }
            }
        }
    }
    // combine the 16 centre groups: smallest distance, ties -> smallest centre index
#pragma unroll
    for (int a = 0; a < 4; ++a) { s_bd[tk][tc * 4 + a] = best_d[a]; s_bj[tk][tc * 4 + a] = best_j[a]; }
    __syncthreads();
    if (threadIdx.x < kAsCols) {
        int cc = threadIdx.x;
        float bd = FLT_MAX;
        int bj = -1;
        for (int g = 0; g < 16; ++g) {
            float d = s_bd[g][cc];
            int j = s_bj[g][cc];
            if (j < 0) continue;
            if (d < bd || (d == bd && (bj < 0 || j < bj))) { bd = d; bj = j; }
        }
        int64_t c = col0 + cc;
        if (c < cols) assign[c] = bj;
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

void launch_assign_full(rzk_ctx* ctx, int rows, int64_t cols, const float* d_R, int ncl, const int* d_centers,
                        const float* d_weights, const float* d_mags, float* d_K, float* d_Knorm,
                        int* d_assign) {
    if (cols <= 0 || ncl <= 0) return;
    hipLaunchKernelGGL(k_build_K, dim3(ncl), dim3(256), 0, ctx->stream, rows, ncl, d_R, d_centers, d_weights,
                       d_K, d_Knorm);
    hipLaunchKernelGGL(k_assign, dim3(nblk(cols, kAsCols)), dim3(256), 0, ctx->stream, rows, cols, ncl, d_R, d_K,
                       d_Knorm, d_mags, d_assign);
    RZK_HIP(hipGetLastError());
}
