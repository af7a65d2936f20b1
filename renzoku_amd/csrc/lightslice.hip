// lightslice.hip — LightSlice with pixel-tile slices (lightslice.h,
// lightslice.cc:589-1345): whole-image primary hits, the reduced matrix on
// the slice-representative rows, alphas, centre sampling, assignment,
// representatives, the per-slice refinement of the clusters (refine.hip),
// per-slice cluster pdfs and luminance maps, and the final per-pixel gather
// over each slice's cluster representatives.  Also the plain whole-image VPL
// render (every VPL a weight-1 cluster).
#pragma clang fp contract(off)

#include <algorithm>
#include <cstring>
#include <memory>

#include "host_sampling.h"
#include "kdslice.h"
#include "pathtrace.h"
#include "rzk_device.h"

void launch_alphas_full(rzk_ctx* ctx, int rows, int64_t cols, const float* d_R, float* d_alphas,
                        float* d_mags, float* d_norms, float* d_u, float* d_v, float* d_c0, float* d_mT);
void launch_assign_full(rzk_ctx* ctx, int rows, int64_t cols, const float* d_R, int ncl, const int* d_centers,
                        const float* d_weights, const float* d_mags, float* d_K, float* d_Knorm,
                        unsigned long long* d_key, int* d_assign);
void launch_column_norms_R(rzk_ctx* ctx, int rows, int64_t cols, const float* d_R, float* d_out);
int64_t refine_draws_needed(int S, int ncl, int maxC);
int launch_refine_slices(rzk_ctx* ctx, const float* d_R, int S, const int* d_nb, int nnb, int ncl, const int* d_off,
                          const int* d_mem, int nmem, const int* d_grep, const float* d_gw, const float* d_mags,
                          const float* d_norms_R, int maxC, const std::vector<float>& h_draws, int* d_rep,
                          float* d_w, int* d_count);

namespace rzk {

__global__ void k_pick_rows(const float* __restrict__ recv, const int* __restrict__ rmat,
                            const int* __restrict__ pix, int rows, float* __restrict__ rrecv,
                            int* __restrict__ rrmat) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    int64_t i = pix[r];
    for (int k = 0; k < RZK_RECV_FLOATS; ++k) rrecv[RZK_RECV_FLOATS * r + k] = recv[RZK_RECV_FLOATS * i + k];
    rrmat[r] = rmat[i];
}

// lightslice.cc:1011-1030 for each (slice s, cluster c): slice_wi = direction
// to the representative, upper = visible && above the shading normal && non-black.
// Entries are [s * sr.width() + c]; clusters beyond a slice's count get upper = 0.
__global__ void k_slice_info(const float* __restrict__ mat, const float* __restrict__ rrecv, int rows,
                             const float* __restrict__ V, const float* __restrict__ vpl,
                             const int* __restrict__ vmat, SliceReps sr, int mode, float thr,
                             float* __restrict__ slice_wi, int* __restrict__ upper) {
    const int C = sr.width();
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)rows * C) return;
    int s = (int)(t / C), c = (int)(t - (int64_t)s * C);
    if (c >= sr.count(s)) {
        st3(slice_wi + 3 * t, mk(0.f, 0.f, 0.f));
        upper[t] = 0;
        return;
    }
    int k = sr.rep(s, c);
    const float* rc = rrecv + RZK_RECV_FLOATS * s;
    const float* vp = vpl + (int64_t)RZK_VPL_FLOATS * k;
    float vis = V[(int64_t)k * rows + s];
    v3 inc = incident_radiance(mat, rc, vp, vmat[k], mode, thr);
    v3 wi = unit(ld3(vp) - ld3(rc));
    st3(slice_wi + 3 * t, wi);
    bool nonblack = inc.x != 0.f || inc.y != 0.f || inc.z != 0.f;
    upper[t] = (vis > 0.0f && dot(ld3(rc + 6), wi) > kZeroEps && nonblack) ? 1 : 0;
}

// lightslice.cc:1114-1179: cluster weight per slice = norm of the
// representative column over the neighbour rows (matrix_c.cc:74-88 order)
// x weight, 0 below the horizon
__global__ void k_slice_pdf_weights(const float* __restrict__ R, int rows, const int* __restrict__ nb,
                                    int nnb, SliceReps sr, const int* __restrict__ upper, float* __restrict__ wout) {
    const int C = sr.width();
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)rows * C) return;
    int s = (int)(t / C), c = (int)(t - (int64_t)s * C);
    if (c >= sr.count(s)) { wout[t] = 0.f; return; }
    const float* col = R + 3 * (int64_t)sr.rep(s, c) * rows;
    float mag = 0.f;
    for (int q = 0; q < nnb; ++q) {
        const float* e = col + 3 * nb[s * nnb + q];
        mag += e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
    }
    float v = sqrtf(mag);
    v *= sr.weight(s, c);
    if (!upper[t]) v = 0.0f;
    wout[t] = v;
}

// DiscretePdf::set_distribution per slice (pdf.cc:18-52): sum of the weights
// in order, pdf = w / sum, cdf[i] = cdf[i-1] + pdf[i], cdf[last] = 1.  One
// wave per slice: the weights pass through LDS 64 at a time, every lane runs
// the same ordered chains (sum, then cdf) and lane k keeps the cdf value of
// element k of the chunk, so writes are coalesced and the adds keep their order.
__global__ __launch_bounds__(64) void k_slice_pdf_norm(int rows, SliceReps sr, const float* __restrict__ w,
                                                       float* __restrict__ pdf, float* __restrict__ cdf) {
    __shared__ float s_v[64];
    const int s = blockIdx.x, lane = threadIdx.x;
    if (s >= rows) return;
    const int C = sr.width(), nreps = sr.count(s);
    const float* a = w + (int64_t)s * C;
    float* p = pdf + (int64_t)s * C;
    float* c = cdf + (int64_t)s * C;
    for (int i = nreps + lane; i < C; i += 64) { p[i] = 0.f; c[i] = 0.f; }
    float sum = 0.f;
    for (int base = 0; base < nreps; base += 64) {
        const int i = base + lane, n = min(64, nreps - base);
        s_v[lane] = i < nreps ? a[i] : 0.f;
        __syncthreads();
        for (int k = 0; k < n; ++k) sum += s_v[k];
        __syncthreads();
    }
    if (sum == 0.0f) {
        for (int i = lane; i < nreps; i += 64) { p[i] = 0.f; c[i] = 0.f; }
        return;
    }
    float run = 0.f;
    for (int base = 0; base < nreps; base += 64) {
        const int i = base + lane, n = min(64, nreps - base);
        const float pv = i < nreps ? a[i] / sum : 0.f;
        if (i < nreps) p[i] = pv;
        s_v[lane] = pv;
        __syncthreads();
        float mine = 0.f;
        for (int k = 0; k < n; ++k) {
            run = (base + k == 0) ? s_v[0] : run + s_v[k];
            if (k == lane) mine = run;
        }
        if (i < nreps) c[i] = i == nreps - 1 ? 1.0f : mine;
        __syncthreads();
    }
}

__global__ void k_fill_int(int* __restrict__ p, int64_t n, int v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// tile-centre pixel of every slice (the row of the reduced matrix)
__global__ void k_tile_centres(int tx, int ty, int tile, int W, int H, int* __restrict__ px, int* __restrict__ py) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= tx * ty) return;
    int x = s % tx, y = s / tx;
    px[s] = min(x * tile + tile / 2, W - 1);
    py[s] = min(y * tile + tile / 2, H - 1);
}

// build_map (lightslice.cc:2058-2147) + DiscretePdf2D::set_distribution
// (pdf.cc:195-233).  One wave per slice.  The clusters' map bins and
// luminances are computed across the lanes (kMapChunk at a time, staged in
// LDS); the scatter then walks the clusters in order, each bin updated only by
// the lane that owns it (bin mod 64) and the total summed by every lane, so
// every map value and the total keep the reference's cluster-order sums.  Rows
// are normalised in parallel.  LDS: res*res bins, res + 1 sums, 2 * kMapChunk.
constexpr int kMapChunk = 256;

__global__ __launch_bounds__(64) void k_slice_maps(int res, SliceReps sr, const float* __restrict__ R, int rows,
                             const float* __restrict__ rrecv, const float* __restrict__ slice_wi,
                             const int* __restrict__ upper, float eps, int valid_all,
                             float* __restrict__ maps, float* __restrict__ col_cdf,
                             float* __restrict__ row_pdf, float* __restrict__ row_cdf,
                             float* __restrict__ totals) {
    extern __shared__ float s_map[];
    float* s_rows = s_map + res * res;
    float* s_lum = s_rows + res + 1;
    int* s_bin = reinterpret_cast<int*>(s_lum + kMapChunk);
    const int s = blockIdx.x, lane = threadIdx.x;
    const int C = sr.width(), nreps = sr.count(s);
    for (int e = lane; e < res * res; e += 64) s_map[e] = 0.f;
    __syncthreads();
    if (valid_all) {
        const float* rc = rrecv + RZK_RECV_FLOATS * s;
        Onb b;
        b.u = ld3(rc + 12); b.v = ld3(rc + 15); b.n = ld3(rc + 18);
        float sum = 0.f;
        for (int c0 = 0; c0 < nreps; c0 += kMapChunk) {
            const int cnt = min(kMapChunk, nreps - c0);
            for (int k = lane; k < cnt; k += 64) {
                const int c = c0 + k;
                const int64_t t = (int64_t)s * C + c;
                int bin = -1;
                float lum = 0.f;
                if (upper[t]) {
                    v3 lwi = world_to_local(b, ld3(slice_wi + 3 * t));
                    if (!(lwi.z <= 0.0f)) {
                        const float* e = R + 3 * ((int64_t)sr.rep(s, c) * rows + s);
                        v3 rad = ld3(e) * sr.weight(s, c);
                        float sx, sy;
                        direction_to_square(lwi, sx, sy);
                        int x = (int)(sx * res), y = (int)(sy * res);
                        if (x >= res) x = res - 1;
                        if (y >= res) y = res - 1;
                        if (x < 0) x = 0;
                        if (y < 0) y = 0;
                        bin = y * res + x;
                        lum = rgb_value(rad);
                    }
                }
                s_bin[k] = bin;
                s_lum[k] = lum;
            }
            __syncthreads();
            for (int k = 0; k < cnt; ++k) {
                const int bin = s_bin[k];
                if (bin < 0) continue;
                const float lum = s_lum[k];
                if ((bin & 63) == lane) s_map[bin] += lum;
                sum += lum;
            }
            __syncthreads();
        }
        if (lane == 0) s_rows[0] = sum;
        __syncthreads();
    }
    if (valid_all) {
        float add = eps * s_rows[0];
        __syncthreads();
        for (int e = threadIdx.x; e < res * res; e += blockDim.x) s_map[e] += add;
        __syncthreads();
    }
    float* mp = maps + (int64_t)s * res * res;
    float* cc = col_cdf + (int64_t)s * res * res;
    for (int i = threadIdx.x; i < res; i += blockDim.x) {
        float rs = 0.f;
        for (int j = 0; j < res; ++j) rs += s_map[i * res + j];
        float run = 0.f;
        for (int j = 0; j < res; ++j) {
            run += s_map[i * res + j];
            cc[i * res + j] = run / rs;
            mp[i * res + j] = s_map[i * res + j];
        }
        s_rows[1 + i] = rs;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float tot = 0.f;
        for (int i = 0; i < res; ++i) tot += s_rows[1 + i];
        totals[s] = tot;
        float sum = 0.f;
        for (int i = 0; i < res; ++i) sum += s_rows[1 + i];
        float* rp = row_pdf + (int64_t)s * res;
        float* rcd = row_cdf + (int64_t)s * res;
        if (sum == 0.0f) {
            for (int i = 0; i < res; ++i) { rp[i] = 0.f; rcd[i] = 0.f; }
        } else {
            for (int i = 0; i < res; ++i) rp[i] = s_rows[1 + i] / sum;
            rcd[0] = rp[0];
            for (int i = 1; i < res; ++i) rcd[i] = rcd[i - 1] + rp[i];
            rcd[res - 1] = 1.0f;
        }
    }
}

}  // namespace rzk

using namespace rzk;

template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    void alloc(size_t count) {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = count;
        RZK_HIP(hipMalloc((void**)&p, sizeof(T) * (count ? count : 1)));
    }
    ~DBuf() { if (p) (void)hipFree(p); }
};

struct rzk_lightslice {
    rzk_ctx* ctx = nullptr;
    const rzk_scene* scene = nullptr;
    const rzk_vpls* vpls = nullptr;
    rzk_lightslice_params prm;
    int W = 0, H = 0, T = 0, tx = 0, ty = 0, S = 0;
    int64_t V = 0;
    bool initialized = false;
    rzk_lightslice_info info;
    // whole image
    DBuf<float> rays, recv, image;
    DBuf<float4> hits;
    DBuf<int> mat, light;
    // rows
    DBuf<int> row_pix, rows_mat;
    DBuf<uint64_t> vis_bits;          // per (representative, 8x8 pixel tile) blocked masks
    DBuf<float> rows_recv, R, Vm;
    // clustering
    DBuf<float> alphas, mags, norms, u, v, c0, center_w, K, Knorm, rep_w, mT;
    DBuf<unsigned long long> akey;                 // assignment keys (one per VPL)
    DBuf<int> centers, assign, reps;
    // slices
    DBuf<int> nb, upper;
    DBuf<float> slice_wi, slice_w, slice_pdf, slice_cdf, maps, col_cdf, row_pdf, row_cdf, totals;
    DBuf<float> pt_image;
    DBuf<unsigned long long> pt_work;
    int64_t h_pt_work[4] = {0, 0, 0, 0};   // last counted render_guided: closest, shadow rays, nodes, tris
    std::vector<int> h_centers, h_reps;
    std::vector<float> h_center_w, h_rep_w;
    PinnedArray<float> h_alphas, h_mags;           // device-to-host staging of the clustering inputs
    PinnedArray<int> h_assign;
    std::vector<int> h_cnt, h_off, h_mem, h_pos;   // cluster members (reused between initializations)
    DiscretePdf h_apdf;                            // the alphas' pdf
    // per-slice clusters (refinement): S x ld representatives / weights and counts
    DBuf<int> sreps, scount, cl_off, cl_mem;
    DBuf<float> srep_w, norms_R;
    std::vector<int> h_scount;
    int ld = 0;                        // 0: every slice gathers the kept clusters
    int cw = 0;                        // row length of the slice_* arrays
    // kd-tree slices: the gather points (pixels with a material, pixel order),
    // the tree, each slice's representative (kd position), each pixel's slice
    bool kd = false;
    std::unique_ptr<rzk_kdslices> kdt;
    DBuf<int> gp_pix, rep_pos, pix_slice;
    std::vector<int> h_gp_pix, h_row_pix, h_pix_slice;
    SliceReps slice_reps() const {
        const int* ps = kd ? pix_slice.p : nullptr;
        if (ld) return SliceReps{sreps.p, srep_w.p, scount.p, ld, ld, T, tx, ps};
        return SliceReps{reps.p, rep_w.p, nullptr, (int)h_reps.size(), 0, T, tx, ps};
    }
    int slice_of_pixel(int64_t i) const {
        return kd ? h_pix_slice[i] : (int)(((i / W) / T) * tx + (i % W) / T);
    }
};

static void ls_sync(rzk_ctx* ctx) { RZK_HIP(hipStreamSynchronize(ctx->stream)); }

template <typename F>
static int ls_guard(rzk_ctx* ctx, F&& f) {
    try {
        f();
        if (ctx) ctx->last_error.clear();
        return RZK_OK;
    } catch (const RzkError& e) {
        if (ctx) ctx->last_error = e.msg;
        return e.code;
    } catch (const std::exception& e) {
        if (ctx) ctx->last_error = e.what();
        return RZK_E_INVALID;
    }
}

// buffers sized by the slice count S
static void ls_alloc_slices(rzk_lightslice* ls) {
    const rzk_lightslice_params& p = ls->prm;
    const size_t S = (size_t)ls->S, V = (size_t)ls->V, C = (size_t)p.num_clusters, res = (size_t)p.map_resolution;
    const size_t CW = std::max<size_t>(C, (size_t)p.max_slice_clusters);
    ls->row_pix.alloc(S); ls->rows_mat.alloc(S); ls->rows_recv.alloc(RZK_RECV_FLOATS * S);
    ls->R.alloc(3 * S * V); ls->Vm.alloc(S * V); ls->u.alloc(S); ls->mT.alloc(S * V); ls->K.alloc(3 * S * C);
    ls->nb.alloc(S * p.num_neighbor_slices);
    ls->upper.alloc(S * CW); ls->slice_wi.alloc(3 * S * CW); ls->slice_w.alloc(S * CW);
    ls->slice_pdf.alloc(S * CW); ls->slice_cdf.alloc(S * CW);
    if (p.max_slice_clusters > 0) {
        ls->sreps.alloc(S * p.max_slice_clusters); ls->srep_w.alloc(S * p.max_slice_clusters); ls->scount.alloc(S);
    }
    ls->maps.alloc(S * res * res); ls->col_cdf.alloc(S * res * res);
    ls->row_pdf.alloc(S * res); ls->row_cdf.alloc(S * res); ls->totals.alloc(S);
}

// The reference's slices (generate_slices, lightslice.cc:252-303): gather
// points = pixels whose primary hit has a material, in pixel order; the 6-D
// kd-tree over their map_vec6 points; per slice a representative point drawn
// from the MT stream (start + rd() * size, then the colour's draw); rows =
// the representatives' pixels; neighbours = find_nearest_cluster of the
// representative (furthest first); each pixel's slice = find_cluster of its
// point.  Uses the primary hits already in ls->recv / ls->mat.
static void ls_kd_slices(rzk_lightslice* ls, Mt19937& rd) {
    rzk_ctx* ctx = ls->ctx;
    hipStream_t st = ctx->stream;
    const rzk_lightslice_params& p = ls->prm;
    const int64_t n = (int64_t)ls->W * ls->H;
    std::vector<int> hm((size_t)n);
    RZK_HIP(hipMemcpyAsync(hm.data(), ls->mat.p, sizeof(int) * n, hipMemcpyDeviceToHost, st));
    ls_sync(ctx);
    ls->h_gp_pix.clear();
    for (int64_t i = 0; i < n; ++i)
        if (hm[i] >= 0) ls->h_gp_pix.push_back((int)i);
    const int np = (int)ls->h_gp_pix.size();
    RZK_CHECK(np >= 2, RZK_E_STATE, "fewer than two gather points");
    RZK_HIP(hipMemcpyAsync(ls->gp_pix.p, ls->h_gp_pix.data(), sizeof(int) * np, hipMemcpyHostToDevice, st));
    // BoundingBox::diagonal of the scene box (boundingbox.h:109-111)
    const float* rb = ls->scene->bvh.root_box;
    const float dx = rb[3] - rb[0], dy = rb[4] - rb[1], dz = rb[5] - rb[2];
    const float diag = sqrtf(dx * dx + dy * dy + dz * dz);
    ls->kdt.reset(new rzk_kdslices());
    rzk_kdslices* kd = ls->kdt.get();
    kd->ctx = ctx;
    kd->n = np;
    kd->max_size = p.kd_max_slice_size;
    kd->seed = p.seed;
    kd->P.alloc((size_t)6 * np);
    SceneView sv = scene_view(ls->scene);
    kd_launch_map_vec6(ctx, sv.mat, ls->recv.p, ls->mat.p, ls->gp_pix.p, np, diag, kd->P.p);
    if (p.kd_pivots == 1) {
        // the reference's pivot stream: the tree's draws come first on rd
        std::vector<float> planes((size_t)6 * np);
        RZK_HIP(hipMemcpyAsync(planes.data(), kd->P.p, sizeof(float) * planes.size(), hipMemcpyDeviceToHost, st));
        ls_sync(ctx);
        (void)kd_build_host_mt(kd, planes.data(), rd);
    } else {
        kd_build_device(kd);
    }
    const int S = kd->S;
    RZK_CHECK(S >= p.num_neighbor_slices, RZK_E_STATE,
              "fewer kd-tree slices (" + std::to_string(S) + ") than neighbour slices");
    if (S != ls->S || !ls->row_pix.p) {
        ls->S = S;
        ls_alloc_slices(ls);
    }
    std::vector<int> order((size_t)np);
    RZK_HIP(hipMemcpyAsync(order.data(), kd->ord.p, sizeof(int) * np, hipMemcpyDeviceToHost, st));
    ls_sync(ctx);
    std::vector<int> rpos(S);
    ls->h_row_pix.resize(S);
    for (int s = 0; s < S; ++s) {
        const int a = kd->h_slices[2 * s], b = kd->h_slices[2 * s + 1];
        const float u = rd.next_float();
        int r = (int)((float)a + u * (float)(b - a + 1));
        r = std::min(r, np - 1);          // float rounding can give b + 1; the last slice stays in the array
        (void)rd.next_float();            // Rgb::from_hsv(rd() * 360, ...) colour
        rpos[s] = r;
        ls->h_row_pix[s] = ls->h_gp_pix[order[r]];
    }
    ls->rep_pos.alloc(S);
    RZK_HIP(hipMemcpyAsync(ls->rep_pos.p, rpos.data(), sizeof(int) * S, hipMemcpyHostToDevice, st));
    RZK_HIP(hipMemcpyAsync(ls->row_pix.p, ls->h_row_pix.data(), sizeof(int) * S, hipMemcpyHostToDevice, st));
    kd_launch_nearest_at(kd, ls->rep_pos.p, S, p.num_neighbor_slices, ls->nb.p);
    hipLaunchKernelGGL(k_fill_int, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ls->pix_slice.p, n, -1);
    kd_launch_find_points(kd, ls->gp_pix.p, ls->pix_slice.p);
    RZK_HIP(hipGetLastError());
    ls->h_pix_slice.resize((size_t)n);
    RZK_HIP(hipMemcpyAsync(ls->h_pix_slice.data(), ls->pix_slice.p, sizeof(int) * n, hipMemcpyDeviceToHost, st));
    ls_sync(ctx);
    ls->info.num_slices = S;
}

// rows = pixel-tile slice representatives (the tile centre)
static void ls_rows(rzk_lightslice* ls, std::vector<int>& pix) {
    pix.resize(ls->S);
    for (int y = 0; y < ls->ty; ++y)
        for (int x = 0; x < ls->tx; ++x) {
            int px = std::min(x * ls->T + ls->T / 2, ls->W - 1);
            int py = std::min(y * ls->T + ls->T / 2, ls->H - 1);
            pix[y * ls->tx + x] = py * ls->W + px;
        }
}

// the num_neighbor_slices nearest slices of every slice (tile-centre distance,
// ties by slice index), the slice itself first
static void ls_neighbours(rzk_lightslice* ls, std::vector<int>& nb, int k) {
    nb.assign((size_t)ls->S * k, 0);
    std::vector<std::pair<int64_t, int>> d(ls->S);
    for (int s = 0; s < ls->S; ++s) {
        int sx = s % ls->tx, sy = s / ls->tx;
        for (int o = 0; o < ls->S; ++o) {
            int ox = o % ls->tx, oy = o / ls->tx;
            d[o] = {(int64_t)(ox - sx) * (ox - sx) + (int64_t)(oy - sy) * (oy - sy), o};
        }
        std::partial_sort(d.begin(), d.begin() + k, d.end());
        for (int q = 0; q < k; ++q) nb[(size_t)s * k + q] = d[q].second;
    }
}

// ---------------------------------------------------------------------------
// Guides from resident clusters

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

GuideLayout guide_layout(int W, int H, int tile, int res, int C) {
    GuideLayout L;
    L.tx = (W + tile - 1) / tile;
    L.ty = (H + tile - 1) / tile;
    L.S = L.tx * L.ty;
    size_t S = (size_t)L.S, Cs = (size_t)std::max(C, 1), r = (size_t)res;
    size_t tmp[12] = {4 * S, 4 * S, 4 * 6 * S, 16 * S, 4 * RZK_RECV_FLOATS * S, 4 * S, 4 * S,
                      4 * 3 * S * Cs, 4 * S * Cs, 4 * 3 * S * Cs, 4 * S * Cs, 4 * S * r};
    size_t tab[4] = {4 * S * r * r, 4 * S * r * r, 4 * S * r, 4 * S};
    size_t o = 0;
    for (int i = 0; i < 12; ++i) { L.off_tmp[i] = o; o += align256(tmp[i]); }
    L.bytes_tmp = o;
    o = 0;
    for (int i = 0; i < 4; ++i) { L.off_tab[i] = o; o += align256(tab[i]); }
    L.bytes_tab = o;
    return L;
}

GuideBuffers guide_buffers(const GuideLayout& L, char* tmp, char* tab) {
    GuideBuffers b;
    b.px = (int*)(tmp + L.off_tmp[0]);
    b.py = (int*)(tmp + L.off_tmp[1]);
    b.rays = (float*)(tmp + L.off_tmp[2]);
    b.hits = (float4*)(tmp + L.off_tmp[3]);
    b.recv = (float*)(tmp + L.off_tmp[4]);
    b.mat = (int*)(tmp + L.off_tmp[5]);
    b.light = (int*)(tmp + L.off_tmp[6]);
    b.R = (float*)(tmp + L.off_tmp[7]);
    b.V = (float*)(tmp + L.off_tmp[8]);
    b.slice_wi = (float*)(tmp + L.off_tmp[9]);
    b.upper = (int*)(tmp + L.off_tmp[10]);
    b.row_pdf = (float*)(tmp + L.off_tmp[11]);
    b.maps = (float*)(tab + L.off_tab[0]);
    b.col_cdf = (float*)(tab + L.off_tab[1]);
    b.row_cdf = (float*)(tab + L.off_tab[2]);
    b.totals = (float*)(tab + L.off_tab[3]);
    return b;
}

// The per-slice stage of LightSlice::initialize for one camera, with the
// reduced matrix restricted to the resident representatives: the same
// (row, representative) values the full matrix holds, so the same maps.
void enqueue_guide_build(rzk_ctx* ctx, const rzk_scene* scene, const rzk_clusters& cl, const float* cam,
                         const rzk_guide_params& gp, const GuideLayout& L, const GuideBuffers& b) {
    SceneView sv = scene_view(scene);
    hipStream_t st = ctx->stream;
    int W = (int)cam[17], H = (int)cam[18], S = L.S, C = cl.C, res = gp.map_resolution;
    hipLaunchKernelGGL(k_tile_centres, dim3((S + 63) / 64), dim3(64), 0, st, L.tx, L.ty, gp.tile, W, H, b.px, b.py);
    launch_camera_rays(ctx, cam, S, b.px, b.py, W, b.rays);
    launch_trace_closest(ctx, sv, b.rays, S, 1e-3f, 1e6f, b.hits, nullptr);
    launch_fill_receivers(ctx, sv, b.rays, b.hits, S, b.recv, b.mat, b.light);
    launch_reduced_matrix(ctx, sv, b.recv, b.mat, S, cl.reps->d_order, cl.reps->d_vpl, cl.reps->d_mat, C,
                          gp.clamp_mode, gp.incident_clamp, gp.tmin, b.R, b.V, nullptr);
    int64_t sc = (int64_t)S * C;
    // the resident representatives are VPLs 0 .. C-1 of cl.reps, shared by every slice
    const SliceReps gsr{cl.d_ident, cl.d_rep_w, nullptr, C, 0, gp.tile, L.tx};
    hipLaunchKernelGGL(k_slice_info, dim3((unsigned)((sc + 255) / 256)), dim3(256), 0, st, sv.mat, b.recv, S, b.V,
                       cl.reps->d_vpl, cl.reps->d_mat, gsr, gp.clamp_mode, gp.incident_clamp, b.slice_wi, b.upper);
    size_t lds = sizeof(float) * ((size_t)res * res + res + 1 + 2 * kMapChunk);
    hipLaunchKernelGGL(k_slice_maps, dim3(S), dim3(64), lds, st, res, gsr, b.R, S, b.recv, b.slice_wi, b.upper,
                       gp.map_epsilon, gp.map_gate ? 0 : 1, b.maps, b.col_cdf, b.row_pdf, b.row_cdf, b.totals);
    RZK_HIP(hipGetLastError());
}

void guide_set(rzk_guide& g, const rzk_scene* scene, const float* cam, const rzk_guide_params& gp,
               const GuideLayout& L, const GuideBuffers& b) {
    g.scene = scene;
    std::memcpy(g.cam, cam, sizeof(g.cam));
    g.W = (int)cam[17]; g.H = (int)cam[18];
    g.tile = gp.tile; g.tiles_x = L.tx; g.tiles_y = L.ty; g.res = gp.map_resolution;
    g.gv.maps = b.maps; g.gv.row_cdf = b.row_cdf; g.gv.col_cdf = b.col_cdf; g.gv.totals = b.totals;
    g.device_bytes = (int64_t)L.bytes_tab;
}

static void check_guide_params(const rzk_guide_params& gp, const float* cam) {
    RZK_CHECK(gp.tile > 0 && gp.map_resolution > 0 && gp.map_resolution <= 64, RZK_E_INVALID,
              "bad guide parameters (tile > 0, 0 < map_resolution <= 64)");
    RZK_CHECK((int)cam[17] > 0 && (int)cam[18] > 0, RZK_E_INVALID, "camera without an image size");
}

extern "C" {

int rzk_lightslice_clusters(const rzk_lightslice* ls, rzk_clusters** out) {
    if (!ls || !out) return RZK_E_INVALID;
    *out = nullptr;
    rzk_ctx* ctx = ls->ctx;
    return ls_guard(ctx, [&] {
        RZK_CHECK(ls->initialized && !ls->h_reps.empty(), RZK_E_STATE, "LightSlice has no clusters");
        std::unique_ptr<rzk_clusters> cl(new rzk_clusters());
        cl->ctx = ctx;
        cl->scene = ls->scene;
        cl->C = (int)ls->h_reps.size();
        cl->h_reps = ls->h_reps;
        cl->h_rep_w = ls->h_rep_w;
        std::vector<float> vr((size_t)cl->C * RZK_VPL_FLOATS);
        std::vector<int> mr(cl->C), ident(cl->C);
        for (int c = 0; c < cl->C; ++c) {
            int k = ls->h_reps[c];
            std::memcpy(&vr[(size_t)c * RZK_VPL_FLOATS], &ls->vpls->h_vpl[(size_t)k * RZK_VPL_FLOATS],
                        sizeof(float) * RZK_VPL_FLOATS);
            mr[c] = ls->vpls->h_mat[k];
            ident[c] = c;
        }
        int rc = rzk_vpls_create(ctx, vr.data(), mr.data(), cl->C, &cl->reps);
        if (rc != RZK_OK) throw RzkError{rc, ctx->last_error};
        RZK_HIP(hipMalloc((void**)&cl->d_rep_w, sizeof(float) * cl->C));
        RZK_HIP(hipMalloc((void**)&cl->d_ident, sizeof(int) * cl->C));
        RZK_HIP(hipMemcpy(cl->d_rep_w, cl->h_rep_w.data(), sizeof(float) * cl->C, hipMemcpyHostToDevice));
        RZK_HIP(hipMemcpy(cl->d_ident, ident.data(), sizeof(int) * cl->C, hipMemcpyHostToDevice));
        const rzk_lightslice_params& p = ls->prm;
        cl->n_emitted = p.n_emitted;
        cl->gp = rzk_guide_params{p.tile, p.map_resolution, p.map_epsilon, p.clamp_mode, p.incident_clamp, p.tmin,
                                  p.map_gate};
        ctx_retain(ctx);
        *out = cl.release();
    });
}

int rzk_clusters_destroy(rzk_clusters* cl) {
    if (!cl) return RZK_OK;
    (void)hipDeviceSynchronize();
    rzk_vpls_destroy(cl->reps);
    if (cl->d_rep_w) (void)hipFree(cl->d_rep_w);
    if (cl->d_ident) (void)hipFree(cl->d_ident);
    rzk_ctx* ctx = cl->ctx;
    delete cl;
    ctx_release(ctx);
    return RZK_OK;
}

int rzk_guide_create(rzk_ctx* ctx, const rzk_scene* scene, const rzk_clusters* cl, const float* camera,
                     const rzk_guide_params* params, rzk_guide** out) {
    if (!ctx || !scene || !cl || !camera || !out) return RZK_E_INVALID;
    *out = nullptr;
    return ls_guard(ctx, [&] {
        RZK_CHECK(cl->scene == scene, RZK_E_INVALID, "clusters belong to another scene");
        rzk_guide_params gp = params ? *params : cl->gp;
        check_guide_params(gp, camera);
        GuideLayout L = guide_layout((int)camera[17], (int)camera[18], gp.tile, gp.map_resolution, cl->C);
        char *tmp = nullptr, *tab = nullptr;
        RZK_HIP(hipMalloc((void**)&tmp, L.bytes_tmp));
        hipError_t e = hipMalloc((void**)&tab, L.bytes_tab);
        if (e != hipSuccess) { (void)hipFree(tmp); RZK_HIP(e); }
        std::unique_ptr<rzk_guide> g(new rzk_guide());
        g->ctx = ctx;
        g->owned.push_back(tab);
        GuideBuffers b = guide_buffers(L, tmp, tab);
        try {
            enqueue_guide_build(ctx, scene, *cl, camera, gp, L, b);
            RZK_HIP(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            (void)hipFree(tmp);
            throw;
        }
        RZK_HIP(hipFree(tmp));
        guide_set(*g, scene, camera, gp, L, b);
        ctx_retain(ctx);
        *out = g.release();
    });
}

int rzk_guide_get(const rzk_guide* g, const char* name, void* dst, int64_t bytes) {
    if (!g || !name || !dst) return RZK_E_INVALID;
    rzk_ctx* ctx = g->ctx;
    return ls_guard(ctx, [&] {
        std::string nm(name);
        int64_t S = (int64_t)g->tiles_x * g->tiles_y, r = g->res;
        const void* src = nullptr;
        int64_t need = 0;
        if (nm == "maps") { src = g->gv.maps; need = 4 * S * r * r; }
        else if (nm == "col_cdf") { src = g->gv.col_cdf; need = 4 * S * r * r; }
        else if (nm == "row_cdf") { src = g->gv.row_cdf; need = 4 * S * r; }
        else if (nm == "totals") { src = g->gv.totals; need = 4 * S; }
        RZK_CHECK(src != nullptr, RZK_E_INVALID, "unknown guide array: " + nm);
        RZK_CHECK(bytes == need, RZK_E_INVALID, "size mismatch for " + nm);
        RZK_HIP(hipMemcpyAsync(dst, src, (size_t)need, hipMemcpyDeviceToHost, ctx->stream));
        ls_sync(ctx);
    });
}

int rzk_render_guided(rzk_ctx* ctx, const rzk_guide* g, const rzk_pt_params* p, float* out_rgb, int64_t* out_rays) {
    if (!ctx || !g || !p) return RZK_E_INVALID;
    return ls_guard(ctx, [&] {
        rzk_render_job j{g, p->w > 0 ? p->x0 : 0, p->h > 0 ? p->y0 : 0, p->w > 0 ? p->w : g->W, p->h > 0 ? p->h : g->H,
                         p->spp, p->max_bounce, p->direct_lighting, p->guide_mode, p->seed, p->tmin};
        check_job(g, j);
        PtParams pp = job_params(*g, j);
        int64_t n = (int64_t)g->W * g->H;
        DBuf<float> img;
        img.alloc(3 * n);
        DBuf<unsigned long long> work;
        work.alloc(4);
        hipStream_t st = ctx->stream;
        if ((int64_t)pp.w * pp.h != n) RZK_HIP(hipMemsetAsync(img.p, 0, sizeof(float) * 3 * n, st));
        if (out_rays) RZK_HIP(hipMemsetAsync(work.p, 0, 4 * sizeof(unsigned long long), st));
        launch_guided_pt(ctx, scene_view(g->scene), g->cam, pp, g->gv, img.p, out_rays ? work.p : nullptr);
        if (out_rgb) RZK_HIP(hipMemcpyAsync(out_rgb, img.p, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, st));
        unsigned long long h[2] = {0, 0};
        if (out_rays) RZK_HIP(hipMemcpyAsync(h, work.p, sizeof(h), hipMemcpyDeviceToHost, st));
        ls_sync(ctx);
        if (out_rays) { out_rays[0] = (int64_t)h[0]; out_rays[1] = (int64_t)h[1]; }
    });
}

int rzk_render_guided_rays(rzk_ctx* ctx, const rzk_guide* g, const rzk_pt_params* p, const float* rays,
                           const int32_t* pix, const int32_t* smp, int64_t n, float* out_rgb) {
    if (!ctx || !g || !p) return RZK_E_INVALID;
    return ls_guard(ctx, [&] {
        RZK_CHECK(n >= 0 && (n == 0 || (rays && pix && smp && out_rgb)), RZK_E_INVALID, "bad ray arguments");
        RZK_CHECK(p->max_bounce >= 0 && (p->guide_mode == 0 || p->guide_mode == 1), RZK_E_INVALID,
                  "bad path tracing parameters");
        RZK_CHECK(g->tile > 0 && g->tiles_x > 0 && g->tiles_y > 0, RZK_E_INVALID, "guide without slices");
        for (int64_t i = 0; i < n; ++i)
            RZK_CHECK(pix[2 * i] >= 0 && pix[2 * i] < g->W && pix[2 * i + 1] >= 0 && pix[2 * i + 1] < g->H &&
                          smp[i] >= 0,
                      RZK_E_INVALID, "ray pixel outside the guide's image");
        if (n == 0) return;
        rzk_render_job j{g, 0, 0, g->W, g->H, 1, p->max_bounce, p->direct_lighting, p->guide_mode, p->seed, p->tmin};
        PtParams pp = job_params(*g, j);
        DBuf<float> d_rays, d_out;
        DBuf<int> d_pix, d_smp;
        d_rays.alloc(6 * n); d_out.alloc(3 * n); d_pix.alloc(2 * n); d_smp.alloc(n);
        hipStream_t st = ctx->stream;
        RZK_HIP(hipMemcpyAsync(d_rays.p, rays, sizeof(float) * 6 * n, hipMemcpyHostToDevice, st));
        RZK_HIP(hipMemcpyAsync(d_pix.p, pix, sizeof(int) * 2 * n, hipMemcpyHostToDevice, st));
        RZK_HIP(hipMemcpyAsync(d_smp.p, smp, sizeof(int) * n, hipMemcpyHostToDevice, st));
        launch_guided_pt_rays(ctx, scene_view(g->scene), pp, g->gv, d_rays.p, d_pix.p, d_smp.p, n, d_out.p);
        RZK_HIP(hipMemcpyAsync(out_rgb, d_out.p, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, st));
        ls_sync(ctx);
    });
}

int rzk_render_vpl(rzk_ctx* ctx, const rzk_scene* s, const rzk_vpls* v, const float* camera, int clamp_mode,
                   float clamp, float tmin, float n_emitted, float* out_rgb) {
    return ls_guard(ctx, [&] {
        RZK_CHECK(s && v && camera, RZK_E_INVALID, "null argument");
        RZK_CHECK(s->bvh.ntri > 0, RZK_E_INVALID, "empty scene");
        int W = (int)camera[17], H = (int)camera[18];
        RZK_CHECK(W > 0 && H > 0, RZK_E_INVALID, "bad image size");
        int64_t n = (int64_t)W * H;
        DBuf<float> rays, recv, img;
        DBuf<float4> hits;
        DBuf<int> mat, light, order;
        rays.alloc(6 * n); recv.alloc(RZK_RECV_FLOATS * n); img.alloc(3 * n);
        hits.alloc(n); mat.alloc(n); light.alloc(n); order.alloc(n);
        // 8x8 pixel tiles per wave keep the 64 segments of a step coherent
        std::vector<int> ord(n);
        int64_t t = 0;
        for (int by = 0; by < H; by += 8)
            for (int bx = 0; bx < W; bx += 8)
                for (int yy = by; yy < std::min(by + 8, H); ++yy)
                    for (int xx = bx; xx < std::min(bx + 8, W); ++xx) ord[t++] = yy * W + xx;
        RZK_HIP(hipMemcpyAsync(order.p, ord.data(), sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
        SceneView sv = scene_view(s);
        launch_camera_image_rays(ctx, camera, W, H, rays.p);
        launch_trace_closest_image(ctx, sv, rays.p, W, H, 1e-3f, 1e6f, hits.p, nullptr);
        launch_fill_receivers(ctx, sv, rays.p, hits.p, n, recv.p, mat.p, light.p);
        // the gather as visibility bits then the shading sum, VPLs in chunks whose
        // bit matrix stays within 256 MiB; the running sums carry across chunks in
        // VPL order, so the image is that of one launch over all VPLs
        const int64_t ngroups = (n + 63) / 64;
        const int64_t budget_words = (int64_t)32 << 20;
        int64_t chunk = std::max<int64_t>(64, budget_words / ngroups / 64 * 64);
        if (v->n > 0) chunk = std::min(chunk, v->n);
        DBuf<uint64_t> bits;
        bits.alloc(ngroups * chunk);
        DBuf<float> acc;
        if (v->n > chunk) acc.alloc(3 * n);
        if (v->n == 0)
            launch_gather_shade(ctx, sv, recv.p, mat.p, n, order.p, v->d_vpl, v->d_mat, 0, bits.p, clamp_mode, clamp,
                                n_emitted, nullptr, nullptr, img.p);
        for (int64_t k0 = 0; k0 < v->n; k0 += chunk) {
            const int64_t cnt = std::min(chunk, v->n - k0);
            const bool first = k0 == 0, last = k0 + cnt >= v->n;
            launch_visibility(ctx, sv, recv.p, mat.p, n, order.p, v->d_vpl + (int64_t)RZK_VPL_FLOATS * k0, cnt, tmin,
                              bits.p, nullptr);
            launch_gather_shade(ctx, sv, recv.p, mat.p, n, order.p, v->d_vpl + (int64_t)RZK_VPL_FLOATS * k0,
                                v->d_mat + k0, cnt, bits.p, clamp_mode, clamp, n_emitted, first ? nullptr : acc.p,
                                last ? nullptr : acc.p, img.p);
        }
        if (out_rgb) RZK_HIP(hipMemcpyAsync(out_rgb, img.p, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, ctx->stream));
        ls_sync(ctx);
    });
}

int rzk_refine_slices(rzk_ctx* ctx, int32_t S, int64_t V, const float* R, int32_t nnb, const int32_t* nb,
                      int32_t ncl, const int32_t* cl_off, const int32_t* cl_mem, const int32_t* cl_rep,
                      const float* cl_w, int32_t max_slice_clusters, const float* draws, int64_t ndraws,
                      int32_t* out_count, int32_t* out_rep, float* out_w) {
    return ls_guard(ctx, [&] {
        RZK_CHECK(ctx && R && nb && cl_off && cl_mem && cl_rep && cl_w && out_count && out_rep && out_w,
                  RZK_E_INVALID, "null argument");
        RZK_CHECK(S > 0 && V > 0 && V < INT32_MAX && ncl > 0 && max_slice_clusters >= ncl, RZK_E_INVALID,
                  "bad refinement sizes");
        RZK_CHECK(ndraws >= refine_draws_needed(S, ncl, max_slice_clusters) && (ndraws == 0 || draws), RZK_E_INVALID,
                  "too few random numbers for the refinement");
        for (int s = 0; s < S * nnb; ++s) RZK_CHECK(nb[s] >= 0 && nb[s] < S, RZK_E_INVALID, "bad neighbour row");
        RZK_CHECK(cl_off[0] == 0, RZK_E_INVALID, "bad cluster offsets");
        for (int c = 0; c < ncl; ++c) RZK_CHECK(cl_off[c + 1] > cl_off[c], RZK_E_INVALID, "empty cluster");
        const int nmem = cl_off[ncl];
        RZK_CHECK(nmem <= V, RZK_E_INVALID, "more members than VPLs");
        for (int e = 0; e < nmem; ++e) RZK_CHECK(cl_mem[e] >= 0 && cl_mem[e] < V, RZK_E_INVALID, "bad member");
        hipStream_t st = ctx->stream;
        DBuf<float> dR, mags, norms, u, v, c0, alphas, mT, normsR, w;
        DBuf<int> dnb, doff, dmem, drep, rep, count;
        dR.alloc(3 * (size_t)S * V); dnb.alloc((size_t)S * nnb); doff.alloc(ncl + 1); dmem.alloc(nmem);
        drep.alloc(ncl); w.alloc(ncl);
        mags.alloc(V); norms.alloc(V); u.alloc(S); v.alloc(V); c0.alloc(1); alphas.alloc(V); mT.alloc((size_t)S * V);
        normsR.alloc(V);
        rep.alloc((size_t)S * max_slice_clusters); count.alloc(S);
        DBuf<float> rw;
        rw.alloc((size_t)S * max_slice_clusters);
        RZK_HIP(hipMemcpyAsync(dR.p, R, sizeof(float) * 3 * (size_t)S * V, hipMemcpyHostToDevice, st));
        RZK_HIP(hipMemcpyAsync(dnb.p, nb, sizeof(int) * (size_t)S * nnb, hipMemcpyHostToDevice, st));
        RZK_HIP(hipMemcpyAsync(doff.p, cl_off, sizeof(int) * (ncl + 1), hipMemcpyHostToDevice, st));
        RZK_HIP(hipMemcpyAsync(dmem.p, cl_mem, sizeof(int) * nmem, hipMemcpyHostToDevice, st));
        RZK_HIP(hipMemcpyAsync(drep.p, cl_rep, sizeof(int) * ncl, hipMemcpyHostToDevice, st));
        RZK_HIP(hipMemcpyAsync(w.p, cl_w, sizeof(float) * ncl, hipMemcpyHostToDevice, st));
        // flattened column magnitudes (the kept clusters' pdfs) and column norms
        launch_alphas_full(ctx, S, V, dR.p, alphas.p, mags.p, norms.p, u.p, v.p, c0.p, mT.p);
        launch_column_norms_R(ctx, S, V, dR.p, normsR.p);
        std::vector<float> h(draws, draws + ndraws);
        launch_refine_slices(ctx, dR.p, S, dnb.p, nnb, ncl, doff.p, dmem.p, nmem, drep.p, w.p, mags.p, normsR.p,
                             max_slice_clusters, h, rep.p, rw.p, count.p);
        RZK_HIP(hipMemcpyAsync(out_count, count.p, sizeof(int) * S, hipMemcpyDeviceToHost, st));
        RZK_HIP(hipMemcpyAsync(out_rep, rep.p, sizeof(int) * (size_t)S * max_slice_clusters, hipMemcpyDeviceToHost, st));
        RZK_HIP(hipMemcpyAsync(out_w, rw.p, sizeof(float) * (size_t)S * max_slice_clusters, hipMemcpyDeviceToHost, st));
        ls_sync(ctx);
    });
}

int rzk_lightslice_create(rzk_ctx* ctx, const rzk_scene* s, const rzk_vpls* v, const rzk_lightslice_params* p,
                          rzk_lightslice** out) {
    return ls_guard(ctx, [&] {
        RZK_CHECK(ctx && s && v && p && out, RZK_E_INVALID, "null argument");
        *out = nullptr;
        RZK_CHECK(p->kd_max_slice_size == 0 || p->kd_max_slice_size >= 3, RZK_E_INVALID,
                  "kd_max_slice_size must be 0 (pixel tiles) or at least 3");
        RZK_CHECK(p->kd_pivots == 0 || p->kd_pivots == 1, RZK_E_INVALID,
                  "kd_pivots must be 0 (per-node Philox) or 1 (one MT stream in recursion order)");
        RZK_CHECK(p->width > 0 && p->height > 0 && (p->tile > 0 || p->kd_max_slice_size > 0), RZK_E_INVALID,
                  "bad image or tile size");
        RZK_CHECK(p->num_clusters > 0 && p->num_neighbor_slices > 0 && p->map_resolution > 0, RZK_E_INVALID,
                  "bad clustering parameters");
        RZK_CHECK(v->n > 0, RZK_E_INVALID, "There are no VPLs to cluster");
        RZK_CHECK(v->n >= p->num_clusters, RZK_E_INVALID, "fewer VPLs than clusters (cluster_builder.cc:91-101)");
        RZK_CHECK(s->bvh.ntri > 0, RZK_E_INVALID, "empty scene");
        std::unique_ptr<rzk_lightslice> ls(new rzk_lightslice());
        ls->ctx = ctx; ls->scene = s; ls->vpls = v; ls->prm = *p;
        ls->W = p->width; ls->H = p->height;
        ls->kd = p->kd_max_slice_size > 0;
        ls->T = ls->kd ? std::max(1, p->tile) : p->tile;
        ls->tx = (ls->W + ls->T - 1) / ls->T; ls->ty = (ls->H + ls->T - 1) / ls->T;
        ls->S = ls->kd ? 0 : ls->tx * ls->ty;           // kd-tree slices: known after initialize
        ls->V = v->n;
        RZK_CHECK(ls->kd || p->num_neighbor_slices <= ls->S, RZK_E_INVALID, "more neighbour slices than slices");
        RZK_CHECK(p->num_neighbor_slices <= 32 || !ls->kd, RZK_E_INVALID,
                  "at most 32 neighbour slices with kd-tree slices");
        int64_t n = (int64_t)ls->W * ls->H;
        int S = ls->S, C = p->num_clusters;
        int64_t V = ls->V;
        RZK_CHECK(p->max_slice_clusters >= 0 && (int64_t)p->max_slice_clusters <= V, RZK_E_INVALID,
                  "max_slice_clusters must lie in [0, number of VPLs]");
        const int CW = std::max(C, p->max_slice_clusters);      // widest per-slice cluster list
        ls->rays.alloc(6 * n); ls->recv.alloc(RZK_RECV_FLOATS * n); ls->image.alloc(3 * n);
        ls->hits.alloc(n); ls->mat.alloc(n); ls->light.alloc(n);
        ls->vis_bits.alloc((size_t)CW * ((ls->W + 7) / 8) * ((ls->H + 7) / 8));
        ls->alphas.alloc(V); ls->mags.alloc(V); ls->norms.alloc(V); ls->v.alloc(V); ls->c0.alloc(1);
        ls->centers.alloc(C); ls->center_w.alloc(C); ls->Knorm.alloc(C);
        ls->assign.alloc(V); ls->reps.alloc(C); ls->rep_w.alloc(C); ls->akey.alloc(V);
        if (p->max_slice_clusters > 0) { ls->norms_R.alloc(V); ls->cl_off.alloc(C + 1); ls->cl_mem.alloc(V); }
        std::memset(&ls->info, 0, sizeof(ls->info));
        if (!ls->kd) {
            ls_alloc_slices(ls.get());
            std::vector<int> pix, nbh;
            ls_rows(ls.get(), pix);
            ls_neighbours(ls.get(), nbh, p->num_neighbor_slices);
            RZK_HIP(hipMemcpy(ls->row_pix.p, pix.data(), sizeof(int) * S, hipMemcpyHostToDevice));
            RZK_HIP(hipMemcpy(ls->nb.p, nbh.data(), sizeof(int) * nbh.size(), hipMemcpyHostToDevice));
            ls->h_row_pix = pix;
        } else {
            ls->pix_slice.alloc(n);
            ls->gp_pix.alloc(n);
        }
        ls->info.num_slices = S;
        ctx_retain(ctx);
        *out = ls.release();
    });
}

int rzk_lightslice_destroy(rzk_lightslice* ls) {
    if (!ls) return RZK_OK;
    rzk_ctx* ctx = ls->ctx;
    delete ls;
    ctx_release(ctx);
    return RZK_OK;
}

int rzk_lightslice_initialize(rzk_lightslice* ls, const float* camera) {
    if (!ls) return RZK_E_INVALID;
    rzk_ctx* ctx = ls->ctx;
    return ls_guard(ctx, [&] {
        RZK_CHECK(camera, RZK_E_INVALID, "null camera");
        RZK_CHECK((int)camera[17] == ls->W && (int)camera[18] == ls->H, RZK_E_INVALID,
                  "camera image size differs from the LightSlice parameters");
        const rzk_lightslice_params& p = ls->prm;
        SceneView sv = scene_view(ls->scene);
        int64_t V = ls->V, n = (int64_t)ls->W * ls->H;
        hipStream_t st = ctx->stream;
        // one MT stream for the slice representatives, the centres and the refinement
        Mt19937 rd(p.seed);
        ls->initialized = false;
        // primary hits for every pixel (camera.cc:52-65 pixel centres)
        launch_camera_image_rays(ctx, camera, ls->W, ls->H, ls->rays.p);
        launch_trace_closest_image(ctx, sv, ls->rays.p, ls->W, ls->H, 1e-3f, 1e6f, ls->hits.p, nullptr);
        launch_fill_receivers(ctx, sv, ls->rays.p, ls->hits.p, n, ls->recv.p, ls->mat.p, ls->light.p);
        if (ls->kd) ls_kd_slices(ls, rd);
        int S = ls->S;
        hipLaunchKernelGGL(k_pick_rows, dim3((S + 63) / 64), dim3(64), 0, st, ls->recv.p, ls->mat.p, ls->row_pix.p, S,
                           ls->rows_recv.p, ls->rows_mat.p);
        // reduced matrix (lightslice.cc:721-748)
        launch_reduced_matrix(ctx, sv, ls->rows_recv.p, ls->rows_mat.p, S, ls->vpls->d_order, ls->vpls->d_vpl, ls->vpls->d_mat, V,
                              p.clamp_mode, p.incident_clamp, p.tmin, ls->R.p, ls->Vm.p, nullptr);
        // alphas (light_transport_matrix.cc:72-106)
        launch_alphas_full(ctx, S, V, ls->R.p, ls->alphas.p, ls->mags.p, ls->norms.p, ls->u.p, ls->v.p, ls->c0.p,
                           ls->mT.p);
        ls->h_alphas.reserve(V);
        RZK_HIP(hipMemcpyAsync(ls->h_alphas.p, ls->alphas.p, sizeof(float) * V, hipMemcpyDeviceToHost, st));
        ls_sync(ctx);
        // centre sampling (cluster_builder.cc:114-156) on the host MT stream
        DiscretePdf& apdf = ls->h_apdf;
        apdf.build(ls->h_alphas.p, (size_t)V);
        int C = p.num_clusters;
        ls->h_centers.clear();
        ls->h_center_w.clear();
        int64_t draws = 0, max_draws = 1000LL * C;
        while ((int)ls->h_centers.size() < C && draws < max_draws) {
            draws++;
            float pk;
            int k = apdf.sample(rd.next_float(), pk);
            if (pk == 0.0f) continue;
            auto it = std::find(ls->h_centers.begin(), ls->h_centers.end(), k);
            if (it == ls->h_centers.end()) {
                ls->h_centers.push_back(k);
                ls->h_center_w.push_back(1.0f / pk);
            } else {
                ls->h_center_w[it - ls->h_centers.begin()] += 1.0f / pk;
            }
        }
        int ncl = (int)ls->h_centers.size();
        RZK_CHECK(ncl > 0, RZK_E_STATE, "no cluster centre could be sampled (all alphas zero)");
        RZK_HIP(hipMemcpyAsync(ls->centers.p, ls->h_centers.data(), sizeof(int) * ncl, hipMemcpyHostToDevice, st));
        RZK_HIP(hipMemcpyAsync(ls->center_w.p, ls->h_center_w.data(), sizeof(float) * ncl, hipMemcpyHostToDevice, st));
        launch_assign_full(ctx, S, V, ls->R.p, ncl, ls->centers.p, ls->center_w.p, ls->mags.p, ls->K.p, ls->Knorm.p,
                           ls->akey.p, ls->assign.p);
        ls->h_assign.reserve(V);
        ls->h_mags.reserve(V);
        RZK_HIP(hipMemcpyAsync(ls->h_assign.p, ls->assign.p, sizeof(int) * V, hipMemcpyDeviceToHost, st));
        RZK_HIP(hipMemcpyAsync(ls->h_mags.p, ls->mags.p, sizeof(float) * V, hipMemcpyDeviceToHost, st));
        ls_sync(ctx);
        // members in increasing VPL order, representatives (cluster_builder.cc:230-272)
        std::vector<int>& cnt = ls->h_cnt;
        std::vector<int>& off = ls->h_off;
        std::vector<int>& mem = ls->h_mem;
        cnt.assign(ncl + 1, 0);
        off.assign(ncl + 1, 0);
        mem.resize(V);
        for (int64_t i = 0; i < V; ++i)
            if (ls->h_assign.p[i] >= 0) cnt[ls->h_assign.p[i]]++;
        for (int c = 0; c < ncl; ++c) off[c + 1] = off[c] + cnt[c];
        std::vector<int>& pos = ls->h_pos;
        pos.assign(off.begin(), off.end());
        for (int64_t i = 0; i < V; ++i) {
            int c = ls->h_assign.p[i];
            if (c >= 0) mem[pos[c]++] = (int)i;
        }
        ls->h_reps.clear();
        ls->h_rep_w.clear();
        for (int c = 0; c < ncl; ++c) {
            const int m = off[c + 1] - off[c];
            const int* mc = mem.data() + off[c];
            float u = rd.next_float();
            float pj = 0.f;
            const int j = sample_members(ls->h_mags.p, mc, m, u, pj);
            int rep;
            float w;
            if (j < 0) { rep = m > 0 ? mc[m / 2] : -1; w = 0.0f; }
            else { rep = mc[j]; w = 1.0f / pj; }
            if (rep >= 0) { ls->h_reps.push_back(rep); ls->h_rep_w.push_back(w); }
        }
        int nreps = (int)ls->h_reps.size();
        RZK_HIP(hipMemcpyAsync(ls->reps.p, ls->h_reps.data(), sizeof(int) * nreps, hipMemcpyHostToDevice, st));
        RZK_HIP(hipMemcpyAsync(ls->rep_w.p, ls->h_rep_w.data(), sizeof(float) * nreps, hipMemcpyHostToDevice, st));
        // per-slice refinement (lightslice.cc:837-902) of the kept clusters, the
        // MT stream continuing after the representatives
        ls->ld = 0;
        ls->info.refine_passes = 0;
        if (p.max_slice_clusters > nreps && nreps > 0) {
            std::vector<int> koff(1, 0), kmem;
            kmem.reserve(V);
            for (int c = 0; c < ncl; ++c) {
                if (cnt[c] == 0) continue;                 // dropped: no representative
                kmem.insert(kmem.end(), mem.begin() + off[c], mem.begin() + off[c] + cnt[c]);
                koff.push_back((int)kmem.size());
            }
            RZK_HIP(hipMemcpyAsync(ls->cl_off.p, koff.data(), sizeof(int) * koff.size(), hipMemcpyHostToDevice, st));
            RZK_HIP(hipMemcpyAsync(ls->cl_mem.p, kmem.data(), sizeof(int) * kmem.size(), hipMemcpyHostToDevice, st));
            launch_column_norms_R(ctx, S, V, ls->R.p, ls->norms_R.p);
            std::vector<float> draws((size_t)refine_draws_needed(S, nreps, p.max_slice_clusters));
            for (auto& x : draws) x = rd.next_float();
            ls->info.refine_passes =
                launch_refine_slices(ctx, ls->R.p, S, ls->nb.p, p.num_neighbor_slices, nreps, ls->cl_off.p,
                                     ls->cl_mem.p, (int)kmem.size(), ls->reps.p, ls->rep_w.p, ls->mags.p,
                                     ls->norms_R.p, p.max_slice_clusters, draws, ls->sreps.p, ls->srep_w.p,
                                     ls->scount.p);
            ls->ld = p.max_slice_clusters;
            ls->h_scount.resize(S);
            RZK_HIP(hipMemcpyAsync(ls->h_scount.data(), ls->scount.p, sizeof(int) * S, hipMemcpyDeviceToHost, st));
        } else {
            ls->h_scount.assign(S, nreps);
        }
        const SliceReps sr = ls->slice_reps();
        ls->cw = sr.width();
        // per-slice cluster data, pdfs and luminance maps
        int64_t sc = (int64_t)S * ls->cw;
        if (nreps > 0) {
            hipLaunchKernelGGL(k_slice_info, dim3((unsigned)((sc + 255) / 256)), dim3(256), 0, st, sv.mat, ls->rows_recv.p,
                               S, ls->Vm.p, ls->vpls->d_vpl, ls->vpls->d_mat, sr, p.clamp_mode, p.incident_clamp,
                               ls->slice_wi.p, ls->upper.p);
            hipLaunchKernelGGL(k_slice_pdf_weights, dim3((unsigned)((sc + 255) / 256)), dim3(256), 0, st, ls->R.p, S,
                               ls->nb.p, p.num_neighbor_slices, sr, ls->upper.p, ls->slice_w.p);
            hipLaunchKernelGGL(k_slice_pdf_norm, dim3(S), dim3(64), 0, st, S, sr, ls->slice_w.p,
                               ls->slice_pdf.p, ls->slice_cdf.p);
        }
        int res = p.map_resolution;
        size_t lds = sizeof(float) * ((size_t)res * res + res + 1 + 2 * kMapChunk);
        hipLaunchKernelGGL(k_slice_maps, dim3(S), dim3(64), lds, st, res, sr, ls->R.p, S, ls->rows_recv.p,
                           ls->slice_wi.p, ls->upper.p, p.map_epsilon, p.map_gate ? 0 : 1,
                           ls->maps.p, ls->col_cdf.p, ls->row_pdf.p, ls->row_cdf.p, ls->totals.p);
        RZK_HIP(hipGetLastError());
        ls_sync(ctx);
        ls->info.num_centers = ncl;
        ls->info.num_clusters = nreps;
        ls->info.draws = (int32_t)draws;
        ls->info.reduced_pairs = (int64_t)S * V;
        ls->initialized = true;
    });
}

int rzk_lightslice_gather(rzk_lightslice* ls, float* out_rgb) {
    if (!ls) return RZK_E_INVALID;
    rzk_ctx* ctx = ls->ctx;
    return ls_guard(ctx, [&] {
        RZK_CHECK(ls->initialized, RZK_E_STATE, "rzk_lightslice_initialize has not run");
        const rzk_lightslice_params& p = ls->prm;
        SceneView sv = scene_view(ls->scene);
        launch_slice_gather(ctx, sv, ls->recv.p, ls->mat.p, ls->W, ls->H, ls->vpls->d_vpl, ls->vpls->d_mat,
                            ls->slice_reps(), p.clamp_mode, p.clamp, p.tmin, p.n_emitted, ls->image.p,
                            ls->vis_bits.p, nullptr);
        int64_t n = (int64_t)ls->W * ls->H;
        if (out_rgb) {
            RZK_HIP(hipMemcpyAsync(out_rgb, ls->image.p, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, ctx->stream));
            ls_sync(ctx);
        }
    });
}

int rzk_lightslice_render_guided(rzk_lightslice* ls, const float* camera, const rzk_pt_params* p,
                                 float* out_rgb, int64_t* out_rays) {
    if (!ls || !camera || !p) return RZK_E_INVALID;
    rzk_ctx* ctx = ls->ctx;
    return ls_guard(ctx, [&] {
        RZK_CHECK(ls->initialized, RZK_E_STATE, "rzk_lightslice_initialize has not run");
        RZK_CHECK((int)camera[17] == ls->W && (int)camera[18] == ls->H, RZK_E_INVALID, "camera size differs");
        RZK_CHECK(p->spp > 0 && p->max_bounce >= 0 && (p->guide_mode == 0 || p->guide_mode == 1), RZK_E_INVALID,
                  "bad path tracing parameters");
        PtParams pp;
        pp.width = ls->W; pp.height = ls->H; pp.spp = p->spp; pp.max_bounce = p->max_bounce;
        pp.direct = p->direct_lighting ? 1 : 0; pp.guide_mode = p->guide_mode; pp.tile = ls->T;
        pp.res = ls->prm.map_resolution; pp.tiles_x = ls->tx; pp.tiles_y = ls->ty;
        pp.pix_slice = ls->kd ? ls->pix_slice.p : nullptr;
        pp.x0 = p->w > 0 ? p->x0 : 0; pp.y0 = p->h > 0 ? p->y0 : 0;
        pp.w = p->w > 0 ? p->w : ls->W; pp.h = p->h > 0 ? p->h : ls->H;
        RZK_CHECK(pp.x0 >= 0 && pp.y0 >= 0 && pp.x0 + pp.w <= ls->W && pp.y0 + pp.h <= ls->H, RZK_E_INVALID,
                  "window outside the image");
        pp.seed = p->seed; pp.tmin = p->tmin;
        GuideView gv;
        gv.maps = ls->maps.p; gv.row_cdf = ls->row_cdf.p; gv.col_cdf = ls->col_cdf.p; gv.totals = ls->totals.p;
        int64_t n = (int64_t)ls->W * ls->H;
        if (ls->pt_image.n != (size_t)(3 * n)) ls->pt_image.alloc(3 * n);
        if (ls->pt_work.n == 0) ls->pt_work.alloc(4);
        if ((int64_t)pp.w * pp.h != n)   // pixels outside the window read back as 0
            RZK_HIP(hipMemsetAsync(ls->pt_image.p, 0, sizeof(float) * 3 * n, ctx->stream));
        if (out_rays) RZK_HIP(hipMemsetAsync(ls->pt_work.p, 0, 4 * sizeof(unsigned long long), ctx->stream));
        launch_guided_pt(ctx, scene_view(ls->scene), camera, pp, gv, ls->pt_image.p, out_rays ? ls->pt_work.p : nullptr);
        if (out_rgb) RZK_HIP(hipMemcpyAsync(out_rgb, ls->pt_image.p, sizeof(float) * 3 * n, hipMemcpyDeviceToHost,
                                            ctx->stream));
        if (out_rays) {
            unsigned long long h[4];
            RZK_HIP(hipMemcpyAsync(h, ls->pt_work.p, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
            ls_sync(ctx);
            out_rays[0] = (int64_t)h[0];
            out_rays[1] = (int64_t)h[1];
            for (int i = 0; i < 4; ++i) ls->h_pt_work[i] = (int64_t)h[i];
        }
        if (out_rgb) ls_sync(ctx);
    });
}

int rzk_lightslice_guide(rzk_lightslice* ls, const float* camera, rzk_guide** out) {
    if (!ls || !camera || !out) return RZK_E_INVALID;
    *out = nullptr;
    rzk_ctx* ctx = ls->ctx;
    return ls_guard(ctx, [&] {
        RZK_CHECK(ls->initialized, RZK_E_STATE, "rzk_lightslice_initialize has not run");
        RZK_CHECK((int)camera[17] == ls->W && (int)camera[18] == ls->H, RZK_E_INVALID, "camera size differs");
        RZK_CHECK(!ls->kd, RZK_E_STATE, "guides are laid out by pixel tiles; this LightSlice has kd-tree slices");
        std::unique_ptr<rzk_guide> g(new rzk_guide());
        g->ctx = ctx;
        g->scene = ls->scene;
        std::memcpy(g->cam, camera, sizeof(g->cam));
        g->W = ls->W; g->H = ls->H; g->tile = ls->T; g->tiles_x = ls->tx; g->tiles_y = ls->ty;
        g->res = ls->prm.map_resolution;
        g->gv.maps = ls->maps.p; g->gv.row_cdf = ls->row_cdf.p; g->gv.col_cdf = ls->col_cdf.p;
        g->gv.totals = ls->totals.p;
        ctx_retain(ctx);
        *out = g.release();
    });
}

int rzk_guide_destroy(rzk_guide* g) {
    if (!g) return RZK_OK;
    (void)hipStreamSynchronize(g->ctx->stream);
    for (void* p : g->owned) (void)hipFree(p);
    rzk_ctx* ctx = g->ctx;
    delete g;
    ctx_release(ctx);
    return RZK_OK;
}

int rzk_lightslice_count_work(rzk_lightslice* ls, int64_t* out) {
    if (!ls || !out) return RZK_E_INVALID;
    rzk_ctx* ctx = ls->ctx;
    return ls_guard(ctx, [&] {
        RZK_CHECK(ls->initialized, RZK_E_STATE, "rzk_lightslice_initialize has not run");
        const rzk_lightslice_params& p = ls->prm;
        SceneView sv = scene_view(ls->scene);
        int S = ls->S;
        int64_t V = ls->V, n = (int64_t)ls->W * ls->H;
        int nreps = (int)ls->h_reps.size();
        DBuf<unsigned long long> tot;
        tot.alloc(4);
        RZK_HIP(hipMemsetAsync(tot.p, 0, sizeof(unsigned long long) * 4, ctx->stream));
        launch_reduced_matrix(ctx, sv, ls->rows_recv.p, ls->rows_mat.p, S, ls->vpls->d_order, ls->vpls->d_vpl, ls->vpls->d_mat, V,
                              p.clamp_mode, p.incident_clamp, p.tmin, ls->R.p, ls->Vm.p, (int*)tot.p);
        launch_slice_gather(ctx, sv, ls->recv.p, ls->mat.p, ls->W, ls->H, ls->vpls->d_vpl, ls->vpls->d_mat,
                            ls->slice_reps(), p.clamp_mode, p.clamp, p.tmin, p.n_emitted, ls->image.p,
                            ls->vis_bits.p, (int*)(tot.p + 2));
        unsigned long long h[4];
        std::vector<int> rm(S), pm(n);
        RZK_HIP(hipMemcpyAsync(h, tot.p, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
        RZK_HIP(hipMemcpyAsync(rm.data(), ls->rows_mat.p, sizeof(int) * S, hipMemcpyDeviceToHost, ctx->stream));
        RZK_HIP(hipMemcpyAsync(pm.data(), ls->mat.p, sizeof(int) * n, hipMemcpyDeviceToHost, ctx->stream));
        ls_sync(ctx);
        int64_t rows_ok = 0, pairs = 0;
        for (int v : rm) rows_ok += v >= 0;
        for (int64_t i = 0; i < n; ++i)        // pixels with a material x their slice's clusters
            if (pm[i] >= 0) pairs += ls->h_scount[ls->slice_of_pixel(i)];
        out[0] = n;
        out[1] = rows_ok * V;
        out[2] = (int64_t)h[0];
        out[3] = (int64_t)h[1];
        out[4] = pairs;
        out[5] = (int64_t)h[2];
        out[6] = (int64_t)h[3];
        out[7] = nreps;
        ls->info.gather_pairs = out[4];
    });
}

int rzk_lightslice_get_info(const rzk_lightslice* ls, rzk_lightslice_info* out) {
    if (!ls || !out) return RZK_E_INVALID;
    *out = ls->info;
    return RZK_OK;
}

int rzk_lightslice_get(const rzk_lightslice* ls, const char* name, void* dst, int64_t bytes) {
    if (!ls || !name || !dst) return RZK_E_INVALID;
    rzk_ctx* ctx = ls->ctx;
    return ls_guard(ctx, [&] {
        std::string nm(name);
        int S = ls->S, res = ls->prm.map_resolution;
        int64_t V = ls->V, n = (int64_t)ls->W * ls->H;
        const int64_t cw = ls->cw;
        const int nreps = (int)ls->h_reps.size();
        const void* src = nullptr;
        int64_t need = 0;
        bool host = false;
        if (nm == "R") { src = ls->R.p; need = 12 * (int64_t)S * V; }
        else if (nm == "V") { src = ls->Vm.p; need = 4 * (int64_t)S * V; }
        else if (nm == "alphas") { src = ls->h_alphas.p; need = 4 * V; host = true; }
        else if (nm == "mags") { src = ls->h_mags.p; need = 4 * V; host = true; }
        else if (nm == "centers") { src = ls->h_centers.data(); need = 4 * (int64_t)ls->h_centers.size(); host = true; }
        else if (nm == "center_w") { src = ls->h_center_w.data(); need = 4 * (int64_t)ls->h_center_w.size(); host = true; }
        else if (nm == "assign") { src = ls->h_assign.p; need = 4 * V; host = true; }
        else if (nm == "reps") { src = ls->h_reps.data(); need = 4 * (int64_t)nreps; host = true; }
        else if (nm == "rep_w") { src = ls->h_rep_w.data(); need = 4 * (int64_t)nreps; host = true; }
        else if (nm == "rows_recv") { src = ls->rows_recv.p; need = 4 * (int64_t)RZK_RECV_FLOATS * S; }
        else if (nm == "rows_mat") { src = ls->rows_mat.p; need = 4 * (int64_t)S; }
        else if (nm == "slice_pdf") { src = ls->slice_pdf.p; need = 4 * (int64_t)S * cw; }
        else if (nm == "slice_cdf") { src = ls->slice_cdf.p; need = 4 * (int64_t)S * cw; }
        else if (nm == "slice_upper") { src = ls->upper.p; need = 4 * (int64_t)S * cw; }
        else if (nm == "slice_wi") { src = ls->slice_wi.p; need = 12 * (int64_t)S * cw; }
        else if (nm == "slice_count") { src = ls->h_scount.data(); need = 4 * (int64_t)ls->h_scount.size(); host = true; }
        else if (nm == "slice_reps" || nm == "slice_rep_w") {
            need = 4 * (int64_t)S * cw;
            RZK_CHECK(bytes == need, RZK_E_INVALID, "size mismatch for " + nm + ": need " + std::to_string(need));
            if (ls->ld) {
                RZK_HIP(hipMemcpyAsync(dst, nm == "slice_reps" ? (const void*)ls->sreps.p : (const void*)ls->srep_w.p,
                                       (size_t)need, hipMemcpyDeviceToHost, ctx->stream));
                ls_sync(ctx);
            } else {
                for (int s = 0; s < S; ++s)
                    std::memcpy((char*)dst + (size_t)s * cw * 4,
                                nm == "slice_reps" ? (const void*)ls->h_reps.data() : (const void*)ls->h_rep_w.data(),
                                (size_t)cw * 4);
            }
            return;
        }
        else if (nm == "maps") { src = ls->maps.p; need = 4 * (int64_t)S * res * res; }
        else if (nm == "map_totals") { src = ls->totals.p; need = 4 * (int64_t)S; }
        else if (nm == "row_cdf") { src = ls->row_cdf.p; need = 4 * (int64_t)S * res; }
        else if (nm == "col_cdf") { src = ls->col_cdf.p; need = 4 * (int64_t)S * res * res; }
        else if (nm == "pt_work") { src = ls->h_pt_work; need = 8 * 4; host = true; }
        else if (nm == "neighbors") { src = ls->nb.p; need = 4 * (int64_t)S * ls->prm.num_neighbor_slices; }
        else if (nm == "image") { src = ls->image.p; need = 12 * n; }
        else if (nm == "recv") { src = ls->recv.p; need = 4 * (int64_t)RZK_RECV_FLOATS * n; }
        else if (nm == "recv_mat") { src = ls->mat.p; need = 4 * n; }
        else if (nm == "hits") { src = ls->hits.p; need = 16 * n; }
        else if (nm == "row_pix") { src = ls->h_row_pix.data(); need = 4 * (int64_t)ls->h_row_pix.size(); host = true; }
        else if (nm == "pixel_slice") {
            RZK_CHECK(ls->kd, RZK_E_INVALID, "pixel_slice: this LightSlice has pixel-tile slices");
            src = ls->h_pix_slice.data(); need = 4 * (int64_t)ls->h_pix_slice.size(); host = true;
        }
        else if (nm == "gather_pix") {
            RZK_CHECK(ls->kd, RZK_E_INVALID, "gather_pix: this LightSlice has pixel-tile slices");
            src = ls->h_gp_pix.data(); need = 4 * (int64_t)ls->h_gp_pix.size(); host = true;
        }
        else throw RzkError{RZK_E_INVALID, "unknown array name " + nm};
        RZK_CHECK(bytes == need, RZK_E_INVALID, "size mismatch for " + nm + ": need " + std::to_string(need));
        if (need == 0) return;
        RZK_CHECK(src != nullptr, RZK_E_STATE, nm + " is not available before initialize");
        if (host) std::memcpy(dst, src, (size_t)need);
        else {
            RZK_HIP(hipMemcpyAsync(dst, src, (size_t)need, hipMemcpyDeviceToHost, ctx->stream));
            ls_sync(ctx);
        }
    });
}

}  // extern "C"
