/* refine.cpp — CPU restatement of LightSlice's per-slice cluster refinement
 * (lightslice.cc:837-902 with compute_distance :336-357, choose_representative
 * :385-400, split_cluster :405-488, pick :326-331), for the test oracle only.
 *
 * It is C++ because the reference's results depend on two standard-library
 * algorithms whose tie orders are not specified by the C++ standard:
 * std::sort (not stable) of the projected columns and std::priority_queue of
 * the cluster costs.  Calling them here, built by the same toolchain's
 * libstdc++ the reference links, reproduces those orders by construction.
 * Draws come from the oracle's global MT stream (orc_mt_rand), continuing
 * where cluster_by_sampling left it, as the reference's scene Random does.
 *
 * Test infrastructure: only tests/, smoke() and bench.py's CPU leg use it.
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <queue>
#include <utility>
#include <vector>

extern "C" {
float orc_mt_rand(void);
int orc_pdf_build(int n, const float *arr, float *pdf, float *cdf);
void orc_pdf_sample(int n, const float *pdf, const float *cdf, float u, int *k, float *pk);
}

namespace {

/* DiscretePdf (pdf.cc:18-77) */
struct Pdf {
    std::vector<float> pdf, cdf;
    void set(const std::vector<float> &a) {
        pdf.assign(a.size(), 0.f);
        cdf.assign(a.size(), 0.f);
        if (!a.empty()) orc_pdf_build((int)a.size(), a.data(), pdf.data(), cdf.data());
    }
    /* one draw per call, whatever the distribution (pdf.cc:59-60) */
    void sample(int &k, float &pk) const {
        float u = orc_mt_rand();
        orc_pdf_sample((int)cdf.size(), pdf.data(), cdf.data(), u, &k, &pk);
    }
};

struct Cluster {
    std::vector<int> indices;
    Pdf pdf;
    int representative = -1;
    float weight = 0.f;
};

struct ClusterCost {
    int index;
    float cost;
    ClusterCost(int i, float c) : index(i), cost(c) {}
};
struct ClusterCostLess {
    bool operator()(const ClusterCost &a, const ClusterCost &b) const { return a.cost < b.cost; }
};

/* the slice matrix L: rows nb[0..nnb) of R, flattened to 3*nnb rows
 * (MatrixC<Rgb>::subrow + flatten, matrix_c.h:251-259, matrix_c.cc:62-72) */
struct SliceMatrix {
    const float *R;
    int S, nnb;
    const int *nb;
    float at(int row, int col) const {      /* flattened row = 3 * q + channel */
        int q = row / 3, ch = row - 3 * q;
        return R[((int64_t)col * S + nb[q]) * 3 + ch];
    }
    int rows() const { return 3 * nnb; }
};

/* compute_distance (lightslice.cc:336-357) on the columns `idx` of L */
float compute_distance(const SliceMatrix &L, const std::vector<int> &idx) {
    const int M = L.rows(), n = (int)idx.size();
    std::vector<float> norms(n), u(M, 0.f), v(n, 0.f);
    for (int j = 0; j < n; ++j) {               /* MatrixC<Float>::get_column_norm, matrix_c.cc:90-102 */
        float mag = 0.f;
        for (int r = 0; r < M; ++r) {
            float x = L.at(r, idx[j]);
            mag += x * x;
        }
        norms[j] = std::sqrt(mag);
    }
    for (int j = 0; j < n; ++j)                 /* row_sum, matrix_c.h:236-245 */
        for (int r = 0; r < M; ++r) u[r] += L.at(r, idx[j]);
    for (int r = 0; r < M; ++r)                 /* Rt.mul(u), matrix_c.h:224-233 */
        for (int j = 0; j < n; ++j) v[j] += L.at(r, idx[j]) * u[r];
    float c = 0.f;                              /* sum(column_norms), lightslice.cc:319-324 */
    for (int j = 0; j < n; ++j) c += norms[j];
    float cost = 0.f;
    for (int j = 0; j < n; ++j) cost += norms[j] * c - v[j];
    return cost;
}

void choose_representative(Cluster &c) {       /* lightslice.cc:385-400 */
    int j;
    float pj;
    c.pdf.sample(j, pj);
    if (j < 0) {
        c.representative = c.indices[c.indices.size() / 2];
        c.weight = 0.0f;
    } else {
        c.representative = c.indices[j];
        c.weight = 1.0f / pj;
    }
}

bool pair_less(const std::pair<int, float> &a, const std::pair<int, float> &b) { return a.second < b.second; }

/* split_cluster (lightslice.cc:405-488) */
void split_cluster(const SliceMatrix &L, const Cluster &cl, std::vector<int> &i1, std::vector<int> &i2) {
    i1.clear();
    i2.clear();
    const int n = (int)cl.indices.size(), M = L.rows();
    if (n == 1) { i1.push_back(cl.indices[0]); return; }
    if (n == 2) { i1.push_back(cl.indices[0]); i2.push_back(cl.indices[1]); return; }
    int i, j;
    float pi, pj;
    cl.pdf.sample(i, pi);
    cl.pdf.sample(j, pj);
    if (i < 0 || j < 0) {
        i1.assign(cl.indices.begin(), cl.indices.begin() + n / 2);
        i2.assign(cl.indices.begin() + n / 2, cl.indices.end());
        return;
    }
    std::vector<float> ab(M);
    for (int r = 0; r < M; ++r) ab[r] = L.at(r, cl.indices[i]) - L.at(r, cl.indices[j]);
    std::vector<std::pair<int, float>> line;
    line.reserve(n);
    for (int k = 0; k < n; ++k) {
        if (k == i || k == j) continue;
        float t = 0.0f;
        for (int r = 0; r < M; ++r) t += (L.at(r, cl.indices[i]) - L.at(r, cl.indices[k])) * ab[r];
        line.push_back(std::pair<int, float>(k, t));
    }
    std::sort(line.begin(), line.end(), pair_less);
    float max_dist = 0.0f;
    int max_k = -1;
    for (int k = 1; k < (int)line.size(); ++k) {
        float d = line[k].second - line[k - 1].second;
        if (d > max_dist) { max_dist = d; max_k = k; }
    }
    if (max_k == -1) {
        i1.assign(cl.indices.begin(), cl.indices.begin() + n / 2);
        i2.assign(cl.indices.begin() + n / 2, cl.indices.end());
    } else {
        i1.push_back(cl.indices[i]);
        i2.push_back(cl.indices[j]);
        for (int k = 0; k < max_k; ++k) i1.push_back(cl.indices[line[k].first]);
        for (int k = max_k; k < (int)line.size(); ++k) i2.push_back(cl.indices[line[k].first]);
    }
}

}  // namespace

/* compute_distance of the columns idx[0..n) of slice matrix rows nb[0..nnb)
 * (exported for the oracle's own checks) */
extern "C" float orc_slice_cost(int S, const float *R, int nnb, const int *nb, int n, const int *idx) {
    SliceMatrix L{R, S, nnb, nb};
    return compute_distance(L, std::vector<int>(idx, idx + n));
}

/* Refine the kept global clusters per slice (lightslice.cc:837-902), slices
 * 0 .. n_slices - 1 (a prefix of the sequential run: same numbers, same results).
 *   R        : column-major Rgb reduced matrix, R[(k * S + r) * 3 + c]
 *   nb       : S x nnb neighbour rows of every slice (the slice matrix L)
 *   cl_off/cl_mem : the ncl kept clusters' members, in order (CSR)
 *   cl_rep, cl_w  : their representatives and weights
 *   mags     : per-VPL flattened column magnitudes (the kept clusters' pdfs,
 *              cluster_builder.cc:228-236)
 * Outputs per slice s (max_slice_clusters entries each): count, rep, weight,
 * size; perm (optional, S x V): the members cluster by cluster; draws: MT
 * numbers consumed.  Returns 0, or -1 on bad sizes. */
extern "C" int orc_refine_slices_n(int S, int n_slices, int V, const float *R, int nnb, const int *nb, int ncl,
                                   const int *cl_off, const int *cl_mem, const int *cl_rep, const float *cl_w,
                                   const float *mags, int max_slice_clusters, int *out_count, int *out_rep,
                                   float *out_w, int *out_size, int *out_perm, int *out_draws) {
    if (S <= 0 || V <= 0 || nnb <= 0 || ncl <= 0 || max_slice_clusters < ncl || n_slices < 0 || n_slices > S)
        return -1;
    /* norms of R's columns (LightTransportMatrix::get_column_norms,
     * light_transport_matrix.cc:117-130): per row r^2 + g^2 + b^2, then summed */
    std::vector<float> norms_R(V);
    for (int k = 0; k < V; ++k) {
        float mag = 0.f;
        for (int r = 0; r < S; ++r) {
            const float *e = R + ((int64_t)k * S + r) * 3;
            mag += e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
        }
        norms_R[k] = std::sqrt(mag);
    }
    std::vector<Cluster> global(ncl);
    for (int c = 0; c < ncl; ++c) {
        Cluster &g = global[c];
        g.indices.assign(cl_mem + cl_off[c], cl_mem + cl_off[c + 1]);
        std::vector<float> m(g.indices.size());
        for (size_t t = 0; t < m.size(); ++t) m[t] = mags[g.indices[t]];
        g.pdf.set(m);
        g.representative = cl_rep[c];
        g.weight = cl_w[c];
    }
    std::vector<int> i1, i2;
    std::vector<float> nrm;
    for (int s = 0; s < n_slices; ++s) {
        SliceMatrix L{R, S, nnb, nb + (int64_t)s * nnb};
        std::vector<Cluster> cls(global);
        std::priority_queue<ClusterCost, std::vector<ClusterCost>, ClusterCostLess> heap;
        for (int c = 0; c < (int)cls.size(); ++c) heap.push(ClusterCost(c, compute_distance(L, cls[c].indices)));
        int draws = 0;
        while (heap.size() > 0 && (int)cls.size() < max_slice_clusters) {
            ClusterCost cc = heap.top();
            if (cls[cc.index].indices.size() <= 1) break;
            if (cls[cc.index].indices.size() > 2) draws += 2;
            split_cluster(L, cls[cc.index], i1, i2);
            Cluster &c1 = cls[cc.index];
            float cost1 = compute_distance(L, i1);
            c1.indices.assign(i1.begin(), i1.end());
            nrm.resize(i1.size());
            for (size_t t = 0; t < i1.size(); ++t) nrm[t] = norms_R[i1[t]];
            c1.pdf.set(nrm);
            choose_representative(c1);
            Cluster c2;
            float cost2 = compute_distance(L, i2);
            c2.indices.assign(i2.begin(), i2.end());
            nrm.resize(i2.size());
            for (size_t t = 0; t < i2.size(); ++t) nrm[t] = norms_R[i2[t]];
            c2.pdf.set(nrm);
            choose_representative(c2);
            draws += 2;
            int new_index = (int)cls.size();
            heap.pop();
            heap.push(ClusterCost(cc.index, cost1));
            heap.push(ClusterCost(new_index, cost2));
            cls.push_back(c2);
        }
        out_count[s] = (int)cls.size();
        out_draws[s] = draws;
        int64_t base = (int64_t)s * max_slice_clusters, pos = 0;
        for (int c = 0; c < (int)cls.size(); ++c) {
            out_rep[base + c] = cls[c].representative;
            out_w[base + c] = cls[c].weight;
            out_size[base + c] = (int)cls[c].indices.size();
            if (out_perm)
                for (int m : cls[c].indices) out_perm[(int64_t)s * V + pos++] = m;
        }
    }
    return 0;
}

/* every slice: orc_refine_slices_n with n_slices = S */
extern "C" int orc_refine_slices(int S, int V, const float *R, int nnb, const int *nb, int ncl, const int *cl_off,
                                 const int *cl_mem, const int *cl_rep, const float *cl_w, const float *mags,
                                 int max_slice_clusters, int *out_count, int *out_rep, float *out_w, int *out_size,
                                 int *out_perm, int *out_draws) {
    return orc_refine_slices_n(S, S, V, R, nnb, nb, ncl, cl_off, cl_mem, cl_rep, cl_w, mags, max_slice_clusters,
                               out_count, out_rep, out_w, out_size, out_perm, out_draws);
}
