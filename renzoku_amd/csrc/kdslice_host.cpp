// kdslice_host.cpp — the 6-D kd-tree slicing with the reference's pivot
// stream: every pivot of every median selection is the next number of one
// MT19937 stream, taken in the reference's depth-first recursion order
// (KdTreeN::cluster_kdtree + quick_select_k, kdtree_n.h:192, :285-365; the
// scene's Random).  The draw count of a node depends on the points, so node
// i + 1 cannot start before node i has drawn: the build is sequential and runs
// on the host (about 0.1 s for 250k points); the tree it leaves on the device
// has the layout kd_build_device leaves, and the queries are the same kernels.
//
// Each selection works on one index array and the node's keys along its axis:
// a round partitions the active range stably into less / equal / greater than
// the pivot (a counting pass, then a scatter through a second buffer).  The
// reference ping-pongs between two arrays and syncs the settled parts back;
// both arrays end equal, holding exactly what the stable partitions on one
// array leave, so the order of the points matches it position for position.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "host_sampling.h"
#include "kdslice.h"

namespace {

constexpr int kDim = 6;
constexpr float kFltMax = 3.402823466e+38f;

struct HostKd {
    const float* P = nullptr;          // 6 planes of n floats
    int n = 0, max_size = 0;
    rzk::Mt19937* rd = nullptr;
    int64_t draws = 0;
    int64_t cap = 0;                   // heap entries kept for axis / plane
    std::vector<int> idx, tmp_i, axis;
    std::vector<float> key, tmp_k, plane, boxes;
    std::vector<int> slices;
    std::vector<int64_t> slice_heap;

    void node(int64_t heap, bool clustered, int s, int e) {
        if (s >= e) return;
        // the node's box (BoundingBoxN::merge: std::min / std::max, point order)
        float lo[kDim], hi[kDim];
        for (int a = 0; a < kDim; ++a) {
            const float* pa = P + (size_t)a * n;
            float l = kFltMax, h = -kFltMax;
            for (int i = s; i <= e; ++i) {
                const float v = pa[idx[i]];
                l = (v < l) ? v : l;
                h = (h < v) ? v : h;
            }
            lo[a] = l;
            hi[a] = h;
        }
        // the first node of at most max_size points on the path is a slice
        if (!clustered && e - s + 1 <= max_size) {
            slices.push_back(s);
            slices.push_back(e);
            boxes.insert(boxes.end(), lo, lo + kDim);
            boxes.insert(boxes.end(), hi, hi + kDim);
            slice_heap.push_back(heap);
            clustered = true;
        }
        // get_longest_axis: first axis of strictly largest |extent|
        int ax = 0;
        float longest = 0.0f;
        for (int a = 0; a < kDim; ++a) {
            const float x = std::fabs(hi[a] - lo[a]);
            if (x > longest) { ax = a; longest = x; }
        }
        const int m = (s + e) / 2;
        const float* pk = P + (size_t)ax * n;
        for (int i = s; i <= e; ++i) key[i] = pk[idx[i]];
        int l = s, h = e;
        for (;;) {
            const float u = rd->next_float();
            ++draws;
            const float v = key[l + (int)(u * (float)(h - l + 1))];
            int nl = 0, ne = 0;
            for (int i = l; i <= h; ++i) {
                nl += key[i] < v;
                ne += key[i] == v;
            }
            int a = l, b = l + nl, c = l + nl + ne;
            for (int i = l; i <= h; ++i) {
                const float k = key[i];
                const int d = (k < v) ? a++ : (k == v) ? b++ : c++;
                tmp_k[d] = k;
                tmp_i[d] = idx[i];
            }
            std::memcpy(&key[l], &tmp_k[l], sizeof(float) * (size_t)(h - l + 1));
            std::memcpy(&idx[l], &tmp_i[l], sizeof(int) * (size_t)(h - l + 1));
            const int eq_s = l + nl, eq_e = l + nl + ne - 1;
            if (m < eq_s) h = eq_s - 1;
            else if (m <= eq_e) break;
            else l = eq_e + 1;
        }
        if (heap < cap) {
            axis[(size_t)heap] = ax;
            plane[(size_t)heap] = key[m];
        }
        node(2 * heap + 1, clustered, s, m);
        node(2 * heap + 2, clustered, m + 1, e);
    }
};

}  // namespace

int64_t kd_build_host_mt(rzk_kdslices* kd, const float* h_planes, rzk::Mt19937& rd) {
    rzk_ctx* ctx = kd->ctx;
    hipStream_t st = ctx->stream;
    const int n = kd->n;
    for (size_t i = 0; i < (size_t)kDim * n; ++i)
        RZK_CHECK(std::isfinite(h_planes[i]), RZK_E_INVALID, "kd-tree point " + std::to_string(i % (size_t)n) +
                                                                  " is not finite");
    HostKd b;
    b.P = h_planes;
    b.n = n;
    b.max_size = kd->max_size;
    b.rd = &rd;
    b.cap = 4 * (int64_t)std::max(n, 1) + 4;           // heap indices stay below 4n (depth <= log2(n) + 2)
    b.idx.resize((size_t)n);
    for (int i = 0; i < n; ++i) b.idx[i] = i;
    b.tmp_i.resize((size_t)n);
    b.key.resize((size_t)n);
    b.tmp_k.resize((size_t)n);
    b.axis.assign((size_t)b.cap, 127);
    b.plane.assign((size_t)b.cap, kFltMax);
    if (n >= 2) b.node(0, false, 0, n - 1);
    const int S = (int)b.slice_heap.size();
    int64_t H = 0;
    for (int64_t h : b.slice_heap) H = std::max(H, h + 1);
    RZK_CHECK(H <= b.cap, RZK_E_STATE, "kd-tree heap beyond its bound");
    std::vector<int> cl((size_t)std::max<int64_t>(H, 1), -1);
    for (int s = 0; s < S; ++s) cl[(size_t)b.slice_heap[s]] = s;
    kd->S = S;
    kd->H = (int)H;
    kd->h_slices = b.slices;
    kd->ord.alloc(n);
    kd->axis.alloc(H);
    kd->plane.alloc(H);
    kd->cl.alloc(H);
    kd->boxes.alloc((size_t)12 * S);
    kd->slices.alloc((size_t)2 * S);
    if (n > 0) RZK_HIP(hipMemcpyAsync(kd->ord.p, b.idx.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice, st));
    if (H > 0) {
        RZK_HIP(hipMemcpyAsync(kd->axis.p, b.axis.data(), sizeof(int) * (size_t)H, hipMemcpyHostToDevice, st));
        RZK_HIP(hipMemcpyAsync(kd->plane.p, b.plane.data(), sizeof(float) * (size_t)H, hipMemcpyHostToDevice, st));
        RZK_HIP(hipMemcpyAsync(kd->cl.p, cl.data(), sizeof(int) * (size_t)H, hipMemcpyHostToDevice, st));
    }
    if (S > 0) {
        RZK_HIP(hipMemcpyAsync(kd->boxes.p, b.boxes.data(), sizeof(float) * 12 * (size_t)S, hipMemcpyHostToDevice, st));
        RZK_HIP(hipMemcpyAsync(kd->slices.p, b.slices.data(), sizeof(int) * 2 * (size_t)S, hipMemcpyHostToDevice, st));
    }
    RZK_HIP(hipStreamSynchronize(st));
    return b.draws;
}
