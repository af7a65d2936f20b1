// kdslice.cpp — CPU restatement of LightSlice's 6-D kd-tree slicing (test
// infrastructure only: the checker of the device build, never the product).
//
//   map_vec6              lightslice.cc:142-188
//   cluster_kdtree        kdtree_n.h:285-365, :399-429 (partial build that records
//                         every first node of at most max_slice_size points as a
//                         slice, then builds on to single points)
//   quick_select_k        kdtree_n.h:181-242 (ping-pong 3-way partition)
//   get_longest_axis      boundingbox_n.h:74-85 (first axis of strictly largest |extent|)
//   find_cluster_rec      kdtree_n.h:540-572 (PLANE_BIAS = -1e-3, kdtree_n.h:171)
//   find_nearest_cluster  kdtree_n.h:575-647 (std::priority_queue by the squared
//                         distance to the cluster's box; furthest first)
//
// The recursion, the ping-pong buffers and the heap are the reference's own
// shapes.  Two pivot streams: orc_kd_build_mt takes every pivot from the
// oracle's global MT stream (orc_mt_rand) in recursion order, as the reference
// does (kdtree_n.h:192, the scene's Random); orc_kd_build takes round t of the
// selection at node [s, e] from Philox4x32-10 keyed by the build seed with
// counter (s, e, t), so that every node can be built on its own.  Nodes are
// numbered as a binary heap (children 2i+1, 2i+2), which fixes the same tree as
// the reference's allocation order.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <queue>
#include <vector>

extern "C" void orc_philox(uint32_t c[4], uint32_t k0, uint32_t k1);
extern "C" float orc_mt_rand(void);

namespace {

constexpr int kDim = 6;
constexpr float kPlaneBias = -1e-3f;

struct Pt {
    float p[kDim];
    int index;
    int cluster = -1;
};

float pivot_u(uint32_t seed, int s, int e, int t) {
    uint32_t c[4] = {(uint32_t)s, (uint32_t)e, (uint32_t)t, 0x6b64u};
    orc_philox(c, seed ^ 0x2545f491u, 0x6b647472u);
    return (float)(c[0] & 0x00ffffffu) * (1.0f / 16777216.0f);
}

int compare_data(int axis, const Pt& a, const Pt& b) {
    return (a.p[axis] < b.p[axis]) ? -1 : (a.p[axis] == b.p[axis]) ? 0 : 1;
}

struct Box {
    float vmin[kDim], vmax[kDim];
    Box() {
        for (int i = 0; i < kDim; ++i) { vmin[i] = 3.402823466e+38f; vmax[i] = -3.402823466e+38f; }
    }
    void merge(const float* p) {          // std::min / std::max
        for (int i = 0; i < kDim; ++i) {
            vmin[i] = (p[i] < vmin[i]) ? p[i] : vmin[i];
            vmax[i] = (vmax[i] < p[i]) ? p[i] : vmax[i];
        }
    }
    int longest_axis() const {
        int index = 0;
        float longest = 0.0f;
        for (int i = 0; i < kDim; ++i) {
            float a = std::fabs(vmax[i] - vmin[i]);
            if (a > longest) { index = i; longest = a; }
        }
        return index;
    }
    float min_squared_distance(const float* p) const {
        float sum = 0.0f;
        for (int i = 0; i < kDim; ++i) {
            float c = (p[i] < vmin[i]) ? vmin[i] : p[i];    // std::max(p, vmin)
            c = (vmax[i] < c) ? vmax[i] : c;                 // std::min(., vmax)
            float d = c - p[i];
            sum += d * d;
        }
        return sum;
    }
};

struct Builder {
    uint32_t seed;
    int max_size;
    int cap_heap;
    int* axis;
    float* plane;
    int* cl;
    std::vector<int> sl_start, sl_end;
    std::vector<Box> boxes;
    int max_heap = -1;
    bool mt = false;          // pivots from the global MT stream in recursion order
    int64_t draws = 0;        // MT numbers taken

    // quick_select_k: round t of node [s0, e0]
    void quick_select(int s0, int e0, int t, int axis_, int k, Pt* in, int start, int end, Pt* out) {
        if (start > end) return;
        float u;
        if (mt) {
            u = orc_mt_rand();
            ++draws;
        } else {
            u = pivot_u(seed, s0, e0, t);
        }
        int p = start + (int)(u * (float)(end - start + 1));
        const Pt v = in[p];
        int pos = start;
        for (int i = start; i <= end; ++i)
            if (compare_data(axis_, in[i], v) < 0) out[pos++] = in[i];
        int equal_start = pos;
        for (int i = start; i <= end; ++i)
            if (compare_data(axis_, in[i], v) == 0) out[pos++] = in[i];
        int equal_end = pos - 1;
        for (int i = start; i <= end; ++i)
            if (compare_data(axis_, in[i], v) > 0) out[pos++] = in[i];
        if (k >= equal_start && k <= equal_end) {
            for (int i = start; i <= end; ++i) in[i] = out[i];
            return;
        } else if (k < equal_start) {
            for (int i = equal_start; i <= end; ++i) in[i] = out[i];
            quick_select(s0, e0, t + 1, axis_, k, out, start, equal_start - 1, in);
        } else {
            for (int i = start; i <= equal_end; ++i) in[i] = out[i];
            quick_select(s0, e0, t + 1, axis_, k, out, equal_end + 1, end, in);
        }
    }

    // cluster_kdtree
    void build(int64_t heap, bool clustered, Pt* data, int start, int end, Pt* median_data) {
        if (start >= end) return;
        if (end - start + 1 <= max_size && data[start].cluster < 0 && !clustered) {
            const int id = (int)sl_start.size();
            Box bb;
            for (int i = start; i <= end; ++i) {
                data[i].cluster = id;
                median_data[i].cluster = id;
                bb.merge(data[i].p);
            }
            sl_start.push_back(start);
            sl_end.push_back(end);
            boxes.push_back(bb);
            if (heap < cap_heap) cl[heap] = id;
            if (heap > max_heap) max_heap = (int)heap;
            clustered = true;
        }
        Box bb;
        for (int i = start; i <= end; ++i) bb.merge(data[i].p);
        const int ax = bb.longest_axis();
        const int median = (end + start) / 2;
        quick_select(start, end, 0, ax, median, data, start, end, median_data);
        if (heap < cap_heap) {
            axis[heap] = ax;
            plane[heap] = median_data[median].p[ax];
        }
        build(2 * heap + 1, clustered, median_data, start, median, data);
        build(2 * heap + 2, clustered, median_data, median + 1, end, data);
    }
};

// find_nearest_cluster's priority queue: less = nearer box
struct ClusterLess {
    const std::vector<float>* d;
    bool operator()(int a, int b) const { return (*d)[a] < (*d)[b]; }
};

}  // namespace

static int kd_build(bool mt, const float* pts, int n, int max_size, uint32_t seed, int* order, int* slices,
                    float* boxes, int cap_slices, int* axis, float* plane, int* cl, int cap_heap, int* heap_used,
                    int64_t* draws) {
    if (n < 0 || max_size < 3) return -1;
    for (int i = 0; i < cap_heap; ++i) { axis[i] = 127; plane[i] = 3.402823466e+38f; cl[i] = -1; }
    std::vector<Pt> data((size_t)n), median_data;
    for (int i = 0; i < n; ++i) {
        std::memcpy(data[i].p, pts + (size_t)kDim * i, sizeof(float) * kDim);
        data[i].index = i;
    }
    median_data = data;
    Builder b;
    b.seed = seed; b.max_size = max_size; b.cap_heap = cap_heap;
    b.axis = axis; b.plane = plane; b.cl = cl; b.mt = mt;
    if (n > 0) b.build(0, false, data.data(), 0, n - 1, median_data.data());
    if (draws) *draws = b.draws;
    for (int i = 0; i < n; ++i) order[i] = data[i].index;
    const int S = (int)b.sl_start.size();
    if (S > cap_slices) return -1;
    for (int s = 0; s < S; ++s) {
        slices[2 * s] = b.sl_start[s];
        slices[2 * s + 1] = b.sl_end[s];
        std::memcpy(boxes + 12 * s, b.boxes[s].vmin, sizeof(float) * kDim);
        std::memcpy(boxes + 12 * s + kDim, b.boxes[s].vmax, sizeof(float) * kDim);
    }
    *heap_used = b.max_heap + 1;
    return S;
}

extern "C" {

// LightSlice map_vec6 of one receiver record (p at 0, shading normal at 6, wo
// at 9) with its material record (NULL: no material)
void orc_map_vec6(const float* rec, const float* mat, float diagonal, float* out) {
    const float weight = 1.0f / 16.0f * diagonal;
    const float max_gloss_scale = 4.0f, max_gloss_exponential = 1000.0f;
    const float nx = rec[6], ny = rec[7], nz = rec[8], wx = rec[9], wy = rec[10], wz = rec[11];
    float dx = nx, dy = ny, dz = nz;
    const int type = mat ? (int)mat[0] : 0;
    if (type == 1 || type == 2) {
        float g;
        if (type == 1) {
            g = mat[7];
        } else {
            float ix = 1.0f / mat[7], iy = 1.0f / mat[8];
            g = (ix < iy) ? iy : ix;                          // std::max
        }
        float e = g / max_gloss_exponential;
        float mx = (1.0f < e) ? e : 1.0f;                     // std::max(1, e)
        float gs = (mx < max_gloss_scale) ? mx : max_gloss_scale;   // std::min(4, .)
        float d2 = 2.0f * (wx * nx + wy * ny + wz * nz);
        dx = (d2 * nx - wx) * gs;
        dy = (d2 * ny - wy) * gs;
        dz = (d2 * nz - wz) * gs;
    }
    out[0] = rec[0]; out[1] = rec[1]; out[2] = rec[2];
    out[3] = dx * weight; out[4] = dy * weight; out[5] = dz * weight;
}

// The slicing of n points (n x 6).  Outputs: order[n] = point index at each kd
// position; slices[2 * s] = first and last kd position of slice s (DFS order);
// boxes[12 * s] = min[6], max[6]; heap arrays (cap_heap entries, filled below
// the first slice level too where they fit): axis, plane, cl (slice of the node
// where the slice was recorded, else -1).  Returns the number of slices; *heap_used
// = 1 + the largest heap index of a slice node.  -1 on bad arguments.
int orc_kd_build(const float* pts, int n, int max_size, uint32_t seed, int* order, int* slices, float* boxes,
                 int cap_slices, int* axis, float* plane, int* cl, int cap_heap, int* heap_used) {
    return kd_build(false, pts, n, max_size, seed, order, slices, boxes, cap_slices, axis, plane, cl, cap_heap,
                    heap_used, nullptr);
}

// the same with every pivot drawn from the global MT stream (orc_mt_seed32 /
// orc_mt_rand) in the reference's recursion order; *draws = numbers taken
int orc_kd_build_mt(const float* pts, int n, int max_size, int* order, int* slices, float* boxes, int cap_slices,
                    int* axis, float* plane, int* cl, int cap_heap, int* heap_used, int64_t* draws) {
    return kd_build(true, pts, n, max_size, 0, order, slices, boxes, cap_slices, axis, plane, cl, cap_heap,
                    heap_used, draws);
}

// find_cluster for m query points; -1 where the tree has no slice
void orc_kd_find(const int* axis, const float* plane, const int* cl, int heap_used, const float* q, int m,
                 int* out) {
    for (int i = 0; i < m; ++i) {
        const float* p = q + (size_t)kDim * i;
        int node = 0, r = -1;
        while (node < heap_used) {
            if (cl[node] >= 0) { r = cl[node]; break; }
            float dist = p[axis[node]] - plane[node] + kPlaneBias;
            node = dist <= 0.0f ? 2 * node + 1 : 2 * node + 2;
        }
        out[i] = r;
    }
}

// find_nearest_cluster(p, k) for m query points: out[k * i ...] furthest first,
// padded with -1
void orc_kd_nearest(const int* axis, const float* plane, const int* cl, int heap_used, const float* boxes, int nslices,
                    const float* q, int m, int k, int* out) {
    for (int i = 0; i < m; ++i) {
        const float* p = q + (size_t)kDim * i;
        std::vector<float> d((size_t)nslices);
        std::vector<Box> bb((size_t)nslices);
        for (int s = 0; s < nslices; ++s) {
            std::memcpy(bb[s].vmin, boxes + 12 * s, sizeof(float) * kDim);
            std::memcpy(bb[s].vmax, boxes + 12 * s + kDim, sizeof(float) * kDim);
            d[s] = bb[s].min_squared_distance(p);
        }
        std::priority_queue<int, std::vector<int>, ClusterLess> heap(ClusterLess{&d});
        // find_nearest_cluster_rec
        struct Rec {
            const int* axis; const float* plane; const int* cl; int heap_used; const float* p; int k;
            const std::vector<float>* d;
            void go(int node, std::priority_queue<int, std::vector<int>, ClusterLess>& h) const {
                if (node >= heap_used) return;
                if (cl[node] >= 0) {
                    int c = cl[node];
                    if ((int)h.size() < k) h.push(c);
                    else if ((*d)[c] < (*d)[h.top()]) { h.pop(); h.push(c); }
                    return;
                }
                float dist = p[axis[node]] - plane[node] + kPlaneBias;
                int first = dist <= 0.0f ? 2 * node + 1 : 2 * node + 2;
                int second = dist <= 0.0f ? 2 * node + 2 : 2 * node + 1;
                go(first, h);
                if ((int)h.size() < k || (*d)[h.top()] > dist * dist) go(second, h);
            }
        } rec{axis, plane, cl, heap_used, p, k, &d};
        if (nslices > 0) rec.go(0, heap);
        int o = 0;
        while (!heap.empty()) { out[(size_t)k * i + o++] = heap.top(); heap.pop(); }
        for (; o < k; ++o) out[(size_t)k * i + o] = -1;
    }
}

}  // extern "C"
