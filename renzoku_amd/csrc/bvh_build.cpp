// bvh_build.cpp — host Morton LBVH with the reference's topology
// (AggregateBvh::build, aggregate_bvh.cc:48-240; morton.h:28-120), compacted
// into the device layout of rzk_internal.h.
//
// Topology: triangles sorted stably by 30-bit Morton code of their box centre
// (normalised to the scene box, which always contains the origin), one
// triangle per leaf, every range split after the last code that shares more
// than the range's common prefix with its first code (midpoint for equal
// codes).  The tree is therefore node-for-node the reference's; only the
// memory order differs (right-first depth-first preorder, so the child the
// traversal takes first usually sits next to its parent).
#pragma clang fp contract(off)

#include <algorithm>
#include <cstring>
#include <numeric>

#include "rzk_internal.h"

using namespace rzk;

static int clz32(uint32_t v) { return v ? __builtin_clz(v) : 32; }

// Last index i in [first, last) with clz(code[first] ^ code[i]) above the
// range's common prefix; the predicate holds on a prefix of the sorted range.
static int split_range(const uint32_t* code, int first, int last) {
    uint32_t fc = code[first], lc = code[last];
    if (fc == lc) return (first + last) >> 1;
    int cp = clz32(fc ^ lc);
    int lo = first, hi = last;
    while (hi - lo > 1) {
        int mid = lo + ((hi - lo) >> 1);
        if (clz32(fc ^ code[mid]) > cp) lo = mid; else hi = mid;
    }
    return lo;
}


// ---------------------------------------------------------------------------
// The 4-wide tree the any-hit and path-tracing traversals walk: a binned-SAH
// binary tree over the triangles' boxes (one triangle per leaf, exact union
// boxes) collapsed four-wide by repeatedly opening the interior child with
// the largest surface area.  Any-hit answers do not depend on the topology:
// a triangle is tested iff its own box passes the slab test (every enclosing
// box is a superset, and the rounded slab bounds are monotone in the box), so
// the boolean equals the reference tree's.  Closest hits are the same
// triangle, t, beta and gamma whenever the nearest triangle is unique.

namespace {

struct SahNode {
    float lo[3], hi[3];
    int left = -1, right = -1;       // children, or -1
    int tri = -1;                    // leaf: position in the leaf-ordered triangle array
};

inline float half_area(const float lo[3], const float hi[3]) {
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return dx * dy + dy * dz + dz * dx;
}

struct Bin {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    int n = 0;
    void add(const float* l, const float* h) {
        for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], l[a]); hi[a] = std::max(hi[a], h[a]); }
        ++n;
    }
    void add(const Bin& b) {
        for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], b.lo[a]); hi[a] = std::max(hi[a], b.hi[a]); }
        n += b.n;
    }
};

}  // namespace

void build_sah4(int ntri, const float* pos, const std::vector<v3>& bmin, const std::vector<v3>& bmax, BvhHost& out) {
    constexpr int kBins = 32;
    std::vector<float> tb(6 * (size_t)ntri), cen(3 * (size_t)ntri);
    for (int k = 0; k < ntri; ++k) {
        tb[6 * k + 0] = bmin[k].x; tb[6 * k + 1] = bmin[k].y; tb[6 * k + 2] = bmin[k].z;
        tb[6 * k + 3] = bmax[k].x; tb[6 * k + 4] = bmax[k].y; tb[6 * k + 5] = bmax[k].z;
        for (int a = 0; a < 3; ++a) cen[3 * k + a] = 0.5f * (tb[6 * k + a] + tb[6 * k + 3 + a]);
    }
    std::vector<int> idx(ntri);
    std::iota(idx.begin(), idx.end(), 0);
    std::vector<SahNode> nodes;
    nodes.reserve(2 * (size_t)ntri);
    std::vector<int> leaf_tri;           // leaf order -> original triangle
    leaf_tri.reserve(ntri);
    struct Task { int node, b, e; };
    std::vector<Task> work;
    nodes.emplace_back();
    work.push_back({0, 0, ntri});
    while (!work.empty()) {
        Task t = work.back();
        work.pop_back();
        SahNode& nd = nodes[t.node];
        Bin all, cb;
        for (int i = t.b; i < t.e; ++i) {
            int k = idx[i];
            all.add(&tb[6 * k], &tb[6 * k + 3]);
            cb.add(&cen[3 * k], &cen[3 * k]);
        }
        for (int a = 0; a < 3; ++a) { nd.lo[a] = all.lo[a]; nd.hi[a] = all.hi[a]; }
        int n = t.e - t.b;
        if (n == 1) {
            nd.tri = (int)leaf_tri.size();
            leaf_tri.push_back(idx[t.b]);
            continue;
        }
        // binned SAH over the three axes
        float best = INFINITY;
        int best_axis = -1, best_split = -1;
        for (int a = 0; a < 3; ++a) {
            float ext = cb.hi[a] - cb.lo[a];
            if (!(ext > 0.f)) continue;
            Bin bins[kBins];
            float scale = kBins / ext;
            for (int i = t.b; i < t.e; ++i) {
                int k = idx[i];
                int bi = std::min(kBins - 1, (int)((cen[3 * k + a] - cb.lo[a]) * scale));
                bins[bi].add(&tb[6 * k], &tb[6 * k + 3]);
            }
            float right_area[kBins];
            int right_n[kBins];
            Bin acc;
            for (int i = kBins - 1; i > 0; --i) {
                acc.add(bins[i]);
                right_area[i] = acc.n ? half_area(acc.lo, acc.hi) : 0.f;
                right_n[i] = acc.n;
            }
            Bin left;
            for (int i = 0; i < kBins - 1; ++i) {
                left.add(bins[i]);
                if (left.n == 0 || right_n[i + 1] == 0) continue;
                float cost = left.n * half_area(left.lo, left.hi) + right_n[i + 1] * right_area[i + 1];
                if (cost < best) { best = cost; best_axis = a; best_split = i; }
            }
        }
        int mid;
        if (best_axis < 0) {
            mid = t.b + n / 2;                 // all centroids coincide: split the range in half
        } else {
            float scale = kBins / (cb.hi[best_axis] - cb.lo[best_axis]);
            auto it = std::partition(idx.begin() + t.b, idx.begin() + t.e, [&](int k) {
                int bi = std::min(kBins - 1, (int)((cen[3 * k + best_axis] - cb.lo[best_axis]) * scale));
                return bi <= best_split;
            });
            mid = (int)(it - idx.begin());
            if (mid == t.b || mid == t.e) mid = t.b + n / 2;
        }
        int l = (int)nodes.size();
        nodes.emplace_back();
        int r = (int)nodes.size();
        nodes.emplace_back();
        nodes[t.node].left = l;
        nodes[t.node].right = r;
        work.push_back({l, t.b, mid});
        work.push_back({r, mid, t.e});
    }

    // triangles in leaf order
    out.tris4.resize(ntri);
    for (int i = 0; i < ntri; ++i) {
        int ot = leaf_tri[i];
        v3 p0 = ld3(pos + 9 * ot), p1 = ld3(pos + 9 * ot + 3), p2 = ld3(pos + 9 * ot + 6);
        v3 e1 = p0 - p1, e2 = p0 - p2;
        TriGpu tg;
        float idf;
        std::memcpy(&idf, &ot, 4);
        tg.p0 = make_float4(p0.x, p0.y, p0.z, idf);
        tg.e1 = make_float4(e1.x, e1.y, e1.z, 0.f);
        tg.e2 = make_float4(e2.x, e2.y, e2.z, 0.f);
        out.tris4[i] = tg;
    }

    // 4-wide collapse
    auto area = [&](int b) { return half_area(nodes[b].lo, nodes[b].hi); };
    auto leaf = [&](int b) { return nodes[b].left < 0; };
    auto open4 = [&](int b, int c[4]) {
        int cnt = 2;
        c[0] = nodes[b].left;
        c[1] = nodes[b].right;
        while (cnt < 4) {
            int bi = -1;
            float ba = -1.f;
            for (int i = 0; i < cnt; ++i)
                if (!leaf(c[i]) && area(c[i]) > ba) { ba = area(c[i]); bi = i; }
            if (bi < 0) break;
            int x = c[bi];
            for (int j = cnt; j > bi + 1; --j) c[j] = c[j - 1];
            c[bi] = nodes[x].left;
            c[bi + 1] = nodes[x].right;
            ++cnt;
        }
        return cnt;
    };
    std::vector<int> idx4(nodes.size(), -1), order4;
    std::vector<std::pair<int, int>> st4;
    st4.push_back({0, 1});
    int depth4 = 0;
    while (!st4.empty()) {
        auto [b, d] = st4.back();
        st4.pop_back();
        idx4[b] = (int)order4.size();
        order4.push_back(b);
        depth4 = std::max(depth4, d);
        int c[4];
        int cnt = open4(b, c);
        for (int i = cnt - 1; i >= 0; --i)
            if (!leaf(c[i])) st4.push_back({c[i], d + 1});
    }
    out.nodes4.resize(order4.size());
    for (size_t j = 0; j < order4.size(); ++j) {
        int c[4];
        int cnt = open4(order4[j], c);
        float lo[3][4], hi[3][4];
        int rf[4];
        for (int i = 0; i < 4; ++i) {
            if (i < cnt) {
                for (int a = 0; a < 3; ++a) { lo[a][i] = nodes[c[i]].lo[a]; hi[a][i] = nodes[c[i]].hi[a]; }
                rf[i] = leaf(c[i]) ? ~nodes[c[i]].tri : idx4[c[i]];
            } else {
                for (int a = 0; a < 3; ++a) { lo[a][i] = 0.f; hi[a][i] = 0.f; }
                rf[i] = kEmptyRef;
            }
        }
        Bvh4NodeGpu nd;
        nd.lox = make_float4(lo[0][0], lo[0][1], lo[0][2], lo[0][3]);
        nd.loy = make_float4(lo[1][0], lo[1][1], lo[1][2], lo[1][3]);
        nd.loz = make_float4(lo[2][0], lo[2][1], lo[2][2], lo[2][3]);
        nd.hix = make_float4(hi[0][0], hi[0][1], hi[0][2], hi[0][3]);
        nd.hiy = make_float4(hi[1][0], hi[1][1], hi[1][2], hi[1][3]);
        nd.hiz = make_float4(hi[2][0], hi[2][1], hi[2][2], hi[2][3]);
        nd.ref = make_int4(rf[0], rf[1], rf[2], rf[3]);
        nd.pad = make_int4(0, 0, 0, 0);
        out.nodes4[j] = nd;
    }
    out.root4_ref = 0;
    out.depth4 = depth4;
    out.stack4 = 3 * depth4 + 1;
    out.num_nodes4 = (int)out.nodes4.size();
}

void bvh_build_host(int ntri, const float* pos, BvhHost& out) {
    out = BvhHost();
    out.ntri = ntri;
    if (ntri <= 0) return;
    std::vector<v3> bmin(ntri), bmax(ntri);
    v3 smin = mk(0.f, 0.f, 0.f), smax = mk(0.f, 0.f, 0.f);
    for (int k = 0; k < ntri; ++k) {
        v3 p0 = ld3(pos + 9 * k), p1 = ld3(pos + 9 * k + 3), p2 = ld3(pos + 9 * k + 6);
        bmin[k] = vmin(vmin(p0, p1), p2);
        bmax[k] = vmax(vmax(p0, p1), p2);
        smin = vmin(smin, bmin[k]);
        smax = vmax(smax, bmax[k]);
    }
    st3(out.box, smin);
    st3(out.box + 3, smax);

    std::vector<uint64_t> key(ntri);
    v3 ext = smax - smin;
    for (int k = 0; k < ntri; ++k) {
        v3 c = (bmin[k] + bmax[k]) * 0.5f;
        c = (c - smin) / ext;
        uint32_t code = morton3d(c.x, c.y, c.z);
        key[k] = ((uint64_t)code << 32) | (uint32_t)k;   // (code, index): a stable order
    }
    std::sort(key.begin(), key.end());
    out.sorted_tri.resize(ntri);
    out.sorted_morton.resize(ntri);
    for (int i = 0; i < ntri; ++i) {
        out.sorted_tri[i] = (int)(key[i] & 0xffffffffu);
        out.sorted_morton[i] = (uint32_t)(key[i] >> 32);
    }
    const uint32_t* code = out.sorted_morton.data();

    // BFS node array exactly as aggregate_bvh.cc:118-176 lays it out
    int nn = 2 * ntri - 1;
    std::vector<int> child(2 * nn, -1), range(2 * nn, 0), level(nn, 0);
    range[0] = 0; range[1] = ntri - 1;
    int next = 1, depth = 0;
    for (int i = 0; i < next; ++i) {
        int s = range[2 * i], e = range[2 * i + 1];
        if (s == e) continue;
        int sp = split_range(code, s, e);
        int l = next, r = next + 1;
        next += 2;
        child[2 * i] = l; child[2 * i + 1] = r;
        range[2 * l] = s; range[2 * l + 1] = sp;
        range[2 * r] = sp + 1; range[2 * r + 1] = e;
        level[l] = level[r] = level[i] + 1;
        depth = std::max(depth, level[i] + 1);
    }
    out.depth = depth + 1;

    // bottom-up boxes (children always have larger BFS indices than parents)
    std::vector<float> box(6 * nn);
    for (int i = nn - 1; i >= 0; --i) {
        v3 a, b;
        if (child[2 * i] < 0) {
            int ot = out.sorted_tri[range[2 * i]];
            a = bmin[ot]; b = bmax[ot];
        } else {
            int l = child[2 * i], r = child[2 * i + 1];
            a = vmin(ld3(&box[6 * l]), ld3(&box[6 * r]));
            b = vmax(ld3(&box[6 * l + 3]), ld3(&box[6 * r + 3]));
        }
        st3(&box[6 * i], a);
        st3(&box[6 * i + 3], b);
    }
    std::copy(box.begin(), box.begin() + 6, out.root_box);
    out.bfs_child = child;
    out.bfs_range = range;
    out.bfs_box = box;

    // triangles in sorted order
    out.tris.resize(ntri);
    for (int i = 0; i < ntri; ++i) {
        int ot = out.sorted_tri[i];
        v3 p0 = ld3(pos + 9 * ot), p1 = ld3(pos + 9 * ot + 3), p2 = ld3(pos + 9 * ot + 6);
        v3 e1 = p0 - p1, e2 = p0 - p2;
        TriGpu t;
        float idf;
        std::memcpy(&idf, &ot, 4);
        t.p0 = make_float4(p0.x, p0.y, p0.z, idf);
        t.e1 = make_float4(e1.x, e1.y, e1.z, 0.f);
        t.e2 = make_float4(e2.x, e2.y, e2.z, 0.f);
        out.tris[i] = t;
    }

    // compaction: interior nodes in right-first depth-first preorder
    if (nn == 1) {
        out.root_ref = ~0;
        out.num_interior = 0;
        out.root4_ref = ~0;
        out.depth4 = 0;
        out.stack4 = 1;
        out.num_nodes4 = 0;
        out.tris4 = out.tris;
        return;
    }
    std::vector<int> newidx(nn, -1);
    std::vector<int> order;
    order.reserve(ntri);
    std::vector<int> stk;
    stk.push_back(0);
    while (!stk.empty()) {
        int i = stk.back();
        stk.pop_back();
        if (child[2 * i] < 0) continue;
        newidx[i] = (int)order.size();
        order.push_back(i);
        stk.push_back(child[2 * i]);       // left visited after the right subtree
        stk.push_back(child[2 * i + 1]);
    }
    auto ref = [&](int b) { return child[2 * b] < 0 ? ~range[2 * b] : newidx[b]; };
    out.num_interior = (int)order.size();
    out.nodes.resize(order.size());
    for (size_t j = 0; j < order.size(); ++j) {
        int i = order[j];
        int l = child[2 * i], r = child[2 * i + 1];
        const float* L = &box[6 * l];
        const float* R = &box[6 * r];
        BvhNodeGpu nd;
        nd.a = make_float4(L[0], L[1], L[2], L[3]);
        nd.b = make_float4(L[4], L[5], R[0], R[1]);
        nd.c = make_float4(R[2], R[3], R[4], R[5]);
        nd.d = make_int4(ref(l), ref(r), 0, 0);
        out.nodes[j] = nd;
    }
    out.root_ref = 0;

    build_sah4(ntri, pos, bmin, bmax, out);
}
