// kdslice.h — the 6-D kd-tree slicing object shared by kdslice.hip and the
// LightSlice driver (lightslice.hip).
#pragma once

#include <algorithm>
#include <vector>

#include "rzk_internal.h"

namespace rzk { class Mt19937; }

// device buffer owned by a slicing
template <typename T>
struct KdBuf {
    T* p = nullptr;
    void alloc(size_t n) {
        if (p) (void)hipFree(p);
        p = nullptr;
        RZK_HIP(hipMalloc((void**)&p, sizeof(T) * std::max<size_t>(n, 1)));
    }
    ~KdBuf() { if (p) (void)hipFree(p); }
};

struct rzk_kdslices {
    rzk_ctx* ctx = nullptr;
    int n = 0, max_size = 0, S = 0, H = 0;
    uint32_t seed = 0;
    KdBuf<float> P, boxes, plane;      // P: 6 planes of n floats
    KdBuf<int> ord, axis, cl, slices;
    std::vector<int> h_slices;         // S x (first, last) kd position
};

// the build over kd->P (n, max_size and seed set); on the context's stream,
// synchronized at the end
void kd_build_device(rzk_kdslices* kd);
// the build with every pivot drawn from rd in the reference's recursion order,
// on the host over h_planes (kd->P's points, 6 planes of n floats); leaves the
// same device arrays as kd_build_device; returns the numbers drawn
int64_t kd_build_host_mt(rzk_kdslices* kd, const float* h_planes, rzk::Mt19937& rd);
// map_vec6 of receiver records idx[i] (i < np) into planes of pitch np
void kd_launch_map_vec6(rzk_ctx* ctx, const float* d_mat, const float* d_recv, const int* d_rmat, const int* d_idx,
                        int np, float diagonal, float* d_planes);
// find_nearest_cluster(k) of the points at kd positions d_pos[0 .. m-1]: d_out[k * i ...]
void kd_launch_nearest_at(const rzk_kdslices* kd, const int* d_pos, int m, int k, int* d_out);
// find_cluster of every point i (< n): d_out[d_dst[i]] = its slice
void kd_launch_find_points(const rzk_kdslices* kd, const int* d_dst, int* d_out);
