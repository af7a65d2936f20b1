// rzk_internal.h — host-side objects behind the C ABI (include/rzk.h).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/rzk.h"
#include "rzk_math.h"

// Device BVH layout.
//   Interior node: 64 B = both child boxes + both child refs
//     a = (L.min.xyz, L.max.x)  b = (L.max.yz, R.min.xy)  c = (R.min.z, R.max.xyz)
//     d = (left_ref, right_ref, 0, 0)
//   ref >= 0 : interior node index;  ref < 0 : leaf holding sorted triangle ~ref
//   Triangle record: 48 B = (p0.xyz, original id) (p0-p1, 0) (p0-p2, 0)
struct BvhNodeGpu { float4 a, b, c; int4 d; };
struct TriGpu { float4 p0, e1, e2; };

// 4-wide node of the same tree collapsed two levels at a time, for any-hit
// (shadow) traversal: 128 B = one L2 line.  Child boxes as SoA float4 lanes,
// refs as above, kEmptyRef for unused slots (always the trailing ones).
struct Bvh4NodeGpu { float4 lox, loy, loz, hix, hiy, hiz; int4 ref; int4 pad; };
constexpr int kEmptyRef = (int)0x80000000;

// Emitting triangle for light sampling: vertices, normal (n.w = area) and the
// cumulative / own area ratio within its light (Quad::sample, quad.cc:75-87)
struct LightTri { float4 p0, p1, p2, n; float cum, ratio; int light, tri; };   // tri: original triangle id
// Area light: radiance, triangle range, selection probability (Scene::sample_light)
struct LightGpu { float le[3]; int first, count; float sel_cdf, sel_pdf, pad; };
constexpr int kLdsStack = 16;   // stack entries per lane kept in LDS; deeper ones spill to HBM

struct BvhHost {
    int ntri = 0;
    int num_interior = 0;
    int depth = 0;            // deepest root-to-leaf path (nodes), bounds the traversal stack
    int root_ref = 0;
    float box[6] = {0, 0, 0, 0, 0, 0};   // scene box (origin included, boundingbox.h:100-103)
    float root_box[6] = {0, 0, 0, 0, 0, 0};
    std::vector<BvhNodeGpu> nodes;
    std::vector<Bvh4NodeGpu> nodes4;
    int root4_ref = 0;
    int depth4 = 0;           // 4-wide nodes on the deepest path
    int stack4 = 0;           // bound on any-hit stack entries (3 per 4-wide level)
    int num_nodes4 = 0;       // 4-wide nodes (nodes4 holds them after a host build)
    bool device_built = false;           // trees built on the device: nodes/tris vectors and the BFS view stay empty
    std::vector<TriGpu> tris;            // reference-tree leaf (Morton) order
    std::vector<TriGpu> tris4;           // 4-wide tree leaf order
    std::vector<int> sorted_tri;         // sorted position -> original triangle
    std::vector<uint32_t> sorted_morton;
    // reference-order (BFS) view for parity checks
    std::vector<int> bfs_child;          // 2 per BFS node
    std::vector<int> bfs_range;          // 2 per BFS node
    std::vector<float> bfs_box;          // 6 per BFS node
};

void bvh_build_host(int ntri, const float* pos, BvhHost& out);

struct rzk_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string last_error;
    // reusable device scratch (grown on demand)
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    int num_cus = 0;
    // held by the creator and by every object created on the context; the
    // stream and scratch go when the last holder lets go, so handles may be
    // destroyed in any order
    std::atomic<int> refs{1};
};

void ctx_retain(rzk_ctx* ctx);
void ctx_release(rzk_ctx* ctx);

struct rzk_scene {
    rzk_ctx* ctx = nullptr;
    BvhHost bvh;
    int nsurf = 0, nmat = 0, nlight = 0;
    // device
    BvhNodeGpu* d_nodes = nullptr;
    Bvh4NodeGpu* d_nodes4 = nullptr;
    TriGpu* d_tris = nullptr;
    TriGpu* d_tris4 = nullptr;
    float4* d_tri_n = nullptr;        // per original triangle: (n.xyz, surface id bits)
    float4* d_tri_tan = nullptr;      // per original triangle: (tangent.xyz, 0)
    int4* d_surf = nullptr;           // per surface: (material, light, flags, 0)
    float* d_mat = nullptr;           // 12 floats per material
    LightTri* d_ltris = nullptr;
    LightGpu* d_lights = nullptr;
    int nltri = 0;
    std::vector<LightTri> h_ltris;
    std::vector<LightGpu> h_lights;
    std::vector<float> h_mat;
    std::vector<int> h_surf_material, h_surf_light, h_surf_flags, h_tri_surface;
    std::vector<float> h_light_radiance;
};

// Builds the scene's two trees on the device from host positions (bvh_device.hip):
// the Morton binary tree node-for-node as bvh_build_host lays it out, and a
// 4-wide tree collapsed from it.  Fills s->bvh's scalars and sorted arrays and
// s->d_nodes, d_nodes4, d_tris, d_tris4, and the per-triangle frames d_tri_n,
// d_tri_tan.  Needs ntri >= 2.
void bvh_build_device(rzk_ctx* ctx, int ntri, const float* pos, const float* tangents, const int* tri_surface,
                      rzk_scene* s);

struct rzk_vpls {
    rzk_ctx* ctx = nullptr;
    int64_t n = 0;
    float* d_vpl = nullptr;           // 21 floats per VPL (include/rzk.h)
    int* d_mat = nullptr;
    int* d_order = nullptr;           // VPL indices in Morton order of their positions
    std::vector<float> h_vpl;
    std::vector<int> h_mat;
};

// Scene constants passed by value to kernels
struct SceneView {
    const BvhNodeGpu* nodes;
    const Bvh4NodeGpu* nodes4;
    const TriGpu* tris;               // leaves of the reference tree (nodes)
    const TriGpu* tris4;              // leaves of the 4-wide tree (nodes4)
    const float4* tri_n;
    const float4* tri_tan;
    const int4* surf;
    const float* mat;
    int root_ref;
    int depth;
    int root4_ref;
    int stack4;
    float rbmin[3], rbmax[3];
    const LightTri* ltris;
    const LightGpu* lights;
    int nlights;
    int* spill;               // any-hit stack overflow, [entry - kLdsStack][spill_pitch]
    int64_t spill_pitch;      // threads of the launch using `spill`
};

SceneView scene_view(const rzk_scene* s);

// error plumbing
struct RzkError { int code; std::string msg; };
#define RZK_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            throw RzkError{RZK_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)}; \
    } while (0)
#define RZK_CHECK(cond, code, msg) \
    do { if (!(cond)) throw RzkError{code, msg}; } while (0)

void* ctx_scratch(rzk_ctx* ctx, size_t bytes);
// resolves the device address of the no-allocation gather's static scratch on
// `device` (current), once per device (gather_noalloc.hip)
void gather_noalloc_setup(int device);
// Page-locked host array (grown on demand, never shrunk): device-to-host
// copies into it run as direct DMA.
template <typename T>
struct PinnedArray {
    T* p = nullptr;
    size_t n = 0;
    PinnedArray() = default;
    PinnedArray(const PinnedArray&) = delete;
    PinnedArray& operator=(const PinnedArray&) = delete;
    void reserve(size_t count) {
        if (count <= n && p) return;
        if (p) RZK_HIP(hipHostFree(p));
        p = nullptr;
        n = 0;
        RZK_HIP(hipHostMalloc((void**)&p, sizeof(T) * (count > 0 ? count : 1), hipHostMallocDefault));
        n = count > 0 ? count : 1;
    }
    ~PinnedArray() { if (p) (void)hipHostFree(p); }
};

// size sv.spill for a launch of `threads` lanes whose traversal stack holds up
// to `entries` ints (0: sv.stack4, the any-hit bound; closest_hit4 needs 2 * stack4)
void bind_spill(rzk_ctx* ctx, SceneView& sv, int64_t threads, int entries = 0, int lds_entries = kLdsStack);

// kernel launchers (traverse.hip, shade.hip, cluster.hip)
void launch_trace_closest(rzk_ctx* ctx, const SceneView& sv, const float* d_rays, int64_t n,
                          float tmin, float tmax, float4* d_hits, int* d_stats);
// closest hits of a width x height row-major image of rays, one wave per 8x8 tile
// (same results as launch_trace_closest)
void launch_trace_closest_image(rzk_ctx* ctx, const SceneView& sv, const float* d_rays, int width, int height,
                                float tmin, float tmax, float4* d_hits, int* d_stats);
void launch_trace_any(rzk_ctx* ctx, const SceneView& sv, const float* d_rays, int64_t n,
                      float tmin, float tmax, uint8_t* d_out, int* d_stats);
void launch_trace_segments(rzk_ctx* ctx, const SceneView& sv, const float* d_p, const float* d_q,
                           int64_t n, float tmin, uint8_t* d_out, int* d_stats);
void launch_fill_receivers(rzk_ctx* ctx, const SceneView& sv, const float* d_rays,
                           const float4* d_hits, int64_t n, float* d_recv, int* d_mat, int* d_light);
void launch_camera_rays(rzk_ctx* ctx, const float* cam, int64_t n, const int* d_px,
                        const int* d_py, int width, float* d_rays);
void launch_camera_image_rays(rzk_ctx* ctx, const float* cam, int width, int height, float* d_rays);
// pair visibility bits of npts receivers (in d_order's order when given) x nvpl VPLs
void launch_visibility(rzk_ctx* ctx, const SceneView& sv, const float* d_recv, const int* d_rmat,
                       int64_t npts, const int* d_order, const float* d_vpl, int64_t nvpl, float tmin,
                       uint64_t* d_bits, int* d_stats);
// the gather's shading sum over nvpl VPLs with those bits; d_acc_in / d_acc_out
// (optional) carry running sums between the chunks of a longer VPL list
void launch_gather_shade(rzk_ctx* ctx, const SceneView& sv, const float* d_recv, const int* d_rmat,
                         int64_t npts, const int* d_order, const float* d_vpl, const int* d_vmat, int64_t nvpl,
                         const uint64_t* d_bits, int clamp_mode, float clamp, float n_emitted,
                         const float* d_acc_in, float* d_acc_out, float* d_out);
void launch_reduced_matrix(rzk_ctx* ctx, const SceneView& sv, const float* d_recv,
                           const int* d_rmat, int rows, const int* d_vorder, const float* d_vpl,
                           const int* d_vmat,
                           int64_t nvpl, int clamp_mode, float clamp, float tmin, float* d_R,
                           float* d_V, int* d_stats, int term = 0);
void launch_alphas(rzk_ctx* ctx, int rows, int64_t cols, const float* d_R, float* d_alphas,
                   float* d_mags);
void launch_cluster_assign(rzk_ctx* ctx, int rows, int64_t cols, const float* d_R, int ncl,
                           const int* d_centers, const float* d_weights, int* d_assign);
// The clusters each slice gathers from: one list shared by every slice
// (ld = 0: nreps representatives), or a list per slice after refinement
// (row s of length ld holds counts[s] representatives).  Slices are tile x
// tile pixel squares, tiles_x per image row.
struct SliceReps {
    const int* reps;
    const float* rep_w;
    const int* counts;
    int nreps;
    int ld;
    int tile, tiles_x;
    const int* pix_slice = nullptr;    // kd-tree slices: each pixel's slice (-1: no gather point)
    __host__ __device__ int width() const { return ld ? ld : nreps; }
    __device__ int slice_of(int64_t pix, int img_width) const {
        if (pix_slice) return pix_slice[pix];
        int px = (int)(pix % img_width), py = (int)(pix / img_width);
        return (py / tile) * tiles_x + px / tile;
    }
    __device__ int count(int s) const { return ld ? counts[s] : nreps; }
    __device__ int rep(int s, int c) const { return reps[(int64_t)s * ld + c]; }
    __device__ float weight(int s, int c) const { return rep_w[(int64_t)s * ld + c]; }
};

void launch_slice_gather(rzk_ctx* ctx, const SceneView& sv, const float* d_recv,
                         const int* d_rmat, int width, int height, const float* d_vpl, const int* d_vmat,
                         const SliceReps& sr, int clamp_mode, float clamp,
                         float tmin, float n_emitted, float* d_out, uint64_t* d_bits, int* d_stats);
