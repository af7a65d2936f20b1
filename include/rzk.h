/*
 * rzk.h — C ABI of the MI355X (gfx950) engine for Renzoku's data-parallel hot
 * path: BVH traversal (closest hit, any hit, segment visibility), the clustered
 * VPL gather, the LightSlice reduced matrix / clustering / per-slice guiding
 * distributions.
 *
 * Plain C types only.  Every call returns an RZK_* status; on failure
 * rzk_ctx_last_error() holds the message.  Nothing throws across the ABI.
 * Host pointers are owned by the caller; device buffers by the engine (the
 * *_dev entry points take device pointers owned by the caller).  A context
 * owns one HIP stream and must be used by one thread at a time; distinct
 * contexts may run concurrently.  There is no CPU fallback: every compute
 * entry point runs gfx950 kernels and fails with RZK_E_HIP without a GPU.
 *
 * Record layouts (fp32, little endian):
 *   ray      : 6 floats  {o[3], d[3]}      d as Ray's constructor leaves it (ray.h:17-36)
 *   hit      : 4 x 32 bit {t, beta, gamma, tri (int32, original index, -1 miss)}
 *   receiver : 21 floats {p, n_geometric, n_shading, wo, uvn.u, uvn.v, uvn.n}
 *              + int32 material (-1 none) + int32 light (-1 none)   (local_geometry.h:11-30)
 *   vpl      : 21 floats {p, n, uvn.u, uvn.v, uvn.n, wi, power} + int32 material
 *              (-1: the VPL lies on a light, its BRDF is white; brdf_point_light.h:235-240)
 *   material : 12 floats {type (0 Lambert, 1 Phong, 2 Ward), kd[3], ks[3],
 *              p0 (Phong exponent | Ward ax), p1 (Ward ay), 0, 0, 0}
 *   camera   : 19 floats {eye[3], u[3], v[3], n[3], cop[2], scale[2], focal, width, height}
 *              (camera.cc:52-65, sensor.h:176-182)
 */
#ifndef RZK_H
#define RZK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RZK_OK         0
#define RZK_E_INVALID  1
#define RZK_E_OOM      2
#define RZK_E_HIP      3
#define RZK_E_STATE    4

#define RZK_RAY_FLOATS      6
#define RZK_RECV_FLOATS     21
#define RZK_VPL_FLOATS      21
#define RZK_MAT_FLOATS      12
#define RZK_CAM_FLOATS      19

/* clamp modes: VplRadianceClamp (vpl.cc:214-230) */
#define RZK_CLAMP_NONE          0
#define RZK_CLAMP_FORM_FACTOR   1
#define RZK_CLAMP_REFLECTIVITY  2
#define RZK_CLAMP_RADIANCE      3

/* surface flags (surface.h:109-110) */
#define RZK_SURF_FLIPPED    1
#define RZK_SURF_TWO_SIDED  2

typedef struct rzk_ctx rzk_ctx;
typedef struct rzk_scene rzk_scene;
typedef struct rzk_vpls rzk_vpls;
typedef struct rzk_lightslice rzk_lightslice;

typedef struct rzk_hit { float t, beta, gamma; int32_t tri; } rzk_hit;

typedef struct rzk_scene_desc {
    int32_t num_triangles;
    const float *positions;        /* [num_triangles][9] p0,p1,p2 in surface order (aggregate_bvh.cc:60-75) */
    const float *tangents;         /* [num_triangles][3] Triangle::u / Quad shared tangent (quad.cc:22-24) */
    const int32_t *tri_surface;    /* [num_triangles] */
    int32_t num_surfaces;
    const int32_t *surf_material;  /* [num_surfaces] material index or -1 */
    const int32_t *surf_light;     /* [num_surfaces] light index or -1 */
    const int32_t *surf_flags;     /* [num_surfaces] RZK_SURF_* */
    int32_t num_materials;
    const float *materials;        /* [num_materials][12] */
    int32_t num_lights;
    const float *light_radiance;   /* [num_lights][3] */
} rzk_scene_desc;

typedef struct rzk_scene_info {
    int32_t num_triangles;
    int32_t num_interior_nodes;
    int32_t depth;                 /* nodes on the deepest root-to-leaf path */
    float box[6];                  /* scene box, origin included (boundingbox.h:100-103) */
    int64_t device_bytes;          /* nodes + triangles + per-triangle shading data */
} rzk_scene_info;

/* ---- context ---------------------------------------------------------- */
int rzk_ctx_create(int device, rzk_ctx **out);
int rzk_ctx_destroy(rzk_ctx *ctx);
const char *rzk_ctx_last_error(const rzk_ctx *ctx);
void *rzk_ctx_stream(rzk_ctx *ctx);                  /* the hipStream_t every launch uses */
int rzk_ctx_synchronize(rzk_ctx *ctx);
int rzk_device_alloc(rzk_ctx *ctx, int64_t bytes, void **out);
int rzk_device_free(rzk_ctx *ctx, void *ptr);
int rzk_copy_h2d(rzk_ctx *ctx, void *dst, const void *src, int64_t bytes);
int rzk_copy_d2h(rzk_ctx *ctx, void *dst, const void *src, int64_t bytes);

/* ---- Aggregate (aggregate.h:15-91, aggregate_bvh.h:20-47) -------------- */
/* AggregateBvh::AggregateBvh + build (aggregate_bvh.cc:16-240): Morton LBVH, one
 * triangle per leaf, identical topology to the reference; uploaded as a
 * compacted BVH2 (64 B interior nodes, 48 B triangle records). */
int rzk_scene_create(rzk_ctx *ctx, const rzk_scene_desc *desc, rzk_scene **out);
/* The same scene with its trees built on the device (the reference builds on
 * the host: AggregateBvh::build, aggregate_bvh.cc:48-240).  With
 * RZK_SCENE_BUILD_DEVICE the Morton binary tree is the one rzk_scene_create
 * makes, node for node and bit for bit (Morton codes, a stable radix sort,
 * top-down splits and bottom-up boxes, one depth per launch); the 4-wide tree
 * the any-hit, gather and path-tracing walks use is that binary tree collapsed
 * four-wide instead of the host's binned-SAH tree.  Any-hit and segment answers
 * are the same; closest hits are the same wherever the nearest triangle is
 * unique.  flags = 0 is rzk_scene_create.  Scenes of fewer than two triangles
 * are built on the host. */
#define RZK_SCENE_BUILD_DEVICE 1
int rzk_scene_create_ex(rzk_ctx *ctx, const rzk_scene_desc *desc, int flags, rzk_scene **out);
int rzk_scene_destroy(rzk_scene *scene);
int rzk_scene_get_info(const rzk_scene *scene, rzk_scene_info *out);
/* the reference's BFS node array (aggregate_bvh.cc:118-207) for parity checks */
int rzk_scene_export_bvh(const rzk_scene *scene, int32_t *sorted_tri, uint32_t *sorted_morton,
                         int32_t *node_child, int32_t *node_range, float *node_box,
                         int32_t *num_nodes);  /* RZK_E_STATE for node arrays of a device-built scene */
/* the device trees as the walks read them: binary nodes (64 B each, num_interior_nodes),
 * Morton-order triangles (48 B each), 4-wide nodes (128 B each, *num_nodes4) and the
 * 4-wide tree's triangles; any pointer may be NULL.  No reference counterpart: parity
 * checks of the two builds. */
int rzk_scene_export_trees(const rzk_scene *scene, void *nodes, void *tris, void *nodes4, void *tris4,
                           int32_t *num_nodes4);

/* Aggregate::hit(const Ray&, tmin, tmax, time, HitRecord&) — closest hit,
 * aggregate_bvh.cc:402-448.  Batched, host buffers. */
int rzk_trace_closest(rzk_ctx *ctx, const rzk_scene *scene, const float *rays, int64_t n,
                      float tmin, float tmax, rzk_hit *out);
/* Aggregate::hit(const Ray&, tmin, tmax, time) — any hit, aggregate_bvh.cc:452-478 */
int rzk_trace_any(rzk_ctx *ctx, const rzk_scene *scene, const float *rays, int64_t n,
                  float tmin, float tmax, uint8_t *out_hit);
/* Aggregate::hit(const Vec3 &p, const Vec3 &q, tmin, time) — segment
 * visibility, true = blocked (aggregate.cc:84-90) */
int rzk_trace_segments(rzk_ctx *ctx, const rzk_scene *scene, const float *p, const float *q,
                       int64_t n, float tmin, uint8_t *out_blocked);
/* device-resident forms: pointers are device memory, launches are queued on
 * the context stream and the call returns without synchronising.  d_stats
 * (optional, device uint64[2]) has {interior nodes visited, triangles tested}
 * added to it. */
int rzk_trace_closest_dev(rzk_ctx *ctx, const rzk_scene *scene, const float *d_rays, int64_t n,
                          float tmin, float tmax, rzk_hit *d_out, uint64_t *d_stats);
int rzk_trace_any_dev(rzk_ctx *ctx, const rzk_scene *scene, const float *d_rays, int64_t n,
                      float tmin, float tmax, uint8_t *d_out, uint64_t *d_stats);
int rzk_trace_segments_dev(rzk_ctx *ctx, const rzk_scene *scene, const float *d_p,
                           const float *d_q, int64_t n, float tmin, uint8_t *d_out,
                           uint64_t *d_stats);
/* per-ray traversal counts {interior nodes visited, triangles tested}, host
 * buffers; kind 0 = closest, 1 = any (counted on the binary tree in the
 * reference's visit order, so they equal aggregate_bvh.cc's loop counts) */
int rzk_trace_stats(rzk_ctx *ctx, const rzk_scene *scene, int kind, const float *rays,
                    int64_t n, float tmin, float tmax, int32_t *out_stats);

/* Triangle::fill_hit_record + Surface::fill_hit_record + Receiver(r, rec)
 * (triangle.cc:277-296, surface.cc:62-96, local_geometry.h:21-30) */
int rzk_fill_receivers(rzk_ctx *ctx, const rzk_scene *scene, const float *rays,
                       const rzk_hit *hits, int64_t n, float *recv, int32_t *recv_mat,
                       int32_t *recv_light);
/* Camera::shoot_ray(i, j) for pixel-centre rays (camera.cc:52-65); px = column
 * j, py = row i.  Writes rays, closest hits and receivers (each output optional). */
int rzk_camera_primary(rzk_ctx *ctx, const rzk_scene *scene, const float *camera, int64_t n,
                       const int32_t *px, const int32_t *py, float *rays, rzk_hit *hits,
                       float *recv, int32_t *recv_mat, int32_t *recv_light);

/* ---- VPL set (brdf_point_light.h) ------------------------------------- */
int rzk_vpls_create(rzk_ctx *ctx, const float *vpl, const int32_t *vpl_mat, int64_t n,
                    rzk_vpls **out);
int rzk_vpls_destroy(rzk_vpls *vpls);

/* VplGenerator::spawn with max_bounce = 2 over area lights (vpl_generator.cc:127-181;
 * SubpathGenerator::sample_direct_vertex, path.cc:171-220, 305-334; AreaLight::sample,
 * area_light.cc:26-44): light samples 0, 1, 2, ... each pick a light (Scene::sample_light),
 * a triangle of it by area and a uniform point, and a cosine-weighted direction in the light
 * surface's basis; the closest hit of a sample on a surface with a material (not a light)
 * becomes a bounce-1 VPL (RZK_VPL_FLOATS layout: p, shading normal, uvn, wi = -ray direction,
 * power = Le |n_q . wo| / (pdf_light pdf_q pdf_wo)).  Samples are taken in order until num_vpls
 * VPLs are stored (at most max_samples samples, else RZK_E_STATE).  Each sample's numbers come
 * from Philox4x32-10 keyed by seed with the sample index as counter (the reference draws them
 * from its global MT stream).  out_samples: light samples used (the gather's
 * num_emitted_photons); out_work (optional, 3 values): closest rays traced (batches may trace
 * a few samples past the last one used), interior nodes visited, triangles tested.  Device
 * buffers come from the context's scratch. */
int rzk_generate_vpls(rzk_ctx *ctx, const rzk_scene *scene, int64_t num_vpls, uint32_t seed,
                      int64_t max_samples, float tmin, float *out_vpl, int32_t *out_mat,
                      int64_t *out_samples, int64_t *out_work);

/* ---- 6-D kd-tree slices (LightSlice::generate_slices, lightslice.cc:252-303) --
 * rzk_map_vec6: map_vec6 (lightslice.cc:142-188) of n receiver records with
 * their material ids (-1: none): position, then the directionality (shading
 * normal; for Phong / Ward the mirror direction of wo scaled by
 * min(4, max(1, gloss exponent / 1000))) times diagonal / 16.  out: n x 6. */
int rzk_map_vec6(rzk_ctx *ctx, const rzk_scene *scene, const float *recv, const int32_t *mat, int64_t n,
                 float diagonal, float *out);

/* KdTreeN(max_cluster_size, ...) (kdtree_n.h:399-429, cluster_kdtree :285-365,
 * quick_select_k :181-242): a kd-tree over n points (n x 6, host) split at the
 * median of the longest axis; the first node of each path with at most
 * max_slice_size (>= 3) points is a slice, numbered depth-first.  Round t of
 * the median selection at node [s, e] takes its pivot from Philox4x32-10 keyed
 * by seed with counter (s, e, t) (the reference draws from its global MT
 * stream).  Arrays (rzk_kdslices_get, bytes must match): "order" (n i32: the
 * point at each kd position), "slices" (S x 2 i32: first, last kd position),
 * "boxes" (S x 12 f32: min[6], max[6]), and the query tree as a binary heap of
 * heap_used nodes: "axis" (i32), "plane" (f32), "cl" (i32 slice of a slice
 * node, else -1). */
typedef struct rzk_kdslices rzk_kdslices;
int rzk_kdslices_build(rzk_ctx *ctx, const float *points, int32_t n, int32_t max_slice_size, uint32_t seed,
                       rzk_kdslices **out);
/* the same with every pivot drawn from one MT19937 stream seeded with seed, in
 * cluster_kdtree's depth-first recursion order, as the reference draws them from
 * the scene's Random (kdtree_n.h:192, :399-429).  The draw count of each node
 * depends on its points, so the build is sequential and runs on the host; the
 * tree lands on the device in the same layout.  *out_draws (may be NULL) = the
 * numbers drawn; the representatives' draws (lightslice.cc:293-301) come next. */
int rzk_kdslices_build_mt(rzk_ctx *ctx, const float *points, int32_t n, int32_t max_slice_size, uint32_t seed,
                          int64_t *out_draws, rzk_kdslices **out);
/* the same as rzk_kdslices_build with the points already in device memory (n x 6, on the context's stream) */
int rzk_kdslices_build_dev(rzk_ctx *ctx, const float *d_points, int32_t n, int32_t max_slice_size, uint32_t seed,
                           rzk_kdslices **out);
int rzk_kdslices_destroy(rzk_kdslices *kd);
int rzk_kdslices_info(const rzk_kdslices *kd, int32_t *out_slices, int32_t *out_heap_used);
int rzk_kdslices_get(const rzk_kdslices *kd, const char *name, void *dst, int64_t bytes);
/* find_cluster (kdtree_n.h:540-572): the slice of each of m query points
 * (m x 6), descending by plane - 1e-3 (PLANE_BIAS, kdtree_n.h:171) */
int rzk_kdslices_find(const rzk_kdslices *kd, const float *queries, int32_t m, int32_t *out);
/* find_nearest_cluster (kdtree_n.h:575-647): the k (1..32) slices whose boxes
 * are nearest each query, in std::priority_queue pop order (furthest first),
 * padded with -1.  out: m x k. */
int rzk_kdslices_nearest(const rzk_kdslices *kd, const float *queries, int32_t m, int32_t k, int32_t *out);

/* ---- gather (vpl.cc:661-688 gather_point, summed over the VPL set as
 * lightslice.cc:1325-1339 does over cluster representatives of weight 1,
 * then / num_emitted as vpl.cc:488-489) ----------------------------------- */
int rzk_gather(rzk_ctx *ctx, const rzk_scene *scene, const rzk_vpls *vpls, const float *recv,
               const int32_t *recv_mat, int64_t npts, int clamp_mode, float clamp, float tmin,
               float n_emitted, float *out_rgb);
/* No counterpart in the reference: the rzk_gather above for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Receivers (d_recv, d_recv_mat) and the
 * output (d_out_rgb, npts x 3) are device memory owned by the caller; launches
 * are queued on the context stream and the call returns without synchronising.
 * The visibility bits live in a static 32 MiB device array per GPU, so
 * ceil(npts / 64) x VPLs must be at most 4,194,304 (else RZK_E_INVALID).  Every
 * call on one GPU shares that array: calls that overlap on one GPU (other
 * streams or contexts) race; calls queued on one stream do not. */
int rzk_gather_dev_noalloc(rzk_ctx *ctx, const rzk_scene *scene, const rzk_vpls *vpls, const float *d_recv,
                           const int32_t *d_recv_mat, int64_t npts, int clamp_mode, float clamp, float tmin,
                           float n_emitted, float *d_out_rgb);
/* visibility of every (receiver, VPL) pair as a bit matrix: word (k, i/64) bit
 * i%64 set when the segment receiver i -> VPL k is blocked (vpl.cc:672). */
int rzk_visibility(rzk_ctx *ctx, const rzk_scene *scene, const rzk_vpls *vpls, const float *recv,
                   const int32_t *recv_mat, int64_t npts, float tmin, uint64_t *out_bits);
/* The plain VPL render of a whole image, everything device resident: primary
 * closest hits at pixel centres, then the full gather.  out_rgb: host
 * [height][width][3] or NULL (keep on device). */
int rzk_render_vpl(rzk_ctx *ctx, const rzk_scene *scene, const rzk_vpls *vpls,
                   const float *camera, int clamp_mode, float clamp, float tmin,
                   float n_emitted, float *out_rgb);

/* ---- LightTransportMatrix / ClusterBuilder (reduced matrix and global
 * clustering) ------------------------------------------------------------- */
/* LightSlice reduced matrix, lightslice.cc:721-748: R(r,k) = vis * incident
 * radiance (vpl.cc:411-440), V(r,k) = vis.  R column-major Rgb [(k*rows+r)*3],
 * V [k*rows+r]; rows whose receiver has no material stay zero. */
int rzk_reduced_matrix(rzk_ctx *ctx, const rzk_scene *scene, const rzk_vpls *vpls,
                       const float *recv, const int32_t *recv_mat, int32_t rows, int clamp_mode,
                       float clamp, float tmin, float *out_R, float *out_V);
/* The VPL evaluator over receiver x VPL pairs (IVirtualPointLightEvaluator,
 * vpl_interface.h:126-139, as VirtualPointLight implements it):
 *  term RZK_TERM_RADIANCE: radiance(scene, r, vpl, visible), vpl.cc:405-409 ->
 *    gather_point :661-688: black when VPL k is blocked from receiver r, else the
 *    clamped outgoing radiance (power x receiver BRDF x light BRDF x G) toward the
 *    receiver's wo: the entries LightTransportMatrix::fill stores
 *    (light_transport_matrix.cc:16-50);
 *  term RZK_TERM_INCIDENT: incident_radiance(scene, r, vpl), vpl.cc:411-440: the
 *    clamped incident radiance (geometric normal, white receiver BRDF), no
 *    visibility.
 * M(r,k) column-major Rgb [(k*rows+r)*3], V(r,k) = visibility [k*rows+r]; rows
 * whose recv_mat is negative (no material, or rows the caller skips, such as
 * receivers on a light) stay zero. */
#define RZK_TERM_RADIANCE 0
#define RZK_TERM_INCIDENT 1
int rzk_transport_matrix(rzk_ctx *ctx, const rzk_scene *scene, const rzk_vpls *vpls,
                         const float *recv, const int32_t *recv_mat, int32_t rows, int term, int clamp_mode,
                         float clamp, float tmin, float *out_M, float *out_V);
/* generate_alphas_for_clustering, light_transport_matrix.cc:72-106 (sequential
 * summation orders kept, bit-exact); mags = cluster_builder.cc:14-19 magnitude
 * of each flattened column (optional). */
int rzk_alphas(rzk_ctx *ctx, int32_t rows, int64_t cols, const float *R, float *out_alphas,
               float *out_mags);
/* assignment of every column to the nearest weighted centre column,
 * cluster_builder.cc:158-225, strict < with the first minimum kept. */
int rzk_cluster_assign(rzk_ctx *ctx, int32_t rows, int64_t cols, const float *R, int32_t ncl,
                       const int32_t *centers, const float *weights, int32_t *out_assign);
/* LightSlice per-slice refinement (lightslice.cc:837-902, split_cluster
 * :405-488, compute_distance :336-357, choose_representative :385-400): the
 * ncl kept clusters (members cl_mem[cl_off[c] .. cl_off[c+1]) in order,
 * representatives cl_rep, weights cl_w) are split per slice, most expensive
 * first, until a slice holds max_slice_clusters clusters.  R: column-major Rgb
 * (R[(k*S + r)*3 + ch], V columns, S rows); nb: S x nnb neighbour rows (the
 * slice matrix).  draws: the MT numbers that follow the clustering's in the
 * reference's stream, at least 4 * (max_slice_clusters - ncl) * S of them.
 * Outputs: out_count[S]; out_rep, out_w: S x max_slice_clusters, row s holding
 * slice s's clusters in the reference's order. */
int rzk_refine_slices(rzk_ctx *ctx, int32_t S, int64_t V, const float *R, int32_t nnb, const int32_t *nb,
                      int32_t ncl, const int32_t *cl_off, const int32_t *cl_mem, const int32_t *cl_rep,
                      const float *cl_w, int32_t max_slice_clusters, const float *draws, int64_t ndraws,
                      int32_t *out_count, int32_t *out_rep, float *out_w);

/* ---- LightSlice (lightslice.h, lightslice.cc:589-1345) ------------------ */
typedef struct rzk_lightslice_params {
    int32_t width, height;          /* image */
    int32_t tile;                   /* slice = tile x tile pixels (32) */
    int32_t num_clusters;           /* ClusterBuilder centres (300) */
    int32_t num_neighbor_slices;    /* rows used for the per-slice cluster pdf (16) */
    int32_t map_resolution;         /* luminance map resolution (32, lightslice.cc:57) */
    float map_epsilon;              /* luminance_map_epsilon (1e-3, lightslice.cc:74) */
    int32_t clamp_mode;
    float clamp;                    /* gather clamp (clamp_threshold) */
    float incident_clamp;           /* reduced matrix clamp (incident_radiance_clamp_threshold) */
    float tmin;
    float n_emitted;                /* num_emitted_photons */
    uint32_t seed;                  /* MT19937 seed for centre / representative sampling */
    int32_t map_gate;               /* 1: build maps only where sampling_records[s].is_valid
                                       would be true (lightslice.cc:2070); 0: every
                                       slice with a non-empty map */
    int32_t max_slice_clusters;     /* per-slice refinement bound (lightslice.cc:29, :869):
                                       clusters are split per slice until a slice holds this
                                       many; 0 or <= the kept clusters: no refinement */
    int32_t kd_max_slice_size;      /* 0: slices are pixel tiles; >= 3: the reference's
                                       slices (generate_slices, lightslice.cc:252-303): a 6-D
                                       kd-tree over the gather points, slices of at most this
                                       many points (max_slice_size = 256, lightslice.cc:27),
                                       a random representative per slice and its nearest
                                       slices as neighbours; tile is then unused */
    int32_t kd_pivots;              /* kd-tree pivot stream: 0 = round t of node [s, e] from
                                       Philox keyed by (seed, s, e, t), every node of a depth at
                                       once on the device; 1 = the reference's: every pivot the
                                       next number of the LightSlice's MT stream in depth-first
                                       recursion order (kdtree_n.h:192), before the
                                       representatives' draws; a sequential host build */
} rzk_lightslice_params;

typedef struct rzk_lightslice_info {
    int32_t num_slices;
    int32_t num_centers;            /* sampled centres (clusters before removal) */
    int32_t num_clusters;           /* clusters kept (with a representative) */
    int32_t draws;                  /* centre draws taken */
    int64_t reduced_pairs;          /* rows x VPLs */
    int64_t gather_pairs;           /* pixels with a material x clusters */
    int32_t refine_passes;          /* passes of the per-slice refinement (0: no refinement); a pass
                                       runs slices at guessed stream offsets and keeps those on the
                                       chain of exact offsets (1 when every slice took 4 numbers per
                                       split) */
} rzk_lightslice_info;

/* the LightSlice keeps scene and vpls by pointer: both must outlive it */
int rzk_lightslice_create(rzk_ctx *ctx, const rzk_scene *scene, const rzk_vpls *vpls,
                          const rzk_lightslice_params *params, rzk_lightslice **out);
int rzk_lightslice_destroy(rzk_lightslice *ls);
/* LightSlice::initialize with pixel-tile slices: primary hits for every pixel,
 * reduced matrix on the tile-centre rows, alphas, centre sampling, assignment,
 * representatives, per-slice cluster pdf + luminance map. */
int rzk_lightslice_initialize(rzk_lightslice *ls, const float *camera);
/* LightSlice::gather_clusters for every pixel (lightslice.cc:1315-1345) / n_emitted.
 * out_rgb host [height][width][3] or NULL to keep the image on the device. */
int rzk_lightslice_gather(rzk_lightslice *ls, float *out_rgb);
int rzk_lightslice_get_info(const rzk_lightslice *ls, rzk_lightslice_info *out);
/* Algorithmic work of one step, measured by re-running the reduced matrix and
 * the gather with traversal counters (results unchanged): out[8] = {primary
 * rays, reduced pairs, reduced nodes visited, reduced triangles tested, gather
 * pairs, gather nodes, gather triangles, clusters}. */
int rzk_lightslice_count_work(rzk_lightslice *ls, int64_t *out);
/* copy an intermediate array to host: name is one of
 *  "R" (rows*V*3 f32), "V" (rows*V f32), "alphas" (V f32), "mags" (V f32),
 *  "centers" (i32), "center_w" (f32), "assign" (V i32), "reps" (i32 per kept cluster),
 *  "rep_w" (f32), "rows_recv" (rows*21 f32), "rows_mat" (rows i32),
 *  "slice_pdf" (slices*clusters f32), "slice_upper" (slices*clusters i32),
 *  "slice_wi" (slices*clusters*3 f32), "maps" (slices*res*res f32), "image" (W*H*3 f32),
 *  "recv" (W*H*21 f32), "recv_mat" (W*H i32), "hits" (W*H x {t, beta, gamma, tri as i32 bits}),
 *  "slice_count" (slices i32), "slice_reps" (slices*C i32), "slice_rep_w" (slices*C f32): each
 *  slice's clusters after refinement (the kept clusters when there is none).  C, the row length
 *  of every slice_* array, is max_slice_clusters when refinement ran, else the kept clusters;
 *  "pt_work" (4 i64): the last rzk_lightslice_render_guided that counted rays: {closest rays,
 *  shadow rays, 4-wide nodes visited, triangles tested}, each lane's own traversal.
 * bytes must equal the array size. */
int rzk_lightslice_get(const rzk_lightslice *ls, const char *name, void *dst, int64_t bytes);

/* ---- guided path tracing (VirtualPointLight::trace_wellington, vpl.cc:934-1041,
 * METRO_RANDOM; next-event estimation FirstBounce/AreaLight::first_bounce,
 * area_light.cc:48-142) ---------------------------------------------------- */
typedef struct rzk_pt_params {
    int32_t spp;                    /* samples per pixel (a square count is stratified) */
    int32_t max_bounce;             /* path depth (trace_wellington max_bounce) */
    int32_t direct_lighting;        /* Integrator::is_direct_lighting */
    int32_t guide_mode;             /* 0: uniform-hemisphere direction sampler everywhere (what the
                                       reference's gated sampler does, lightslice.cc:1654-1665);
                                       1: the slice's luminance map at the first bounce */
    uint32_t seed;                  /* Philox key: results depend only on (seed, pixel, sample) */
    float tmin;
    int32_t x0, y0, w, h;           /* window to render; w = h = 0 for the whole image */
} rzk_pt_params;

/* Guided path tracing of the LightSlice's image with its per-slice guiding
 * distributions (rzk_lightslice_initialize must have run).  out_rgb: host
 * [height][width][3] (pixels outside the window are 0) or NULL to keep the
 * image on the device; out_rays (optional, 2 x int64): closest-hit rays and
 * shadow rays traced. */
int rzk_lightslice_render_guided(rzk_lightslice *ls, const float *camera, const rzk_pt_params *params,
                                 float *out_rgb, int64_t *out_rays);

/* ---- guides and batched rendering -------------------------------------
 * A guide is the per-slice guiding data of one camera (the direction sampler
 * LightSlice implements for DirectionSampler, local_geometry.h:73-79).  A
 * batch renders many jobs (windows of guided path tracing, each with its own
 * guide, spp, depth and seed) with one persistent kernel fed from a work
 * queue, launched as one captured HIP graph.  Every job's image is bit-equal
 * to rzk_lightslice_render_guided of the same window: samples are summed per
 * pixel in sample order whatever the batch holds. */
typedef struct rzk_guide rzk_guide;
typedef struct rzk_batch rzk_batch;

/* borrow the guiding data of an initialised LightSlice for `camera` (the
 * camera it was initialised with); valid until the LightSlice is destroyed or
 * initialised again */
int rzk_lightslice_guide(rzk_lightslice *ls, const float *camera, rzk_guide **out);
int rzk_guide_destroy(rzk_guide *guide);

typedef struct rzk_render_job {
    const rzk_guide *guide;         /* camera, image size and slices of the job */
    int32_t x0, y0, w, h;           /* window of the guide's image */
    int32_t spp, max_bounce, direct_lighting, guide_mode;
    uint32_t seed;
    float tmin;
} rzk_render_job;

typedef struct rzk_batch_config {
    int32_t block;                  /* workgroup size: 64, 128, 192 or 256 lanes */
    int32_t blocks_per_cu;          /* grid = compute units x blocks_per_cu */
    int32_t items_per_dequeue;      /* persistent engine: work items a wave takes per atomic */
    int32_t samples_per_item;       /* persistent engine: samples per work item (64 pixels) */
    int32_t engine;                 /* 0: persistent one-lane-per-pixel kernel over a work queue;
                                       1: wavefront engine (paths in slots, compacted ray queues) */
    int32_t path_slots;             /* wavefront engine: paths in flight (queue depth) */
} rzk_batch_config;

/* cfg NULL: defaults {256, 8, 1, 1, 1, 1048576} */
int rzk_batch_create(rzk_ctx *ctx, const rzk_scene *scene, const rzk_batch_config *cfg, rzk_batch **out);
int rzk_batch_destroy(rzk_batch *batch);
int rzk_batch_set_config(rzk_batch *batch, const rzk_batch_config *cfg);
int rzk_batch_get_config(const rzk_batch *batch, rzk_batch_config *out);
/* render n jobs in one graph launch.  out_rgb: host, the jobs' [h][w][3]
 * images concatenated in job order, or NULL to keep them on the device;
 * out_rays (optional, 2 x int64): closest-hit and shadow rays traced;
 * out_ms (optional): device time of the launch.  Synchronous. */
int rzk_batch_render(rzk_batch *batch, const rzk_render_job *jobs, int32_t n, float *out_rgb,
                     int64_t *out_rays, float *out_ms);
/* time candidate configurations on `jobs` (the engine, then coordinate
 * descent over its launch shape: block, blocks_per_cu and path_slots for the
 * wavefront engine; samples_per_item, block, blocks_per_cu and
 * items_per_dequeue for the persistent one), keep the fastest; each candidate
 * is timed once (the fastest of three renders after a warm-up); tried
 * (optional, max_tried entries) receives each candidate with its time in ms
 * in rzk_batch_trial records; returns the number tried in *n_tried. */
typedef struct rzk_batch_trial { rzk_batch_config cfg; float ms; } rzk_batch_trial;
int rzk_batch_autotune(rzk_batch *batch, const rzk_render_job *jobs, int32_t n, rzk_batch_config *best,
                       rzk_batch_trial *tried, int32_t max_tried, int32_t *n_tried);

/* ---- served form: resident clusters, per-camera guides, request loop ----
 * The clusters (representatives and weights of LightSlice's global
 * clustering) stay resident; a guide for any camera is then built from them
 * alone: tile-centre primary hits, the reduced matrix restricted to the
 * representatives (rows x clusters segments), cluster directions and the
 * per-slice luminance maps (lightslice.cc:1011-1030, 2058-2151).  For the
 * LightSlice's own camera the guide equals the LightSlice's maps bit for bit. */
typedef struct rzk_clusters rzk_clusters;
typedef struct rzk_server rzk_server;

typedef struct rzk_guide_params {
    int32_t tile;                   /* slice = tile x tile pixels */
    int32_t map_resolution;         /* 32 (lightslice.cc:57) */
    float map_epsilon;              /* 1e-3 (lightslice.cc:74) */
    int32_t clamp_mode;             /* incident radiance clamp of the reduced matrix */
    float incident_clamp;
    float tmin;
    int32_t map_gate;               /* as rzk_lightslice_params.map_gate */
} rzk_guide_params;

/* keep the LightSlice's kept clusters (representative VPLs, weights) and its
 * guide parameters as a resident set owned by the returned handle; it refers
 * to the LightSlice's scene, which must outlive it (the LightSlice need not) */
int rzk_lightslice_clusters(const rzk_lightslice *ls, rzk_clusters **out);
int rzk_clusters_destroy(rzk_clusters *clusters);
/* params NULL: the parameters of the LightSlice the clusters came from */
int rzk_guide_create(rzk_ctx *ctx, const rzk_scene *scene, const rzk_clusters *clusters, const float *camera,
                     const rzk_guide_params *params, rzk_guide **out);
/* copy a guide table to host: "maps", "col_cdf" (slices*res*res f32),
 * "row_cdf" (slices*res f32), "totals" (slices f32); bytes must match */
int rzk_guide_get(const rzk_guide *guide, const char *name, void *dst, int64_t bytes);
/* the plain per-request drop-in over any guide: rzk_lightslice_render_guided
 * with the guide's camera (window, spp, depth, seed from params) */
int rzk_render_guided(rzk_ctx *ctx, const rzk_guide *guide, const rzk_pt_params *params, float *out_rgb,
                      int64_t *out_rays);

/* MonteCarloIntegrator::radiance(PrimaryRay&) for guided path tracing
 * (monte_carlo_integrator.cc:40-43, vpl.cc:443-461 with trace_wellington,
 * vpl.cc:934-1041) over a batch of given primary rays, the rays of a
 * PrimaryRayBundle (sensor.h:124-149): ray i (a ray record) is sample smp[i] of
 * pixel (pix[2i] = x, pix[2i+1] = y) of the guide's image; out_rgb[3i..] gets
 * its radiance, unaveraged.  Of params only max_bounce, direct_lighting,
 * guide_mode, seed and tmin are read; bounce b of ray i draws from
 * Philox(seed, pixel, smp[i], b) as rzk_render_guided does. */
int rzk_render_guided_rays(rzk_ctx *ctx, const rzk_guide *guide, const rzk_pt_params *params, const float *rays,
                           const int32_t *pix, const int32_t *smp, int64_t n, float *out_rgb);

typedef struct rzk_server_params {
    int64_t memory_budget;          /* bytes of device memory for admitted requests (guide tables,
                                       per-sample radiance, image) */
    int32_t step_iterations;        /* engine step between looks at finished requests (1..32): wavefront
                                       iterations, or work items per wave of the grid for the
                                       persistent kernel */
    rzk_batch_config launch;        /* engine (0: the persistent path kernel, items of 64 pixels x one
                                       sample; 1: the wavefront engine and its path slots) and its
                                       launch shape */
    int32_t use_cluster_params;     /* 1: guide parameters of the clusters' LightSlice, else `guide` */
    rzk_guide_params guide;
} rzk_server_params;

typedef struct rzk_request {
    float camera[19];               /* camera record; width and height give the image size */
    int32_t spp, max_bounce, direct_lighting, guide_mode;
    uint32_t seed;
    float tmin;
} rzk_request;

typedef struct rzk_request_stats {
    double submitted, admitted, finished;   /* seconds on the server's clock */
    int32_t rounds;                 /* engine steps the request was open for */
    int32_t pad;
    int64_t device_bytes;           /* memory the request held while admitted */
} rzk_request_stats;

typedef struct rzk_server_stats {
    int64_t submitted, finished, rounds, items;   /* rounds: engine steps; items: paths finished */
    int64_t peak_bytes;             /* most device memory admitted requests held at once */
    double busy_seconds;            /* summed device time of the round launches */
    double now;                     /* the server clock */
} rzk_server_stats;

/* a server owns a context (its own stream) and a worker thread on `device`;
 * scene and clusters must outlive it */
int rzk_server_create(int device, const rzk_scene *scene, const rzk_clusters *clusters,
                      const rzk_server_params *params, rzk_server **out);
int rzk_server_destroy(rzk_server *server);
/* thread-safe; the request is queued at once and admitted when the memory
 * budget allows, oldest first */
int rzk_server_submit(rzk_server *server, const rzk_request *request, int64_t *ticket);
/* block until the request is finished (timeout_s < 0: no limit; on time-out
 * RZK_E_STATE), copy its [height][width][3] image to out_rgb (may be NULL)
 * and its stats, and release it */
int rzk_server_wait(rzk_server *server, int64_t ticket, double timeout_s, float *out_rgb,
                    rzk_request_stats *stats);
int rzk_server_get_stats(rzk_server *server, rzk_server_stats *out);
/* message of the last failure (a failed worker fails every pending request) */
const char *rzk_server_last_error(rzk_server *server);

#ifdef __cplusplus
}
#endif
#endif
