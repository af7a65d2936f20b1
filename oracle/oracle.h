/*
 * oracle.h — CPU restatement of Renzoku's hot path (BVH traversal, VPL gather,
 * reduced matrix, clustering, per-slice guiding maps).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load liboracle.so, and only as the checker
 * or the timed CPU baseline.  The product path (renzoku_amd/) never links it.
 *
 * Every function follows the reference's arithmetic order (fp32, no FMA:
 * built with -O2 -ffp-contract=off like the reference's -O2 x86-64 build,
 * Makefile:10-11) and cites the reference file:line it restates.
 *
 * Parity pinning status per function is recorded in oracle.c's header.
 */
#ifndef RZK_ORACLE_H
#define RZK_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* record strides (floats) shared with the tests */
#define ORC_RAY_STRIDE  6   /* o[3], d[3] (d = Ray::d after the constructor) */
#define ORC_RECV_STRIDE 21  /* p, n(geometric), shading_n, wo, uvn.u, uvn.v, uvn.n */
#define ORC_VPL_STRIDE  21  /* p, n(dg.n), uvn.u, uvn.v, uvn.n, wi, power */
#define ORC_MAT_STRIDE  12  /* type, kd[3], ks[3], p0 (phong n | ward ax), p1 (ward ay), pad[3] */
#define ORC_CAM_STRIDE  19  /* eye[3], u[3], v[3], n[3], cop[2], scale[2], focal, width, height */

enum { ORC_LAMBERT = 0, ORC_PHONG = 1, ORC_WARD = 2 };
enum { ORC_CLAMP_NONE = 0, ORC_CLAMP_FORM_FACTOR = 1, ORC_CLAMP_REFLECTIVITY = 2, ORC_CLAMP_RADIANCE = 3 };
enum { ORC_SURF_FLIPPED = 1, ORC_SURF_TWO_SIDED = 2 };

typedef struct orc_scene orc_scene;

/* ---- RNG (random.h, mtwist.c) ---- */
void     orc_mt_seed32(uint32_t seed);
uint32_t orc_mt_lrand(void);
float    orc_mt_rand(void);          /* Random::operator() */
void     orc_mt_fill(int n, uint32_t *out);

/* ---- primitive tests ---- */
void orc_ray_make(const float o[3], const float d[3], float out[ORC_RAY_STRIDE]);
int  orc_box_hit(const float bmin[3], const float bmax[3], const float ray[ORC_RAY_STRIDE],
                 float tmin, float tmax);
int  orc_tri_hit(const float p0[3], const float p1[3], const float p2[3],
                 const float ray[ORC_RAY_STRIDE], float tmin, float tmax,
                 float *t, float *beta, float *gamma);
uint32_t orc_morton3d(float x, float y, float z);
int  orc_find_split(const uint32_t *codes, int first, int last);

/* ---- scene + BVH ---- */
orc_scene *orc_scene_create(int ntri, const float *pos, const float *tangent, const int *tri_surface,
                            int nsurf, const int *surf_material, const int *surf_light,
                            const int *surf_flags, int nmat, const float *mat);
void orc_scene_destroy(orc_scene *s);
int  orc_bvh_num_nodes(const orc_scene *s);
int  orc_bvh_depth(const orc_scene *s);
void orc_bvh_export(const orc_scene *s, int *sorted_tri, uint32_t *sorted_morton,
                    int *node_child /* 2 per node */, int *node_range /* 2 per node */,
                    float *node_box /* 6 per node */);
void orc_scene_box(const orc_scene *s, float out[6]);

/* ---- traversal (stats: per ray, [interior nodes visited, triangles tested]) ---- */
void orc_trace_closest(const orc_scene *s, int n, const float *rays, float tmin, float tmax,
                       float *t, float *beta, float *gamma, int *tri, int *stats);
void orc_trace_any(const orc_scene *s, int n, const float *rays, float tmin, float tmax,
                   uint8_t *hit, int *stats);
void orc_trace_segment(const orc_scene *s, int n, const float *p, const float *q, float tmin,
                       uint8_t *blocked, int *stats);
void orc_trace_closest_brute(const orc_scene *s, int n, const float *rays, float tmin, float tmax,
                             float *t, int *tri);

/* ---- hit records / camera ---- */
void orc_fill_receivers(const orc_scene *s, int n, const float *rays, const float *t,
                        const float *beta, const float *gamma, const int *tri,
                        float *recv, int *recv_mat, int *recv_light);
void orc_camera_rays(const float cam[ORC_CAM_STRIDE], int n, const int *px, const int *py,
                     float *rays);

/* ---- shading ---- */
void  orc_bsdf_eval(const float mat[ORC_MAT_STRIDE], const float uvn[9], const float wo[3],
                    const float wi[3], float out[3]);
float orc_form_factor_abs(const float p[3], const float pn[3], const float q[3], const float qn[3]);
void  orc_clamp_term(int mode, float threshold, const float power[3], const float fs[3],
                     const float fl[3], float G, float out[3]);
int   orc_gather_point(const orc_scene *s, const float recv[ORC_RECV_STRIDE], int recv_mat,
                       const float vpl[ORC_VPL_STRIDE], int vpl_mat, int clamp_mode, float clamp,
                       float tmin, float out[3]);
void  orc_incident_radiance(const orc_scene *s, const float recv[ORC_RECV_STRIDE],
                            const float vpl[ORC_VPL_STRIDE], int vpl_mat, int clamp_mode,
                            float clamp, float out[3]);
void  orc_gather(const orc_scene *s, int npts, const float *recv, const int *recv_mat,
                 int nvpl, const float *vpl, const int *vpl_mat, int clamp_mode, float clamp,
                 float tmin, float n_emitted, float *out_rgb, int nthreads);
/* the same with a weight per VPL applied after the clamp (gather_clusters over
 * cluster representatives, lightslice.cc:1325-1339) */
void  orc_gather_weighted(const orc_scene *s, int npts, const float *recv, const int *recv_mat,
                          int nvpl, const float *vpl, const int *vpl_mat, const float *vpl_w,
                          int clamp_mode, float clamp, float tmin, float n_emitted, float *out_rgb,
                          int nthreads);

/* ---- light transport matrix + clustering ---- */
void orc_reduced_matrix(const orc_scene *s, int rows, const float *recv, const int *recv_mat,
                        int nvpl, const float *vpl, const int *vpl_mat, int clamp_mode,
                        float clamp, float tmin, float *R, float *V, int nthreads);
void orc_alphas(int rows, int cols, const float *R, float *alphas);
void orc_column_magnitudes(int rows, int cols, const float *R, float *mags);
void orc_cluster_assign(int rows, int cols, const float *R, int ncl, const int *centers,
                        const float *weights, int *assign, int nthreads);
int  orc_pdf_build(int n, const float *arr, float *pdf, float *cdf);
void orc_pdf_sample(int n, const float *pdf, const float *cdf, float u, int *k, float *pk);
int  orc_cluster_by_sampling(int rows, int cols, const float *R, int num_clusters,
                             int *centers, float *center_w, int *assign, int *reps,
                             float *rep_w, int *kept, int nthreads);
/* per-slice refinement (refine.cpp, lightslice.cc:837-902); draws continue the
 * global MT stream */
int  orc_refine_slices(int S, int V, const float *R, int nnb, const int *nb, int ncl,
                       const int *cl_off, const int *cl_mem, const int *cl_rep, const float *cl_w,
                       const float *mags, int max_slice_clusters, int *out_count, int *out_rep,
                       float *out_w, int *out_size, int *out_perm, int *out_draws);
/* the same for slices 0 .. n_slices - 1 only (a prefix of the sequential run) */
int  orc_refine_slices_n(int S, int n_slices, int V, const float *R, int nnb, const int *nb, int ncl,
                         const int *cl_off, const int *cl_mem, const int *cl_rep, const float *cl_w,
                         const float *mags, int max_slice_clusters, int *out_count, int *out_rep,
                         float *out_w, int *out_size, int *out_perm, int *out_draws);
float orc_slice_cost(int S, const float *R, int nnb, const int *nb, int n, const int *idx);

/* ---- 6-D kd-tree slicing (kdslice.cpp, kdtree_n.h, lightslice.cc:142-303) ---- */
void orc_map_vec6(const float *rec, const float *mat, float diagonal, float *out);
int  orc_kd_build(const float *pts, int n, int max_size, uint32_t seed, int *order, int *slices,
                  float *boxes, int cap_slices, int *axis, float *plane, int *cl, int cap_heap,
                  int *heap_used);
/* the same with the pivots drawn from the global MT stream in recursion order */
int  orc_kd_build_mt(const float *pts, int n, int max_size, int *order, int *slices, float *boxes,
                     int cap_slices, int *axis, float *plane, int *cl, int cap_heap, int *heap_used,
                     int64_t *draws);
void orc_kd_find(const int *axis, const float *plane, const int *cl, int heap_used, const float *q,
                 int m, int *out);
void orc_kd_nearest(const int *axis, const float *plane, const int *cl, int heap_used,
                    const float *boxes, int nslices, const float *q, int m, int k, int *out);

/* ---- fp32 transcendentals of the direction samplers (test hook) ---- */
void orc_xmath(int fn, int n, const float *x, const float *y, float *out);

/* ---- guiding distributions ---- */
void orc_direction_to_square(const float wi[3], float out[2]);
void orc_square_to_direction(const float s[2], float out[3]);
void orc_slice_cluster_info(const orc_scene *s, const float recv[ORC_RECV_STRIDE], int nreps,
                            const int *reps, const float *vpl, const int *vpl_mat,
                            const float *Vrow, int clamp_mode, float clamp,
                            float *slice_wi, int *upper);
void orc_slice_cluster_pdf(int rows, const float *R, int nnb, const int *nb_rows, int nreps,
                           const int *reps, const float *rep_w, const int *upper, float *out);
int  orc_build_map(int res, int nreps, const float *slice_wi, const int *upper,
                   const float *rep_w, const float *R_rep_rgb, const float uvn[9], float eps,
                   int valid, float *map, float *row_cdf, float *row_pdf, float *col_cdf,
                   float *total);

/* ---- guided path tracing (vpl.cc:934-1041, area_light.cc:48-142) ---- */
void orc_scene_set_lights(orc_scene *s, int nlights, const float *radiance);
void orc_philox(uint32_t c[4], uint32_t k0, uint32_t k1);
void orc_bsdf_pdf(const float mat[ORC_MAT_STRIDE], const float uvn[9], const float wo[3], const float wi[3],
                  float *pdf);
void orc_bsdf_sample(const float mat[ORC_MAT_STRIDE], const float uvn[9], const float wo[3], const uint32_t u[3],
                     float wi[3], float *pdf, float brdf[3]);
void orc_guided_pt(const orc_scene *s, const float cam[ORC_CAM_STRIDE], int spp, int max_bounce, int direct,
                   int guide_mode, uint32_t seed, float tmin, int tile, int res, const float *maps,
                   const float *row_cdf, const float *col_cdf, const float *totals, int x0, int y0, int w,
                   int h, float *out, int nthreads);
/* the same path sampler from explicit primary rays: ray i (a ray record) belongs
 * to pixel (pix[2i], pix[2i+1]) of a width x height image and is that pixel's
 * sample smp[i]; out[3i..] is its radiance (MonteCarloIntegrator::radiance(PrimaryRay&),
 * monte_carlo_integrator.cc:40-43) */
void orc_guided_pt_rays(const orc_scene *s, int width, int height, int max_bounce, int direct, int guide_mode,
                        uint32_t seed, float tmin, int tile, int res, const float *maps, const float *row_cdf,
                        const float *col_cdf, const float *totals, int n, const float *rays, const int *pix,
                        const int *smp, float *out, int nthreads);
/* the primary ray records orc_guided_pt draws: sample smp[i] of pixel (px[i], py[i]) at spp samples per pixel */
void orc_camera_sample_rays(const float cam[ORC_CAM_STRIDE], int spp, uint32_t seed, int n, const int *px,
                            const int *py, const int *smp, float *out);

/* bounce-1 VPL generation (vpl_generator.cc:127-181); see oracle.c */
void orc_sincos_turn(float u, float *s, float *c);
int  orc_generate_vpls(const orc_scene *s, long long num_vpls, uint32_t seed, long long max_samples, float tmin,
                       float *out_vpl, int *out_mat, long long *out_samples);

#ifdef __cplusplus
}
#endif

#endif
