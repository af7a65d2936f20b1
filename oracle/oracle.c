/*
 * oracle.c — CPU restatement of Renzoku's hot path.  TEST INFRASTRUCTURE ONLY
 * (see oracle.h): the product library never links or calls this file.
 *
 * Built -O2 -ffp-contract=off, fp32 everywhere, each expression evaluated in the
 * reference's C++ left-to-right order (the reference is built -O2 for x86-64
 * without -mfma, Makefile:10-11, so it never fuses a*b+c).
 *
 * Parity pinning (what checks each function besides re-reading the reference):
 *   MT19937 stream ............ oracle/_ref build of the reference's own mtwist.c
 *                               (tests/test_oracle_rng.py) + published MT19937 KAT
 *   ray/box/triangle tests .... closed forms, hand-built KATs (edge, vertex, |det|<1e-6)
 *   Morton / find_split ....... integer KATs + brute-force split search
 *   BVH closest / any ......... brute force over all triangles (the reference's
 *                               own Aggregate::hit loop, aggregate.cc:56-82)
 *   hit record / Onb .......... orthonormality invariants, flip rules
 *   BSDFs / form factor ....... closed forms (Lambert kd/pi, Phong/Ward peaks),
 *                               reciprocity of Lambert/Ward
 *   clamp_term ................ closed form per branch
 *   gather / incident ......... composition of the above, brute-force visibility
 *   alphas .................... float64 closed form alpha_i = sum_j |R_i-R_j|-ish
 *                               identity (||R_i|| sum_j||R_j|| - R_i.sum_j R_j)
 *   cluster assignment ........ brute-force argmin recomputed in float64 on
 *                               non-tied columns
 *   DiscretePdf / maps ........ CDF invariants (monotone, last = 1, strict <)
 *   concentric map ............ round-trip invariant + Shirley-Chiu closed form
 * Functions whose parity with the reference binary rests only on the reading
 * of the source: glibc powf/expf inside the Phong/Ward evaluations (ULP level,
 * "parity unpinned" beyond the tolerance stated in the tests).  The
 * transcendentals that pick a sampled direction (sin, cos, tan, atan, atan2,
 * acos, log, pow) are restated in fp32 basic operations (xsincos .. xpow below),
 * as the kernels restate them: within 3.5 ulp of the exact functions (pinned
 * against float64 numpy in tests/test_oracle_xmath.py), where glibc's float
 * functions are within 1 ulp, so a sampled direction may differ from the
 * reference binary's by a few ulp.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* common.h:38-55 */
#define A_PI        3.14159265358979f
#define INV_PI      0.31830988618379f
#define TWO_PI      6.28318530717959f
#define INV_2PI     0.15915494309190f
#define ZERO_EPS    1e-6f
#define ONB_EPS     0.01f

typedef struct { float x, y, z; } v3;

static inline v3 mk(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 ld(const float *p) { return mk(p[0], p[1], p[2]); }
static inline void st(float *p, v3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mulv(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 divv(v3 a, v3 b) { return mk(a.x / b.x, a.y / b.y, a.z / b.z); }
static inline v3 scl(v3 a, float t) { return mk(a.x * t, a.y * t, a.z * t); }   /* vec3_cpu.h:254-258 */
static inline v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
static inline float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; } /* vec3_cpu.h:300 */
static inline v3 cross(v3 a, v3 b) {                                            /* vec3_cpu.h:303 */
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float len(v3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); } /* :149 */
static inline v3 unit(v3 a) { float inv = 1.0f / len(a); return mk(a.x * inv, a.y * inv, a.z * inv); } /* :406 */
static inline int isnan3(v3 a) { return a.x != a.x || a.y != a.y || a.z != a.z; }
/* std::min / std::max on floats: (b < a) ? b : a and (a < b) ? b : a */
static inline float fmin_std(float a, float b) { return (b < a) ? b : a; }
static inline float fmax_std(float a, float b) { return (a < b) ? b : a; }
/* Vec3 min/max, vec3_cpu.h:415-429 */
static inline v3 vmin3(v3 a, v3 b) { v3 t = a; if (b.x < a.x) t.x = b.x; if (b.y < a.y) t.y = b.y; if (b.z < a.z) t.z = b.z; return t; }
static inline v3 vmax3(v3 a, v3 b) { v3 t = a; if (b.x > a.x) t.x = b.x; if (b.y > a.y) t.y = b.y; if (b.z > a.z) t.z = b.z; return t; }

/* ======================================================================= */
/* MT19937 (random.h:17-71, mtwist.c:378-406, mtwist.h:415-426).           */
/* The reference seeds with the 1998 Knuth-69069 recurrence and keeps the  */
/* state reversed; its output equals the classic genrand() stream after    */
/* sgenrand(seed).  Restated from Matsumoto & Nishimura (1998).             */
/* ======================================================================= */
#define MT_N 624
#define MT_M 397
static uint32_t mt_x[MT_N];
static int mt_i = MT_N + 1;

void orc_mt_seed32(uint32_t seed) {
    if (seed == 0) seed = 4357;              /* mtwist.c:289, DEFAULT_SEED32_OLD */
    mt_x[0] = seed;
    for (int i = 1; i < MT_N; ++i) mt_x[i] = 69069u * mt_x[i - 1];
    mt_i = MT_N;
}

static void mt_twist(void) {
    for (int k = 0; k < MT_N; ++k) {
        uint32_t y = (mt_x[k] & 0x80000000u) | (mt_x[(k + 1) % MT_N] & 0x7fffffffu);
        mt_x[k] = mt_x[(k + MT_M) % MT_N] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    }
    mt_i = 0;
}

uint32_t orc_mt_lrand(void) {
    if (mt_i > MT_N) orc_mt_seed32(4357);
    if (mt_i >= MT_N) mt_twist();
    uint32_t y = mt_x[mt_i++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

/* random.h:60-63: 24-bit float in [0,1) */
float orc_mt_rand(void) { return (float)(orc_mt_lrand() & 0x00ffffffu) * (1.0f / 16777216.0f); }

void orc_mt_fill(int n, uint32_t *out) { for (int i = 0; i < n; ++i) out[i] = orc_mt_lrand(); }

/* ======================================================================= */
/* Ray, box, triangle                                                      */
/* ======================================================================= */
typedef struct { v3 o, d, inv_d, org_inv_d; } ray_t;

/* ray.h:17-36 — d is normalised, NaN -> 0, |d_i| < 1e-6 uses 1e-6 in inv_d only */
static ray_t ray_ctor(v3 o, v3 d_in) {
    ray_t r;
    r.o = o;
    r.d = unit(d_in);
    if (isnan3(r.d)) r.d = mk(0.f, 0.f, 0.f);
    v3 dd = mk(fabsf(r.d.x) < ZERO_EPS ? ZERO_EPS : r.d.x,
               fabsf(r.d.y) < ZERO_EPS ? ZERO_EPS : r.d.y,
               fabsf(r.d.z) < ZERO_EPS ? ZERO_EPS : r.d.z);
    r.inv_d = divv(mk(1.f, 1.f, 1.f), dd);
    r.org_inv_d = mulv(neg(o), r.inv_d);
    return r;
}

/* A ray record holds Ray::o and Ray::d as the constructor left them; the
 * derived members are recomputed exactly as the constructor does. */
static ray_t ray_from_record(const float *rec) {
    ray_t r;
    r.o = ld(rec);
    r.d = ld(rec + 3);
    v3 dd = mk(fabsf(r.d.x) < ZERO_EPS ? ZERO_EPS : r.d.x,
               fabsf(r.d.y) < ZERO_EPS ? ZERO_EPS : r.d.y,
               fabsf(r.d.z) < ZERO_EPS ? ZERO_EPS : r.d.z);
    r.inv_d = divv(mk(1.f, 1.f, 1.f), dd);
    r.org_inv_d = mulv(neg(r.o), r.inv_d);
    return r;
}

void orc_ray_make(const float o[3], const float d[3], float out[ORC_RAY_STRIDE]) {
    ray_t r = ray_ctor(ld(o), ld(d));
    st(out, r.o);
    st(out + 3, r.d);
}

/* boundingbox.h:238-280 */
static int box_hit(v3 bmin, v3 bmax, const ray_t *r, float tmin, float tmax) {
    v3 v0 = add(mulv(bmin, r->inv_d), r->org_inv_d);
    v3 v1 = add(mulv(bmax, r->inv_d), r->org_inv_d);
    float tminbox = fmax_std(tmin, fmax_std(fmin_std(v0.x, v1.x), fmax_std(fmin_std(v0.y, v1.y), fmin_std(v0.z, v1.z))));
    float tmaxbox = fmin_std(tmax, fmin_std(fmax_std(v0.x, v1.x), fmin_std(fmax_std(v0.y, v1.y), fmax_std(v0.z, v1.z))));
    return (tminbox <= tmaxbox) && (tmax >= tminbox && tmaxbox >= tmin);
}

int orc_box_hit(const float bmin[3], const float bmax[3], const float ray[ORC_RAY_STRIDE],
                float tmin, float tmax) {
    ray_t r = ray_from_record(ray);
    return box_hit(ld(bmin), ld(bmax), &r, tmin, tmax);
}

/* triangle.cc:158-232 (Cramer's rule, Shirley's form) */
static int tri_hit(v3 p0, v3 p1, v3 p2, const ray_t *r, float tmin, float tmax,
                   float *t_out, float *b_out, float *g_out) {
    v3 p01 = sub(p0, p1);
    float A = p01.x, B = p01.y, C = p01.z;
    v3 p02 = sub(p0, p2);
    float D = p02.x, E = p02.y, F = p02.z;
    float G = r->d.x, H = r->d.y, I = r->d.z;
    v3 p0org = sub(p0, r->o);
    float J = p0org.x, K = p0org.y, L = p0org.z;
    float EIHF = E * I - H * F;
    float GFDI = G * F - D * I;
    float DHEG = D * H - E * G;
    float denom = (A * EIHF + B * GFDI + C * DHEG);
    if (-ZERO_EPS < denom && denom < ZERO_EPS) return 0;
    float beta = (J * EIHF + K * GFDI + L * DHEG) / denom;
    if (beta < 0.0f || beta > 1.0f) return 0;
    float AKJB = A * K - J * B;
    float JCAL = J * C - A * L;
    float BLKC = B * L - K * C;
    float gamma = (I * AKJB + H * JCAL + G * BLKC) / denom;
    if (gamma < 0.0f || beta + gamma > 1.0f) return 0;
    float t = -(F * AKJB + E * JCAL + D * BLKC) / denom;
    if (t >= tmin && t <= tmax) {
        if (t_out) *t_out = t;
        if (b_out) *b_out = beta;
        if (g_out) *g_out = gamma;
        return 1;
    }
    return 0;
}

int orc_tri_hit(const float p0[3], const float p1[3], const float p2[3],
                const float ray[ORC_RAY_STRIDE], float tmin, float tmax,
                float *t, float *beta, float *gamma) {
    ray_t r = ray_from_record(ray);
    return tri_hit(ld(p0), ld(p1), ld(p2), &r, tmin, tmax, t, beta, gamma);
}

/* ======================================================================= */
/* Morton LBVH (aggregate_bvh.cc:48-240, morton.h:28-120)                  */
/* ======================================================================= */
static uint32_t spread10(uint32_t v) {
    /* place bit k of a 10-bit value at bit 3k (Morton interleave) */
    uint32_t r = 0;
    for (int k = 0; k < 10; ++k) r |= ((v >> k) & 1u) << (3 * k);
    return r;
}

/* morton.h:37-46 */
uint32_t orc_morton3d(float x, float y, float z) {
    x = fmin_std(fmax_std(x * 1024.0f, 0.0f), 1023.0f);
    y = fmin_std(fmax_std(y * 1024.0f, 0.0f), 1023.0f);
    z = fmin_std(fmax_std(z * 1024.0f, 0.0f), 1023.0f);
    return (spread10((uint32_t)x) << 2) | (spread10((uint32_t)y) << 1) | spread10((uint32_t)z);
}

static int clz32(uint32_t v) { return v ? __builtin_clz(v) : 32; }

/* morton.h:48-88: the split is the last index in [first, last) whose code shares
 * more than the range's common prefix with codes[first]; equal codes split at
 * the midpoint.  Restated as a plain scan-free binary search on that predicate. */
int orc_find_split(const uint32_t *codes, int first, int last) {
    uint32_t fc = codes[first], lc = codes[last];
    if (fc == lc) return (first + last) >> 1;
    int cp = clz32(fc ^ lc);
    int lo = first, hi = last;               /* pred(lo) true, pred(hi) false */
    while (hi - lo > 1) {
        int mid = lo + (hi - lo) / 2;
        if (clz32(fc ^ codes[mid]) > cp) lo = mid; else hi = mid;
    }
    return lo;
}

typedef struct { float bmin[3], bmax[3]; int start, end, left, right, level; } node_t;

typedef struct { v3 p0, p1, p2, n; float area, cum, ratio; int light, tri; } ltri_t;
typedef struct { v3 le; int first, count; float sel_cdf, sel_pdf; } light_t;

struct orc_scene {
    int nlights, nltri;
    light_t *lights;
    ltri_t *ltris;
    int ntri;
    v3 *p0, *p1, *p2, *n, *tan;     /* original order */
    int *tri_surface;
    int nsurf;
    int *surf_material, *surf_light, *surf_flags;
    int nmat;
    float *mat;
    /* bvh */
    int num_nodes, depth;
    node_t *nodes;
    int *sorted_tri;                /* sorted position -> original index */
    uint32_t *sorted_morton;
    v3 box_min, box_max;
};

typedef struct { uint32_t code; int index; } mentry;
/* morton.h:109-112 — codes are < 2^30 so the int difference never overflows */
static int mentry_cmp(const void *a, const void *b) {
    return (int)((const mentry *)a)->code - (int)((const mentry *)b)->code;
}

static void bvh_build(orc_scene *s) {
    int n = s->ntri;
    s->num_nodes = 0;
    s->nodes = NULL;
    if (n <= 0) return;
    v3 *bmin = malloc(sizeof(v3) * n), *bmax = malloc(sizeof(v3) * n);
    v3 smin = mk(0.f, 0.f, 0.f), smax = mk(0.f, 0.f, 0.f);     /* BoundingBox::reset(), boundingbox.h:100-103 */
    for (int k = 0; k < n; ++k) {
        bmin[k] = vmin3(vmin3(s->p0[k], s->p1[k]), s->p2[k]);  /* boundingbox.h:27-30 */
        bmax[k] = vmax3(vmax3(s->p0[k], s->p1[k]), s->p2[k]);
        smin = vmin3(smin, bmin[k]);
        smax = vmax3(smax, bmax[k]);
    }
    s->box_min = smin;
    s->box_max = smax;
    mentry *e = malloc(sizeof(mentry) * n);
    for (int k = 0; k < n; ++k) {
        v3 c = scl(add(bmin[k], bmax[k]), 0.5f);                /* morton.h:98-102 */
        c = divv(sub(c, smin), sub(smax, smin));
        e[k].code = orc_morton3d(c.x, c.y, c.z);
        e[k].index = k;
    }
    qsort(e, n, sizeof(mentry), mentry_cmp);                   /* morton.h:114-120 */
    s->sorted_tri = malloc(sizeof(int) * n);
    s->sorted_morton = malloc(sizeof(uint32_t) * n);
    for (int i = 0; i < n; ++i) { s->sorted_tri[i] = e[i].index; s->sorted_morton[i] = e[i].code; }
    free(e);

    /* BFS split, aggregate_bvh.cc:118-176 */
    int cap = 2 * n;
    node_t *nodes = calloc(cap, sizeof(node_t));
    int *lvl_first = malloc(sizeof(int) * (n + 64)), *lvl_last = malloc(sizeof(int) * (n + 64));
    int nlev = 0;
    int first = 0, last = 1, next_last = 1, level = 0;
    nodes[0].start = 0; nodes[0].end = n - 1;
    lvl_first[0] = 0; lvl_last[0] = 1; nlev = 1;
    int num_nodes = 1;
    for (;;) {
        node_t *nd = &nodes[first];
        nd->level = level;
        if (nd->start == nd->end) {
            nd->left = nd->right = -1;
        } else {
            int split = orc_find_split(s->sorted_morton, nd->start, nd->end);
            nd->left = next_last; nd->right = next_last + 1;
            next_last += 2; num_nodes += 2;
            nodes[nd->left].start = nd->start; nodes[nd->left].end = split;
            nodes[nd->right].start = split + 1; nodes[nd->right].end = nd->end;
        }
        first++;
        if (first == last) {
            if (next_last == last) break;
            last = next_last;
            lvl_first[nlev] = first; lvl_last[nlev] = last; nlev++;
            level++;
        }
    }
    /* bottom-up boxes, aggregate_bvh.cc:185-207 */
    for (int li = nlev - 1; li >= 0; --li) {
        for (int j = lvl_first[li]; j < lvl_last[li]; ++j) {
            node_t *nd = &nodes[j];
            v3 a, b;
            if (nd->left < 0) {
                int ot = s->sorted_tri[nd->start];
                a = bmin[ot]; b = bmax[ot];
            } else {
                a = vmin3(ld(nodes[nd->left].bmin), ld(nodes[nd->right].bmin));
                b = vmax3(ld(nodes[nd->left].bmax), ld(nodes[nd->right].bmax));
            }
            st(nd->bmin, a); st(nd->bmax, b);
        }
    }
    s->nodes = nodes;
    s->num_nodes = num_nodes;
    s->depth = nlev;
    free(lvl_first); free(lvl_last); free(bmin); free(bmax);
}

orc_scene *orc_scene_create(int ntri, const float *pos, const float *tangent, const int *tri_surface,
                            int nsurf, const int *surf_material, const int *surf_light,
                            const int *surf_flags, int nmat, const float *mat) {
    orc_scene *s = calloc(1, sizeof(orc_scene));
    s->ntri = ntri;
    s->p0 = malloc(sizeof(v3) * (ntri + 1)); s->p1 = malloc(sizeof(v3) * (ntri + 1));
    s->p2 = malloc(sizeof(v3) * (ntri + 1)); s->n = malloc(sizeof(v3) * (ntri + 1));
    s->tan = malloc(sizeof(v3) * (ntri + 1));
    s->tri_surface = malloc(sizeof(int) * (ntri + 1));
    for (int k = 0; k < ntri; ++k) {
        s->p0[k] = ld(pos + 9 * k); s->p1[k] = ld(pos + 9 * k + 3); s->p2[k] = ld(pos + 9 * k + 6);
        v3 p01 = sub(s->p0[k], s->p1[k]), p02 = sub(s->p0[k], s->p2[k]);
        s->n[k] = unit(cross(p01, p02));                          /* triangle.cc:17-18 */
        s->tan[k] = ld(tangent + 3 * k);
        s->tri_surface[k] = tri_surface[k];
    }
    s->nsurf = nsurf;
    s->surf_material = malloc(sizeof(int) * (nsurf + 1));
    s->surf_light = malloc(sizeof(int) * (nsurf + 1));
    s->surf_flags = malloc(sizeof(int) * (nsurf + 1));
    memcpy(s->surf_material, surf_material, sizeof(int) * nsurf);
    memcpy(s->surf_light, surf_light, sizeof(int) * nsurf);
    memcpy(s->surf_flags, surf_flags, sizeof(int) * nsurf);
    s->nmat = nmat;
    s->mat = malloc(sizeof(float) * ORC_MAT_STRIDE * (nmat + 1));
    memcpy(s->mat, mat, sizeof(float) * ORC_MAT_STRIDE * nmat);
    bvh_build(s);
    return s;
}

void orc_scene_destroy(orc_scene *s) {
    if (!s) return;
    free(s->p0); free(s->p1); free(s->p2); free(s->n); free(s->tan); free(s->tri_surface);
    free(s->surf_material); free(s->surf_light); free(s->surf_flags); free(s->mat);
    free(s->nodes); free(s->sorted_tri); free(s->sorted_morton);
    free(s->lights); free(s->ltris);
    free(s);
}

int orc_bvh_num_nodes(const orc_scene *s) { return s->num_nodes; }
int orc_bvh_depth(const orc_scene *s) { return s->depth; }
void orc_scene_box(const orc_scene *s, float out[6]) { st(out, s->box_min); st(out + 3, s->box_max); }

void orc_bvh_export(const orc_scene *s, int *sorted_tri, uint32_t *sorted_morton,
                    int *node_child, int *node_range, float *node_box) {
    for (int i = 0; i < s->ntri; ++i) { sorted_tri[i] = s->sorted_tri[i]; sorted_morton[i] = s->sorted_morton[i]; }
    for (int j = 0; j < s->num_nodes; ++j) {
        node_child[2 * j] = s->nodes[j].left; node_child[2 * j + 1] = s->nodes[j].right;
        node_range[2 * j] = s->nodes[j].start; node_range[2 * j + 1] = s->nodes[j].end;
        memcpy(node_box + 6 * j, s->nodes[j].bmin, 3 * sizeof(float));
        memcpy(node_box + 6 * j + 3, s->nodes[j].bmax, 3 * sizeof(float));
    }
}

/* ======================================================================= */
/* Traversal (aggregate_bvh.cc:243-281 stack, :402-448 closest, :452-478 any) */
/* ======================================================================= */
typedef struct { int buf[64]; int top; } stack_t64;
static inline void spush(stack_t64 *q, int v) { if (q->top < 63) q->buf[++q->top] = v; }
static inline int spop(stack_t64 *q) { return q->top >= 0 ? q->buf[q->top--] : -1; }

static int bvh_closest(const orc_scene *s, const ray_t *r, float tmin, float tmax,
                       float *t, float *beta, float *gamma, int *tri, int *stats) {
    int nvis = 0, ntest = 0, hit = 0, best = -1;
    float bt = 0, bb = 0, bg = 0;
    if (s->num_nodes > 0 && box_hit(ld(s->nodes[0].bmin), ld(s->nodes[0].bmax), r, tmin, tmax)) {
        stack_t64 q; q.top = -1;
        spush(&q, 0);
        while (q.top >= 0) {
            const node_t *nd = &s->nodes[spop(&q)];
            if (nd->left < 0 && nd->right < 0) {
                int ot = s->sorted_tri[nd->start];
                float tt, b, g;
                ntest++;
                if (tri_hit(s->p0[ot], s->p1[ot], s->p2[ot], r, tmin, tmax, &tt, &b, &g)) {
                    tmax = tt; bt = tt; bb = b; bg = g; hit = 1; best = ot;
                }
            } else {
                nvis++;
                if (nd->left >= 0 && box_hit(ld(s->nodes[nd->left].bmin), ld(s->nodes[nd->left].bmax), r, tmin, tmax))
                    spush(&q, nd->left);
                if (nd->right >= 0 && box_hit(ld(s->nodes[nd->right].bmin), ld(s->nodes[nd->right].bmax), r, tmin, tmax))
                    spush(&q, nd->right);
            }
        }
    }
    if (t) *t = hit ? bt : 0.f;
    if (beta) *beta = hit ? bb : 0.f;
    if (gamma) *gamma = hit ? bg : 0.f;
    if (tri) *tri = hit ? best : -1;
    if (stats) { stats[0] = nvis; stats[1] = ntest; }
    return hit;
}

static int bvh_any(const orc_scene *s, const ray_t *r, float tmin, float tmax, int *stats) {
    int nvis = 0, ntest = 0, hit = 0;
    if (s->num_nodes > 0 && box_hit(ld(s->nodes[0].bmin), ld(s->nodes[0].bmax), r, tmin, tmax)) {
        stack_t64 q; q.top = -1;
        spush(&q, 0);
        while (q.top >= 0) {
            const node_t *nd = &s->nodes[spop(&q)];
            if (nd->left < 0 && nd->right < 0) {
                int ot = s->sorted_tri[nd->start];
                ntest++;
                if (tri_hit(s->p0[ot], s->p1[ot], s->p2[ot], r, tmin, tmax, NULL, NULL, NULL)) { hit = 1; break; }
            } else {
                nvis++;
                if (nd->left >= 0 && box_hit(ld(s->nodes[nd->left].bmin), ld(s->nodes[nd->left].bmax), r, tmin, tmax))
                    spush(&q, nd->left);
                if (nd->right >= 0 && box_hit(ld(s->nodes[nd->right].bmin), ld(s->nodes[nd->right].bmax), r, tmin, tmax))
                    spush(&q, nd->right);
            }
        }
    }
    if (stats) { stats[0] = nvis; stats[1] = ntest; }
    return hit;
}

/* aggregate.cc:84-90 */
static int segment_blocked(const orc_scene *s, v3 p, v3 q, float tmin, int *stats) {
    ray_t r = ray_ctor(p, sub(q, p));
    float tmax = len(sub(q, p)) - tmin;
    return bvh_any(s, &r, tmin, tmax, stats);
}

void orc_trace_closest(const orc_scene *s, int n, const float *rays, float tmin, float tmax,
                       float *t, float *beta, float *gamma, int *tri, int *stats) {
    for (int i = 0; i < n; ++i) {
        ray_t r = ray_from_record(rays + ORC_RAY_STRIDE * i);
        bvh_closest(s, &r, tmin, tmax, t + i, beta + i, gamma + i, tri + i, stats ? stats + 2 * i : NULL);
    }
}

void orc_trace_any(const orc_scene *s, int n, const float *rays, float tmin, float tmax,
                   uint8_t *hit, int *stats) {
    for (int i = 0; i < n; ++i) {
        ray_t r = ray_from_record(rays + ORC_RAY_STRIDE * i);
        hit[i] = (uint8_t)bvh_any(s, &r, tmin, tmax, stats ? stats + 2 * i : NULL);
    }
}

void orc_trace_segment(const orc_scene *s, int n, const float *p, const float *q, float tmin,
                       uint8_t *blocked, int *stats) {
    for (int i = 0; i < n; ++i)
        blocked[i] = (uint8_t)segment_blocked(s, ld(p + 3 * i), ld(q + 3 * i), tmin, stats ? stats + 2 * i : NULL);
}

/* brute force over every triangle in original order: the reference's own
 * non-BVH Aggregate::hit loop semantics (aggregate.cc:56-71) */
void orc_trace_closest_brute(const orc_scene *s, int n, const float *rays, float tmin, float tmax,
                             float *t, int *tri) {
    for (int i = 0; i < n; ++i) {
        ray_t r = ray_from_record(rays + ORC_RAY_STRIDE * i);
        float tm = tmax; int best = -1; float bt = 0;
        for (int k = 0; k < s->ntri; ++k) {
            float tt;
            if (tri_hit(s->p0[k], s->p1[k], s->p2[k], &r, tmin, tm, &tt, NULL, NULL)) { tm = tt; bt = tt; best = k; }
        }
        t[i] = best >= 0 ? bt : 0.f; tri[i] = best;
    }
}

/* ======================================================================= */
/* Hit records, Onb, camera                                                */
/* ======================================================================= */
/* onb.cc:42-54 */
static void onb_from_n(v3 n, v3 *u, v3 *v, v3 *w) {
    v3 nn = unit(n);
    v3 uu = cross(nn, mk(1.f, 0.f, 0.f));
    if (len(uu) < ONB_EPS) uu = cross(nn, mk(0.f, 1.f, 0.f));
    v3 vv = cross(nn, uu);
    uu = unit(uu);           /* make_unit_vector: e *= 1/length (vec3_cpu.h:155-160) */
    vv = unit(vv);
    *u = uu; *v = vv; *w = nn;
}

/* onb.cc:99-109 */
static void onb_from_nu(v3 n, v3 u_in, v3 *u, v3 *v, v3 *w) {
    v3 c = unit(cross(n, u_in));
    if (isnan3(c)) { onb_from_n(n, u, v, w); return; }
    v3 nn = unit(n);
    *w = nn; *v = c; *u = cross(c, nn);
}

/* triangle.cc:277-296 + surface.cc:62-96 + local_geometry.h:21-30 */
static void fill_receiver(const orc_scene *s, const ray_t *r, float t, float beta, float gamma,
                          int ot, float *rec, int *mat_out, int *light_out) {
    memset(rec, 0, sizeof(float) * ORC_RECV_STRIDE);
    if (ot < 0) { *mat_out = -1; *light_out = -1; return; }
    v3 n = s->n[ot];
    v3 n0 = n, n1 = n, n2 = n;
    v3 sn = unit(add(add(n0, scl(sub(n1, n0), beta)), scl(sub(n2, n0), gamma)));
    int surf = s->tri_surface[ot];
    int m = s->surf_material[surf];
    int flags = s->surf_flags[surf];
    if (flags & ORC_SURF_FLIPPED) { n = neg(n); sn = neg(sn); }
    if ((flags & ORC_SURF_TWO_SIDED) && dot(sn, r->d) > 0.0f) {
        if (m >= 0) { sn = neg(sn); n = neg(n); }   /* no glass on this path */
    }
    v3 u, v, w;
    onb_from_nu(sn, s->tan[ot], &u, &v, &w);
    v3 p = add(r->o, scl(r->d, t));
    st(rec + 0, p); st(rec + 3, n); st(rec + 6, sn); st(rec + 9, neg(r->d));
    st(rec + 12, u); st(rec + 15, v); st(rec + 18, w);
    *mat_out = m;
    *light_out = s->surf_light[surf];
}

void orc_fill_receivers(const orc_scene *s, int n, const float *rays, const float *t,
                        const float *beta, const float *gamma, const int *tri,
                        float *recv, int *recv_mat, int *recv_light) {
    for (int i = 0; i < n; ++i) {
        ray_t r = ray_from_record(rays + ORC_RAY_STRIDE * i);
        fill_receiver(s, &r, t[i], beta[i], gamma[i], tri[i], recv + ORC_RECV_STRIDE * i, recv_mat + i, recv_light + i);
    }
}

/* camera.cc:52-65 (shoot_ray(i, j) with i = row, j = column) */
void orc_camera_rays(const float cam[ORC_CAM_STRIDE], int n, const int *px, const int *py, float *rays) {
    v3 eye = ld(cam), uu = ld(cam + 3), vv = ld(cam + 6), nn = ld(cam + 9);
    float copx = cam[12], copy = cam[13], sx = cam[14], sy = cam[15], focal = cam[16];
    float height = cam[18];
    for (int k = 0; k < n; ++k) {
        int i = py[k], j = px[k];
        float pxf = (float)j + 0.5f;
        float pyf = (height - 1.0f - (float)i) + 0.5f;
        float fx = (pxf - copx) / sx, fy = (pyf - copy) / sy;
        v3 d = add(add(scl(uu, fx), scl(vv, fy)), scl(nn, -focal));   /* onb.cc:117-119 */
        d = unit(d);
        ray_t r = ray_ctor(eye, d);
        st(rays + ORC_RAY_STRIDE * k, r.o);
        st(rays + ORC_RAY_STRIDE * k + 3, r.d);
    }
}

/* ======================================================================= */
/* BSDFs, form factor, clamps, gather                                      */
/* ======================================================================= */
static void bsdf_eval(const float *m, v3 u, v3 v, v3 w, v3 wo, v3 wi, float out[3]) {
    int type = (int)m[0];
    v3 kd = ld(m + 1), ks = ld(m + 4);
    out[0] = out[1] = out[2] = 0.f;
    if (type == ORC_WARD) {                                   /* ward.cc:104-116 */
        float ax = m[7], ay = m[8];
        float ax2 = ax * ax, ay2 = ay * ay;
        float inv4 = 1.0f / (4.0f * A_PI * ax * ay);          /* ward.cc:6-10 */
        v3 lwi = mk(dot(u, wi), dot(v, wi), dot(w, wi));      /* onb.cc:113-115 */
        v3 lwo = mk(dot(u, wo), dot(v, wo), dot(w, wo));
        if (lwi.z <= ZERO_EPS || lwo.z <= ZERO_EPS) return;
        v3 wh = add(lwi, lwo);
        float expo = expf(-(wh.x * wh.x / ax2 + wh.y * wh.y / ay2) / (wh.z * wh.z));
        float sq = sqrtf(lwi.z * lwo.z);
        v3 a = scl(kd, INV_PI);
        v3 b = scl(mk(ks.x * inv4 / sq, ks.y * inv4 / sq, ks.z * inv4 / sq), expo);
        st(out, add(a, b));
        return;
    }
    if (dot(wi, w) <= ZERO_EPS || dot(wo, w) <= ZERO_EPS) return;   /* lambertian.cc:63-68 */
    v3 brdf = scl(kd, INV_PI);
    if (type == ORC_PHONG) {                                  /* phong.cc:74-87 */
        float ne = m[7];
        v3 wr = sub(scl(w, 2.0f * dot(wi, w)), wi);
        float cos_alpha = dot(wo, wr);
        if (cos_alpha > 0.0f) {
            /* ((ks * (n+2)) * INV_2PI) * pow(cos_alpha, n) */
            v3 spec = scl(scl(scl(ks, ne + 2.0f), INV_2PI), powf(cos_alpha, ne));
            brdf = add(brdf, spec);
        }
    }
    st(out, brdf);
}

void orc_bsdf_eval(const float mat[ORC_MAT_STRIDE], const float uvn[9], const float wo[3],
                   const float wi[3], float out[3]) {
    bsdf_eval(mat, ld(uvn), ld(uvn + 3), ld(uvn + 6), ld(wo), ld(wi), out);
}

/* form_factor.h:11-20 */
static float ff_abs(v3 p, v3 pn, v3 q, v3 qn) {
    v3 pq = sub(q, p);
    float d1 = dot(pn, pq);
    float d2 = -dot(qn, pq);
    float dpq = dot(pq, pq);
    return fabsf((d1 * d2) / (dpq * dpq));
}

float orc_form_factor_abs(const float p[3], const float pn[3], const float q[3], const float qn[3]) {
    return ff_abs(ld(p), ld(pn), ld(q), ld(qn));
}

static inline float rgb_value(v3 c) { return sqrtf(c.x * c.x + c.y * c.y + c.z * c.z); } /* rgb.h:85 */

/* vpl_interface.h:50-123, vpl.cc:214-230 */
static v3 clamp_term(int mode, float thr, v3 power, v3 fs, v3 fl, float G) {
    switch (mode) {
    case ORC_CLAMP_FORM_FACTOR:
        if (G > thr) return mulv(scl(mulv(fl, fs), thr), power);
        return mulv(scl(mulv(fl, fs), G), power);
    case ORC_CLAMP_REFLECTIVITY: {
        v3 refl = mulv(scl(fs, G), fl);
        float lum = rgb_value(refl);
        if (lum > thr) { float w = thr / lum; return mulv(scl(refl, w), power); }
        return mulv(refl, power);
    }
    case ORC_CLAMP_RADIANCE: {
        v3 rnc = mulv(mulv(scl(power, G), fs), fl);
        float b = rgb_value(rnc);
        if (b > thr) { float w = thr / b; return scl(rnc, w); }
        return rnc;
    }
    default:
        return mulv(mulv(scl(power, G), fs), fl);
    }
}

void orc_clamp_term(int mode, float threshold, const float power[3], const float fs[3],
                    const float fl[3], float G, float out[3]) {
    st(out, clamp_term(mode, threshold, ld(power), ld(fs), ld(fl), G));
}

/* brdf_point_light.h:235-240: white on the light, else m->eval(dg, wi_stored, wo) */
static v3 vpl_brdf(const orc_scene *s, const float *vpl, int vmat, v3 wo) {
    if (vmat < 0) return mk(1.f, 1.f, 1.f);
    float out[3];
    bsdf_eval(s->mat + ORC_MAT_STRIDE * vmat, ld(vpl + 6), ld(vpl + 9), ld(vpl + 12), ld(vpl + 15), wo, out);
    return ld(out);
}

/* vpl.cc:661-688 */
static int gather_point(const orc_scene *s, const float *rc, int rmat, const float *vpl, int vmat,
                        int mode, float thr, float tmin, v3 *out) {
    *out = mk(0.f, 0.f, 0.f);
    if (rmat < 0) return 0;
    v3 p = ld(rc), pn = ld(rc + 6), q = ld(vpl), qn = ld(vpl + 3);
    if (segment_blocked(s, p, q, tmin, NULL)) return 0;
    float G = ff_abs(p, pn, q, qn);
    v3 wi = unit(sub(q, p));
    float fs[3];
    bsdf_eval(s->mat + ORC_MAT_STRIDE * rmat, ld(rc + 12), ld(rc + 15), ld(rc + 18), ld(rc + 9), wi, fs);
    v3 fl = vpl_brdf(s, vpl, vmat, neg(wi));
    *out = clamp_term(mode, thr, ld(vpl + 18), ld(fs), fl, G);
    return 1;
}

int orc_gather_point(const orc_scene *s, const float recv[ORC_RECV_STRIDE], int recv_mat,
                     const float vpl[ORC_VPL_STRIDE], int vpl_mat, int clamp_mode, float clamp,
                     float tmin, float out[3]) {
    v3 o;
    int vis = gather_point(s, recv, recv_mat, vpl, vpl_mat, clamp_mode, clamp, tmin, &o);
    st(out, o);
    return vis;
}

/* vpl.cc:411-440 (geometric normal, surface BRDF = white, no visibility) */
static v3 incident_radiance(const orc_scene *s, const float *rc, const float *vpl, int vmat,
                            int mode, float thr) {
    v3 p = ld(rc), pn = ld(rc + 3), q = ld(vpl), qn = ld(vpl + 3);
    float G = ff_abs(p, pn, q, qn);
    v3 wi = unit(sub(q, p));
    v3 fl = vpl_brdf(s, vpl, vmat, neg(wi));
    return clamp_term(mode, thr, ld(vpl + 18), mk(1.f, 1.f, 1.f), fl, G);
}

void orc_incident_radiance(const orc_scene *s, const float recv[ORC_RECV_STRIDE],
                           const float vpl[ORC_VPL_STRIDE], int vpl_mat, int clamp_mode,
                           float clamp, float out[3]) {
    st(out, incident_radiance(s, recv, vpl, vpl_mat, clamp_mode, clamp));
}

/* --- threading helper (the CPU baseline runs on all host cores) --- */
typedef struct { void (*fn)(void *, int, int); void *arg; int lo, hi; } job_t;
static void *job_run(void *p) { job_t *j = p; j->fn(j->arg, j->lo, j->hi); return NULL; }
static void parallel_for(int n, int nthreads, void (*fn)(void *, int, int), void *arg) {
    if (nthreads <= 1 || n < 2) { fn(arg, 0, n); return; }
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256]; job_t jb[256];
    int chunk = (n + nthreads - 1) / nthreads;
    int used = 0;
    for (int t = 0; t < nthreads; ++t) {
        int lo = t * chunk, hi = lo + chunk > n ? n : lo + chunk;
        if (lo >= hi) break;
        jb[t].fn = fn; jb[t].arg = arg; jb[t].lo = lo; jb[t].hi = hi;
        pthread_create(&th[t], NULL, job_run, &jb[t]);
        used++;
    }
    for (int t = 0; t < used; ++t) pthread_join(th[t], NULL);
}

typedef struct {
    const orc_scene *s; const float *recv; const int *rmat; int nvpl; const float *vpl; const int *vmat;
    const float *w; int mode; float thr, tmin, nem; float *out;
} gather_args;

/* Σ_k w_k gather_point / num_emitted: LightSlice::gather_clusters (lightslice.cc:1325-1339,
 * radiance += weight * Lo after the clamp inside gather_point), vpl.cc:488-489 for the
 * division; without weights every VPL is its own weight-1 cluster */
static void gather_range(void *a_, int lo, int hi) {
    gather_args *a = a_;
    for (int i = lo; i < hi; ++i) {
        const float *rc = a->recv + ORC_RECV_STRIDE * i;
        v3 acc = mk(0.f, 0.f, 0.f);
        if (a->rmat[i] >= 0) {
            for (int k = 0; k < a->nvpl; ++k) {
                v3 lo3;
                if (gather_point(a->s, rc, a->rmat[i], a->vpl + ORC_VPL_STRIDE * k, a->vmat[k], a->mode, a->thr, a->tmin, &lo3))
                    acc = add(acc, scl(lo3, a->w ? a->w[k] : 1.0f));
            }
        }
        a->out[3 * i] = acc.x / a->nem; a->out[3 * i + 1] = acc.y / a->nem; a->out[3 * i + 2] = acc.z / a->nem;
    }
}

void orc_gather(const orc_scene *s, int npts, const float *recv, const int *recv_mat,
                int nvpl, const float *vpl, const int *vpl_mat, int clamp_mode, float clamp,
                float tmin, float n_emitted, float *out_rgb, int nthreads) {
    gather_args a = {s, recv, recv_mat, nvpl, vpl, vpl_mat, NULL, clamp_mode, clamp, tmin, n_emitted, out_rgb};
    parallel_for(npts, nthreads, gather_range, &a);
}

void orc_gather_weighted(const orc_scene *s, int npts, const float *recv, const int *recv_mat,
                         int nvpl, const float *vpl, const int *vpl_mat, const float *vpl_w, int clamp_mode,
                         float clamp, float tmin, float n_emitted, float *out_rgb, int nthreads) {
    gather_args a = {s, recv, recv_mat, nvpl, vpl, vpl_mat, vpl_w, clamp_mode, clamp, tmin, n_emitted, out_rgb};
    parallel_for(npts, nthreads, gather_range, &a);
}

/* ======================================================================= */
/* Reduced matrix, alphas, assignment, cluster_by_sampling                  */
/* ======================================================================= */
typedef struct {
    const orc_scene *s; int rows; const float *recv; const int *rmat; int nvpl; const float *vpl;
    const int *vmat; int mode; float thr, tmin; float *R, *V;
} rm_args;

/* lightslice.cc:721-748; R column-major Rgb: R[(k*rows + r)*3 + c] */
static void rm_range(void *a_, int lo, int hi) {
    rm_args *a = a_;
    for (int r = lo; r < hi; ++r) {
        const float *rc = a->recv + ORC_RECV_STRIDE * r;
        for (int k = 0; k < a->nvpl; ++k) {
            size_t idx = (size_t)k * a->rows + r;
            if (a->rmat[r] < 0) { a->R[3 * idx] = a->R[3 * idx + 1] = a->R[3 * idx + 2] = 0.f; a->V[idx] = 0.f; continue; }
            const float *vp = a->vpl + ORC_VPL_STRIDE * k;
            int visible = !segment_blocked(a->s, ld(rc), ld(vp), a->tmin, NULL);
            v3 inc = incident_radiance(a->s, rc, vp, a->vmat[k], a->mode, a->thr);
            float fv = (float)visible;
            a->R[3 * idx] = inc.x * fv; a->R[3 * idx + 1] = inc.y * fv; a->R[3 * idx + 2] = inc.z * fv;
            a->V[idx] = fv;
        }
    }
}

void orc_reduced_matrix(const orc_scene *s, int rows, const float *recv, const int *recv_mat,
                        int nvpl, const float *vpl, const int *vpl_mat, int clamp_mode,
                        float clamp, float tmin, float *R, float *V, int nthreads) {
    rm_args a = {s, rows, recv, recv_mat, nvpl, vpl, vpl_mat, clamp_mode, clamp, tmin, R, V};
    parallel_for(rows, nthreads, rm_range, &a);
}

/* light_transport_matrix.cc:72-106 with matrix_c.cc:48-59 flatten_magnitude,
 * matrix_c.h:224-245 mul / row_sum (sequential orders kept) */
void orc_alphas(int rows, int cols, const float *R, float *alphas) {
    float *m = malloc(sizeof(float) * (size_t)rows * cols);
    for (int j = 0; j < cols; ++j)
        for (int i = 0; i < rows; ++i) {
            const float *e = R + 3 * ((size_t)j * rows + i);
            m[(size_t)j * rows + i] = sqrtf(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
        }
    float *norms = malloc(sizeof(float) * cols);
    for (int j = 0; j < cols; ++j) {
        float sum = 0.f;
        for (int i = 0; i < rows; ++i) sum += m[(size_t)j * rows + i] * m[(size_t)j * rows + i];
        norms[j] = sqrtf(sum);
    }
    float *u = calloc(rows, sizeof(float));
    for (int j = 0; j < cols; ++j)
        for (int i = 0; i < rows; ++i) u[i] += m[(size_t)j * rows + i];
    /* v = Rt * u: Rt(j, i) = m(i, j); out[j] += Rt(j,i) * u[i], i outer */
    float *v = calloc(cols, sizeof(float));
    for (int i = 0; i < rows; ++i)
        for (int j = 0; j < cols; ++j) v[j] += m[(size_t)j * rows + i] * u[i];
    float c0 = 0.f;
    for (int j = 0; j < cols; ++j) c0 += norms[j];
    for (int j = 0; j < cols; ++j) alphas[j] = norms[j] * c0 - v[j];
    free(m); free(norms); free(u); free(v);
}

/* cluster_builder.cc:14-19 magnitude() over the flattened 3*rows column */
void orc_column_magnitudes(int rows, int cols, const float *R, float *mags) {
    for (int j = 0; j < cols; ++j) {
        float sum = 0.f;
        const float *c = R + 3 * (size_t)j * rows;
        for (int t = 0; t < 3 * rows; ++t) sum += c[t] * c[t];
        mags[j] = sqrtf(sum);
    }
}

typedef struct {
    int rows, cols, ncl; const float *R; const float *K; const float *Knorm; int *assign;
} as_args;

/* cluster_builder.cc:199-225 */
static void as_range(void *a_, int lo, int hi) {
    as_args *a = a_;
    int n3 = 3 * a->rows;
    for (int i = lo; i < hi; ++i) {
        const float *mi = a->R + (size_t)n3 * i;
        float s2 = 0.f;
        for (int t = 0; t < n3; ++t) s2 += mi[t] * mi[t];
        float norm_mi = sqrtf(s2);
        int min_c = -1; float min_d = FLT_MAX;
        for (int j = 0; j < a->ncl; ++j) {
            const float *mk_ = a->K + (size_t)n3 * j;
            float d = 0.f;
            for (int t = 0; t < n3; ++t) d += mi[t] * mk_[t];
            float dist = fabsf(a->Knorm[j] * norm_mi - d);
            if (dist < min_d) { min_c = j; min_d = dist; }
        }
        a->assign[i] = min_c;
    }
}

static void build_K(int rows, const float *R, int ncl, const int *centers, const float *weights,
                    float *K, float *Knorm) {
    int n3 = 3 * rows;
    for (int j = 0; j < ncl; ++j) {
        const float *c = R + (size_t)n3 * centers[j];
        float *kj = K + (size_t)n3 * j;
        for (int t = 0; t < n3; ++t) kj[t] = c[t] * weights[j];          /* cluster_builder.cc:166-168 */
        float mag = 0.f;                                                  /* light_transport_matrix.cc:117-130 */
        for (int r = 0; r < rows; ++r)
            mag += kj[3 * r] * kj[3 * r] + kj[3 * r + 1] * kj[3 * r + 1] + kj[3 * r + 2] * kj[3 * r + 2];
        Knorm[j] = sqrtf(mag);
    }
}

void orc_cluster_assign(int rows, int cols, const float *R, int ncl, const int *centers,
                        const float *weights, int *assign, int nthreads) {
    float *K = malloc(sizeof(float) * 3 * (size_t)rows * ncl);
    float *Kn = malloc(sizeof(float) * ncl);
    build_K(rows, R, ncl, centers, weights, K, Kn);
    as_args a = {rows, cols, ncl, R, K, Kn, assign};
    parallel_for(cols, nthreads, as_range, &a);
    free(K); free(Kn);
}

/* pdf.cc:18-52; returns 1 when sampleable */
int orc_pdf_build(int n, const float *arr, float *pdf, float *cdf) {
    if (n == 0) return 0;
    float sum = 0.f;
    for (int i = 0; i < n; ++i) sum += arr[i];
    if (sum == 0.0f) { for (int i = 0; i < n; ++i) pdf[i] = cdf[i] = 0.f; return 0; }
    for (int i = 0; i < n; ++i) pdf[i] = arr[i] / sum;
    cdf[0] = pdf[0];
    for (int i = 1; i < n; ++i) cdf[i] = cdf[i - 1] + pdf[i];
    cdf[n - 1] = 1.0f;
    return 1;
}

/* pdf.cc:59-77 (linear scan, strict <) */
void orc_pdf_sample(int n, const float *pdf, const float *cdf, float u, int *k, float *pk) {
    for (int i = 0; i < n; ++i)
        if (u < cdf[i]) { *k = i; *pk = pdf[i]; return; }
    *k = -1; *pk = 0.f;
}

/* cluster_builder.cc:82-273 with the global MT stream.  Returns the number of
 * clusters kept.  centers/center_w: sampled centres (ncenters <= num_clusters);
 * assign[V]: cluster index per VPL (pre-removal numbering); reps/rep_w/kept
 * are indexed by the pre-removal cluster index and kept[c] = 1 when the
 * cluster survives removal.  The reference loops forever when fewer than
 * num_clusters columns have alpha > 0; the restatement stops after
 * 1000*num_clusters draws and returns the clusters found so far. */
int orc_cluster_by_sampling(int rows, int cols, const float *R, int num_clusters,
                            int *centers, float *center_w, int *assign, int *reps,
                            float *rep_w, int *kept, int nthreads) {
    if (cols < num_clusters) return -1;
    float *alphas = malloc(sizeof(float) * cols);
    orc_alphas(rows, cols, R, alphas);
    float *pdf = malloc(sizeof(float) * cols), *cdf = malloc(sizeof(float) * cols);
    orc_pdf_build(cols, alphas, pdf, cdf);
    int count = 0;
    long attempts = 0, max_attempts = 1000L * num_clusters;
    while (count < num_clusters && attempts < max_attempts) {
        attempts++;
        int k; float pk;
        orc_pdf_sample(cols, pdf, cdf, orc_mt_rand(), &k, &pk);
        if (pk == 0.0f) continue;
        int idx = -1;
        for (int i = 0; i < count; ++i) if (centers[i] == k) { idx = i; break; }
        if (idx < 0) { centers[count] = k; center_w[count] = 1.0f / pk; count++; }
        else center_w[idx] += 1.0f / pk;
    }
    int ncl = count;
    orc_cluster_assign(rows, cols, R, ncl, centers, center_w, assign, nthreads);
    float *mags = malloc(sizeof(float) * cols);
    orc_column_magnitudes(rows, cols, R, mags);
    /* members in increasing VPL order (cluster_builder.cc:223, :231-239) */
    int *cnt = calloc(ncl + 1, sizeof(int));
    for (int i = 0; i < cols; ++i) if (assign[i] >= 0) cnt[assign[i]]++;
    int *off = calloc(ncl + 1, sizeof(int));
    for (int c = 0; c < ncl; ++c) off[c + 1] = off[c] + cnt[c];
    int *mem = malloc(sizeof(int) * (cols + 1));
    int *fill = calloc(ncl + 1, sizeof(int));
    for (int i = 0; i < cols; ++i) if (assign[i] >= 0) { int c = assign[i]; mem[off[c] + fill[c]++] = i; }
    int nkept = 0;
    for (int c = 0; c < ncl; ++c) {                                   /* :242-262 */
        int m = cnt[c];
        float *mm = malloc(sizeof(float) * (m + 1)), *pp = malloc(sizeof(float) * (m + 1)), *cc = malloc(sizeof(float) * (m + 1));
        for (int t = 0; t < m; ++t) mm[t] = mags[mem[off[c] + t]];
        int ok = orc_pdf_build(m, mm, pp, cc);
        float u = orc_mt_rand();
        int j = -1; float pj = 0.f;
        if (m > 0 && ok) orc_pdf_sample(m, pp, cc, u, &j, &pj);
        else if (m > 0) orc_pdf_sample(m, pp, cc, u, &j, &pj);   /* zero cdf: never < */
        if (j < 0) {
            reps[c] = m > 0 ? mem[off[c] + m / 2] : -1;
            rep_w[c] = 0.0f;
        } else {
            reps[c] = mem[off[c] + j];
            rep_w[c] = 1.0f / pj;
        }
        kept[c] = reps[c] >= 0;
        nkept += kept[c];
        free(mm); free(pp); free(cc);
    }
    free(alphas); free(pdf); free(cdf); free(mags); free(cnt); free(off); free(mem); free(fill);
    return ncl * 65536 + nkept;
}

/* ======================================================================= */
/* Concentric map and per-slice guiding distributions                       */
/* ======================================================================= */
/* Transcendentals that pick a ray direction, restated from the mathematics in
 * fp32 add, multiply, divide and square root only (each correctly rounded, no
 * contraction), in a fixed order, so the kernels' restatement of the same steps
 * (rzk_math.h) gives the same bits as these for every argument.  Accuracy
 * against the exact functions (tests/test_oracle_xmath.py): a few ulp on the
 * angles and arguments the samplers use.
 *
 *   xsincos  |x| <= 1e4: k = nearest integer of x * 2/pi, r = x - k pi/2 with
 *            pi/2 in three parts of which k times each is exact (Cody-Waite),
 *            Taylor polynomials of degree 9 (sin) and 10 (cos) on |r| <= pi/4
 *            (truncation below 2e-9), the quadrant k mod 4 picking and signing
 *            the pair
 *   xatan    |x| > 1 through pi/2 - atan(1/|x|), |x| > tan(pi/8) through
 *            pi/4 + atan((|x| - 1) / (|x| + 1)), Taylor series to degree 21
 *   xatan2   the quadrant from the signs, atan of the ratio of the smaller to
 *            the larger magnitude (the C library's signed-zero cases)
 *   xacos    |x| <= 1/2: pi/2 - asin(x); else 2 asin(sqrt((1 - |x|) / 2)),
 *            reflected for x < 0; asin's Taylor series to degree 21 on [0, 1/2]
 *   xlog     x = m 2^e with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s),
 *            s = (m - 1) / (m + 1), series to s^9; e ln 2 with ln 2 in two parts
 *   xexp     k = nearest integer of t / ln 2, r = t - k ln 2 (two parts),
 *            Taylor to r^8, times 2^k from the exponent bits
 *   xpow     x >= 0: exp(y log x) */
#define XM_PIO2_1   0x1.921ep+0f          /* pi/2 = PIO2_1 + PIO2_2 + PIO2_3 (+ 2^-58) */
#define XM_PIO2_2   0x1.b544p-16f
#define XM_PIO2_3   0x1.0b46p-34f
#define XM_PIO2_LO  0x1.b54442p-16f      /* pi/2 - PIO2_1 */
#define XM_PIO4_HI  0x1.921ep-1f
#define XM_PIO4_LO  0x1.b54442p-17f
#define XM_PI_HI    0x1.921ep+1f
#define XM_PI_LO    0x1.b54442p-15f
#define XM_LN2_HI   0x1.62e4p-1f
#define XM_LN2_LO   0x1.7f7d1cp-20f

static inline int xm_round(float v) { return (int)(v + (v >= 0.0f ? 0.5f : -0.5f)); }

static void xsincos(float x, float *s, float *c) {
    if (!(fabsf(x) <= 1e4f)) { *s = *c = (x - x) / (x - x); return; }
    float kf = x * 0x1.45f306p-1f;
    int k = xm_round(kf);
    float fk = (float)k;
    float r = ((x - fk * XM_PIO2_1) - fk * XM_PIO2_2) - fk * XM_PIO2_3;
    float r2 = r * r;
    float sr = r + r * r2 * (-0x1.555556p-3f + r2 * (0x1.111112p-7f + r2 * (-0x1.a01a02p-13f + r2 * 0x1.71de3ap-19f)));
    float cr = 1.0f + r2 * (-0.5f + r2 * (0x1.555556p-5f + r2 * (-0x1.6c16c2p-10f + r2 * (0x1.a01a02p-16f +
                                                                                       r2 * -0x1.27e4fcp-22f))));
    switch (k & 3) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
    }
}
static inline float xsin(float x) { float s, c; xsincos(x, &s, &c); return s; }
static inline float xcos(float x) { float s, c; xsincos(x, &s, &c); return c; }
static inline float xtan(float x) { float s, c; xsincos(x, &s, &c); return s / c; }

static float xatan(float x) {
    if (x != x) return x;
    float a = fabsf(x);
    int inv = a > 1.0f, quarter;
    if (inv) a = 1.0f / a;
    quarter = a > 0x1.a8279ap-2f;
    if (quarter) a = (a - 1.0f) / (a + 1.0f);
    float z2 = a * a;
    float p = a + a * z2 * (-0x1.555556p-2f + z2 * (0x1.99999ap-3f + z2 * (-0x1.24924ap-3f + z2 * (0x1.c71c72p-4f +
              z2 * (-0x1.745d18p-4f + z2 * (0x1.3b13b2p-4f + z2 * (-0x1.111112p-4f + z2 * (0x1.e1e1e2p-5f +
              z2 * (-0x1.af286cp-5f + z2 * 0x1.861862p-5f)))))))));
    if (quarter) p = XM_PIO4_HI + (p + XM_PIO4_LO);
    if (inv) p = (XM_PIO2_1 - p) + XM_PIO2_LO;
    return copysignf(p, x);
}

static float xatan2(float y, float x) {
    if (x != x || y != y) return x + y;
    float ax = fabsf(x), ay = fabsf(y), t;
    if (ay == 0.0f && ax == 0.0f) t = signbit(x) ? XM_PI_HI + XM_PI_LO : 0.0f;
    else if (ay <= ax) {
        t = xatan(ay / ax);
        if (signbit(x)) t = (XM_PI_HI - t) + XM_PI_LO;
    } else {
        float u = xatan(ax / ay);
        t = signbit(x) ? (XM_PIO2_1 + u) + XM_PIO2_LO : (XM_PIO2_1 - u) + XM_PIO2_LO;
    }
    return copysignf(t, y);
}

static inline float xasin_series(float z) {     /* |z| <= 1/2 */
    float z2 = z * z;
    return z + z * z2 * (0x1.555556p-3f + z2 * (0x1.333334p-4f + z2 * (0x1.6db6dcp-5f + z2 * (0x1.f1c71cp-6f +
           z2 * (0x1.6e8ba2p-6f + z2 * (0x1.1c4ec4p-6f + z2 * (0x1.c9999ap-7f + z2 * (0x1.7a8788p-7f +
           z2 * (0x1.3fde50p-7f + z2 * 0x1.12ef3cp-7f)))))))));
}

static float xacos(float x) {
    float a = fabsf(x);
    if (!(a <= 1.0f)) return (x - x) / (x - x);
    if (a <= 0.5f) return (XM_PIO2_1 - xasin_series(x)) + XM_PIO2_LO;
    float t = 2.0f * xasin_series(sqrtf((1.0f - a) * 0.5f));
    return x > 0.0f ? t : (XM_PI_HI - t) + XM_PI_LO;
}

static float xlog(float x) {
    if (x != x || x < 0.0f) return (x - x) / (x - x);
    if (x == 0.0f) return -INFINITY;
    if (x == INFINITY) return x;
    uint32_t ix;
    memcpy(&ix, &x, 4);
    int e = (int)(ix >> 23) - 127;
    if (e == -127) {                                   /* subnormal */
        float xs = x * 0x1p23f;
        memcpy(&ix, &xs, 4);
        e = (int)(ix >> 23) - 127 - 23;
    }
    ix = (ix & 0x007fffffu) | 0x3f800000u;
    float m;
    memcpy(&m, &ix, 4);
    if (m > 0x1.6a09e6p+0f) { m = m * 0.5f; e += 1; }
    float sv = (m - 1.0f) / (m + 1.0f);
    float s2 = sv * sv;
    float lm = 2.0f * sv + sv * s2 * (0x1.555556p-1f + s2 * (0x1.99999ap-2f + s2 * (0x1.24924ap-2f + s2 * 0x1.c71c72p-3f)));
    float fe = (float)e;
    return fe * XM_LN2_HI + (lm + fe * XM_LN2_LO);
}

static float xexp(float t) {
    if (t != t) return t;
    if (t > 88.7f) return INFINITY;
    if (t < -104.0f) return 0.0f;
    int k = xm_round(t * 0x1.715476p+0f);
    float fk = (float)k;
    float r = (t - fk * XM_LN2_HI) - fk * XM_LN2_LO;
    float p = 1.0f + (r + r * r * (0.5f + r * (0x1.555556p-3f + r * (0x1.555556p-5f + r * (0x1.111112p-7f +
              r * (0x1.6c16c2p-10f + r * (0x1.a01a02p-13f + r * 0x1.a01a02p-16f)))))));
    if (k < -126) { p = p * 0x1p-64f; k += 64; }
    uint32_t b = (uint32_t)(k + 127) << 23;
    float sc;
    memcpy(&sc, &b, 4);
    return p * sc;
}

static float xpow(float x, float y) {              /* x >= 0 */
    if (x == 1.0f || y == 0.0f) return 1.0f;
    if (x == 0.0f) return y > 0.0f ? 0.0f : INFINITY;
    return xexp(y * xlog(x));
}

/* test hook: fn 0 sin, 1 cos, 2 tan, 3 atan, 4 atan2(x, y), 5 acos, 6 log, 7 exp, 8 pow(x, y) */
void orc_xmath(int fn, int n, const float *x, const float *y, float *out) {
    for (int i = 0; i < n; ++i) {
        float a = x[i], b = y ? y[i] : 0.0f, v = 0.0f;
        switch (fn) {
        case 0: v = xsin(a); break;
        case 1: v = xcos(a); break;
        case 2: v = xtan(a); break;
        case 3: v = xatan(a); break;
        case 4: v = xatan2(a, b); break;
        case 5: v = xacos(a); break;
        case 6: v = xlog(a); break;
        case 7: v = xexp(a); break;
        default: v = xpow(a, b); break;
        }
        out[i] = v;
    }
}

/* concentric_mapping.h:16-47 */
static void square_to_disk(float sx, float sy, float *u, float *v) {
    float phi, r;
    float a = 2.0f * sx - 1.0f, b = 2.0f * sy - 1.0f;
    if (a > -b) {
        if (a > b) { r = a; phi = (A_PI / 4) * (b / a); }
        else { r = b; phi = (A_PI / 4) * (2 - (a / b)); }
    } else {
        if (a < b) { r = -a; phi = (A_PI / 4) * (4 + (b / a)); }
        else { r = -b; if (b != 0.0f) phi = (A_PI / 4) * (6 - (a / b)); else phi = 0.0f; }
    }
    *u = r * xcos(phi);
    *v = r * xsin(phi);
}

/* concentric_mapping.h:49-74 */
static void disk_to_square(float dx, float dy, float *x, float *y) {
    float r = sqrtf(dx * dx + dy * dy);
    float phi = xatan2(dy, dx);
    if (phi < -A_PI / 4) phi += TWO_PI;
    float a, b;
    if (phi < A_PI / 4) { a = r; b = phi * a / (A_PI / 4); }
    else if (phi < 3 * A_PI / 4) { b = r; a = -(phi - A_PI / 2) * b / (A_PI / 4); }
    else if (phi < 5 * A_PI / 4) { a = -r; b = (phi - A_PI) * a / (A_PI / 4); }
    else { b = -r; a = -(phi - 3 * A_PI / 2) * b / (A_PI / 4); }
    *x = a * 0.5f + 0.5f;
    *y = b * 0.5f + 0.5f;
}

/* concentric_mapping.h:83-114 */
void orc_direction_to_square(const float wi[3], float out[2]) {
    float r2 = 1.0f - wi[2];
    float dx = wi[0] / sqrtf(2.0f - r2), dy = wi[1] / sqrtf(2.0f - r2);
    disk_to_square(dx, dy, &out[0], &out[1]);
}

void orc_square_to_direction(const float s[2], float out[3]) {
    float dx, dy;
    square_to_disk(s[0], s[1], &dx, &dy);
    float r2 = dx * dx + dy * dy;
    out[0] = dx * sqrtf(2.0f - r2);
    out[1] = dy * sqrtf(2.0f - r2);
    out[2] = 1.0f - r2;
}

/* lightslice.cc:1011-1030 subset for the cluster representatives: slice_wi and
 * the upper-hemisphere flag (visible, above the shading normal, non-black). */
void orc_slice_cluster_info(const orc_scene *s, const float recv[ORC_RECV_STRIDE], int nreps,
                            const int *reps, const float *vpl, const int *vpl_mat,
                            const float *Vrow, int clamp_mode, float clamp,
                            float *slice_wi, int *upper) {
    for (int c = 0; c < nreps; ++c) {
        int k = reps[c];
        const float *vp = vpl + ORC_VPL_STRIDE * k;
        float vis = Vrow[k];
        v3 inc = incident_radiance(s, recv, vp, vpl_mat[k], clamp_mode, clamp);
        v3 wi = unit(sub(ld(vp), ld(recv)));
        st(slice_wi + 3 * c, wi);
        int nonblack = inc.x != 0.f || inc.y != 0.f || inc.z != 0.f;
        upper[c] = vis > 0.0f && dot(ld(recv + 6), wi) > ZERO_EPS && nonblack;
    }
}

/* lightslice.cc:1114-1179: norms of the representative columns over the
 * neighbour rows (matrix_c.cc:74-88 order) x weight, zero below the horizon */
void orc_slice_cluster_pdf(int rows, const float *R, int nnb, const int *nb_rows, int nreps,
                           const int *reps, const float *rep_w, const int *upper, float *out) {
    for (int c = 0; c < nreps; ++c) {
        float mag = 0.f;
        for (int t = 0; t < nnb; ++t) {
            const float *e = R + 3 * ((size_t)reps[c] * rows + nb_rows[t]);
            mag += e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
        }
        float v = sqrtf(mag);
        v *= rep_w[c];
        if (!upper[c]) v = 0.0f;
        out[c] = v;
    }
}

/* lightslice.cc:2058-2147 build_map + pdf.cc:195-233 DiscretePdf2D.  R_rep_rgb
 * holds R(s, rep_c) per cluster.  Returns 1 when the 2-D pdf is valid. */
int orc_build_map(int res, int nreps, const float *slice_wi, const int *upper,
                  const float *rep_w, const float *R_rep_rgb, const float uvn[9], float eps,
                  int valid, float *map, float *row_cdf, float *row_pdf, float *col_cdf,
                  float *total) {
    v3 u = ld(uvn), v = ld(uvn + 3), w = ld(uvn + 6);
    for (int i = 0; i < res * res; ++i) map[i] = 0.f;
    float sum = 0.f;
    if (valid) {
        for (int c = 0; c < nreps; ++c) {
            if (!upper[c]) continue;
            v3 wi = ld(slice_wi + 3 * c);
            float lwi[3] = {dot(u, wi), dot(v, wi), dot(w, wi)};
            if (lwi[2] <= 0.0f) continue;
            v3 rad = scl(ld(R_rep_rgb + 3 * c), rep_w[c]);
            float sq[2];
            orc_direction_to_square(lwi, sq);
            int x = (int)(sq[0] * res), y = (int)(sq[1] * res);
            if (x >= res) x = res - 1;
            if (y >= res) y = res - 1;
            if (x < 0) x = 0;
            if (y < 0) y = 0;
            float lum = rgb_value(rad);
            map[y * res + x] += lum;
            sum += lum;
        }
        for (int j = 0; j < res; ++j)
            for (int i = 0; i < res; ++i) map[i * res + j] += eps * sum;
    }
    float tot = 0.f;
    float *rs = malloc(sizeof(float) * res);
    for (int i = 0; i < res; ++i) {
        rs[i] = 0.f;
        for (int j = 0; j < res; ++j) rs[i] += map[i * res + j];
        tot += rs[i];
        float s2 = 0.f;
        for (int j = 0; j < res; ++j) { s2 += map[i * res + j]; col_cdf[i * res + j] = s2 / rs[i]; }
    }
    orc_pdf_build(res, rs, row_pdf, row_cdf);
    free(rs);
    *total = tot;
    return tot > 0.0f;
}


/* ======================================================================= */
/* Guided path tracing                                                      */
/* ======================================================================= */
/* light tables: emitting triangles per light in surface order, triangle
 * choice by cumulative area ratio (quad.cc:75-87), light choice by a
 * DiscretePdf over |Le| x area */
void orc_scene_set_lights(orc_scene *s, int nlights, const float *radiance) {
    free(s->lights); free(s->ltris);
    s->nlights = nlights;
    s->lights = calloc(nlights + 1, sizeof(light_t));
    s->ltris = calloc(s->ntri + 1, sizeof(ltri_t));
    s->nltri = 0;
    for (int li = 0; li < nlights; ++li) {
        light_t *L = &s->lights[li];
        L->le = ld(radiance + 3 * li);
        L->first = s->nltri;
        float total = 0.f;
        for (int k = 0; k < s->ntri; ++k) {
            int sid = s->tri_surface[k];
            if (s->surf_light[sid] != li) continue;
            v3 cr = cross(sub(s->p0[k], s->p1[k]), sub(s->p0[k], s->p2[k]));
            ltri_t *t = &s->ltris[s->nltri++];
            t->p0 = s->p0[k]; t->p1 = s->p1[k]; t->p2 = s->p2[k];
            t->area = len(cr) * 0.5f;                       /* triangle.cc:15 */
            t->n = unit(cr);
            if (s->surf_flags[sid] & ORC_SURF_FLIPPED) t->n = neg(t->n);
            t->light = li; t->tri = k;
            total += t->area;
        }
        L->count = s->nltri - L->first;
        float cum = 0.f;
        for (int q = 0; q < L->count; ++q) {
            ltri_t *t = &s->ltris[L->first + q];
            t->ratio = t->area / total;
            cum += t->ratio;
            t->cum = q == L->count - 1 ? 1.0f : cum;
        }
        L->sel_cdf = total * rgb_value(L->le);
    }
    float sum = 0.f;
    for (int i = 0; i < nlights; ++i) sum += s->lights[i].sel_cdf;
    float cum = 0.f;
    for (int i = 0; i < nlights; ++i) {
        light_t *L = &s->lights[i];
        L->sel_pdf = sum > 0.f ? L->sel_cdf / sum : 0.f;
        cum += L->sel_pdf;
        L->sel_cdf = i + 1 == nlights ? 1.0f : cum;
    }
}

/* Philox4x32-10 (Salmon et al., SC'11) */
void orc_philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

static inline float u24(uint32_t r) { return (float)(r & 0x00ffffffu) * (1.0f / 16777216.0f); }
static inline float u24o(uint32_t r) { return (float)(r & 0x00ffffffu) * (1.0f / 16777216.0f) + (1.0f / 33554432.0f); }

static void path_rng(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t bounce, uint32_t u[12]) {
    for (uint32_t b = 0; b < 3; ++b) {
        uint32_t c[4] = {sample, bounce, b, 0x52a7u};
        orc_philox(c, seed ^ 0x9a3c5e17u, pixel);
        u[4 * b] = c[0]; u[4 * b + 1] = c[1]; u[4 * b + 2] = c[2]; u[4 * b + 3] = c[3];
    }
}

static float rgb_max(v3 c) { float t = c.x; if (c.y > t) t = c.y; if (c.z > t) t = c.z; return t; }

static void lobe_weights(const float *m, float *rd, float *rs) {   /* phong.cc:6-16, ward.cc:6-22 */
    *rd = rgb_max(ld(m + 1));
    *rs = rgb_max(ld(m + 4));
    float total = *rd + *rs;
    if (total > 0.0f) { *rd /= total; *rs /= total; }
}

static v3 l2w(v3 u, v3 v, v3 w, v3 l) { return add(add(scl(u, l.x), scl(v, l.y)), scl(w, l.z)); }   /* onb.cc:117-119 */
static v3 w2l(v3 u, v3 v, v3 w, v3 x) { return mk(dot(u, x), dot(v, x), dot(w, x)); }

static v3 sph2w(v3 u, v3 v, v3 w, float theta, float phi) {       /* onb.cc:155-163 */
    float st = xsin(theta), ct = xcos(theta), sp = xsin(phi), cp = xcos(phi);
    return l2w(u, v, w, mk(st * cp, st * sp, ct));
}

/* lambertian.cc:70-75, phong.cc:89-102, ward.cc:118-134 */
static float bsdf_pdf(const float *m, v3 u, v3 v, v3 w, v3 wo, v3 wi) {
    if (dot(wi, w) <= ZERO_EPS || dot(wo, w) <= ZERO_EPS) return 0.0f;
    int type = (int)m[0];
    if (type == ORC_LAMBERT) return dot(w, wi) * INV_PI;
    float rd, rs;
    lobe_weights(m, &rd, &rs);
    if (type == ORC_PHONG) {
        float ne = m[7];
        float pdf = rd * (dot(w, wi) * INV_PI);
        v3 wr = sub(scl(w, 2.0f * dot(wi, w)), wi);
        float ca = dot(wo, wr);
        if (ca > 0.0f) pdf += rs * (ne + 1.0f) * INV_2PI * powf(ca, ne);
        return pdf;
    }
    float ax = m[7], ay = m[8];
    float ax2 = ax * ax, ay2 = ay * ay;
    float inv4 = 1.0f / (4.0f * A_PI * ax * ay);
    float z = dot(w, wi);
    v3 wh = unit(w2l(u, v, w, add(wi, wo)));
    float expo = expf(-(wh.x * wh.x / ax2 + wh.y * wh.y / ay2) / (wh.z * wh.z));
    float ct = wh.z;
    v3 lwo = unit(w2l(u, v, w, wo));
    float ct3 = ct * ct * ct;
    return rd * (z * INV_PI) + rs * (inv4 / (dot(wh, lwo) * ct3) * expo);
}

void orc_bsdf_pdf(const float mat[ORC_MAT_STRIDE], const float uvn[9], const float wo[3], const float wi[3],
                  float *pdf) {
    *pdf = bsdf_pdf(mat, ld(uvn), ld(uvn + 3), ld(uvn + 6), ld(wo), ld(wi));
}

static v3 eval3(const float *m, v3 u, v3 v, v3 w, v3 wo, v3 wi) {
    float o[3];
    bsdf_eval(m, u, v, w, wo, wi, o);
    return ld(o);
}

/* lambertian.cc:5-27, phong.cc:21-72, ward.cc:24-101 (u[0] lobe, u[1..2] direction) */
static v3 bsdf_sample(const float *m, v3 u, v3 v, v3 w, v3 wo, const uint32_t *r, v3 *wi, float *pdf) {
    *wi = mk(0.f, 0.f, 0.f);
    *pdf = 0.0f;
    if (dot(wo, w) <= ZERO_EPS) return mk(0.f, 0.f, 0.f);
    int type = (int)m[0];
    if (type == ORC_LAMBERT) {
        float phi = TWO_PI * u24(r[1]);
        float st = sqrtf(u24(r[2]));
        float x = st * xcos(phi), y = st * xsin(phi);
        float z = sqrtf(1.0f - st * st);
        *wi = unit(l2w(u, v, w, mk(x, y, z)));
        *pdf = z * INV_PI;
        return scl(ld(m + 1), INV_PI);
    }
    float rd, rs;
    lobe_weights(m, &rd, &rs);
    if (u24(r[0]) < rd) {
        float phi = TWO_PI * u24(r[1]);
        float st = sqrtf(u24(r[2]));
        float x = st * xcos(phi), y = st * xsin(phi);
        float z = sqrtf(1.0f - st * st);
        *wi = unit(l2w(u, v, w, mk(x, y, z)));
    } else if (type == ORC_PHONG) {
        float ne = m[7];
        v3 ou, ov, ow;
        onb_from_nu(wo, w, &ou, &ov, &ow);
        float phi = TWO_PI * u24(r[1]);
        float alpha = xacos(xpow(u24(r[2]), 1.0f / (ne + 1.0f)));
        v3 wr = sph2w(ou, ov, ow, alpha, phi);
        float d = dot(wr, w);
        if (d < 0.0f) return mk(0.f, 0.f, 0.f);
        *wi = sub(scl(w, 2.0f * d), wr);
    } else {
        float ax = m[7], ay = m[8];
        float ax2 = ax * ax, ay2 = ay * ay, ay_ax = ay / ax;
        float s0 = u24o(r[1]), s1 = u24o(r[2]);
        v3 lwo = unit(w2l(u, v, w, wo));
        float angle = TWO_PI * s0;
        float phi = xatan(ay_ax * xtan(angle));
        if (xsin(phi) * xsin(angle) < 0 || xcos(phi) * xcos(angle) < 0) phi += A_PI;
        float cp = xcos(phi), sp = xsin(phi);
        float theta = xatan(sqrtf(-xlog(s1) / (cp * cp / ax2 + sp * sp / ay2)));
        float st = xsin(theta), ct = xcos(theta);
        v3 wh = mk(st * cp, st * sp, ct);
        v3 lwi = sub(scl(wh, 2.0f * dot(lwo, wh)), lwo);
        *wi = l2w(u, v, w, lwi);
    }
    *pdf = bsdf_pdf(m, u, v, w, wo, *wi);
    return eval3(m, u, v, w, wo, *wi);
}

void orc_bsdf_sample(const float mat[ORC_MAT_STRIDE], const float uvn[9], const float wo[3], const uint32_t u[3],
                     float wi[3], float *pdf, float brdf[3]) {
    v3 w_;
    v3 f = bsdf_sample(mat, ld(uvn), ld(uvn + 3), ld(uvn + 6), ld(wo), u, &w_, pdf);
    st(wi, w_);
    st(brdf, f);
}

static v3 light_query(const light_t *L, v3 n, v3 wo) {          /* area_light.cc:19-25 */
    if (dot(wo, n) < ZERO_EPS) return mk(0.f, 0.f, 0.f);
    return L->le;
}

static float balance(float a, float b) { return a / (a + b); } /* mis.h:11-13 */

/* FirstBounce::radiance -> AreaLight::first_bounce (area_light.cc:48-142) */
static v3 next_event(const orc_scene *s, const float *rc, int m, const uint32_t *u, float tmin) {
    float ul = u24(u[1]);
    int k = -1;
    for (int i = 0; i < s->nlights; ++i) if (ul < s->lights[i].sel_cdf) { k = i; break; }
    if (k < 0 || s->lights[k].count <= 0) return mk(0.f, 0.f, 0.f);
    const light_t *L = &s->lights[k];
    const float *mat = s->mat + ORC_MAT_STRIDE * m;
    v3 p = ld(rc), pn = ld(rc + 6), wo = ld(rc + 9), bu = ld(rc + 12), bv = ld(rc + 15), bw = ld(rc + 18);
    v3 Lo = mk(0.f, 0.f, 0.f);
    float ut = u24(u[2]);
    int ti = L->first + L->count - 1;
    for (int q = 0; q < L->count; ++q) if (ut <= s->ltris[L->first + q].cum) { ti = L->first + q; break; }
    const ltri_t *T = &s->ltris[ti];
    float tmp = sqrtf(1.0f - u24(u[3]));
    float bb = 1.0f - tmp, cc = u24(u[4]) * tmp;
    float aa = 1.0f - bb - cc;
    v3 q = add(add(scl(T->p0, aa), scl(T->p1, bb)), scl(T->p2, cc));   /* triangle.cc:298-308 */
    v3 qn = T->n;
    float pdf = T->ratio * (1.0f / T->area);
    if (pdf > 0.0f && !segment_blocked(s, p, q, tmin, NULL)) {
        v3 wi = unit(sub(q, p));
        v3 Le = light_query(L, qn, neg(wi));
        float G = ff_abs(p, pn, q, qn);
        if (G > 0.0f) {
            v3 brdf = eval3(mat, bu, bv, bw, wo, wi);
            float dpdf = bsdf_pdf(mat, bu, bv, bw, wo, wi);
            v3 pq = sub(q, p);
            float w1 = balance(pdf * dot(pq, pq) / fabsf(dot(unit(sub(p, q)), qn)), dpdf);
            Lo = add(Lo, mulv(scl(Le, (w1 * G) / pdf), brdf));
        }
    }
    v3 wi;
    float dpdf;
    bsdf_sample(mat, bu, bv, bw, wo, u + 5, &wi, &dpdf);
    if (dpdf > 0.0f) {
        ray_t r2 = ray_ctor(p, wi);
        float t = 0, tb = 0, tg = 0;
        int tri = -1;
        if (bvh_closest(s, &r2, tmin, 1e6f, &t, &tb, &tg, &tri, NULL)) {
            int sid = s->tri_surface[tri];
            if (s->surf_light[sid] == k) {
                v3 qh = add(p, scl(wi, t));
                v3 n = s->n[tri];
                v3 qhn = unit(add(add(n, scl(sub(n, n), tb)), scl(sub(n, n), tg)));
                if (s->surf_flags[sid] & ORC_SURF_FLIPPED) qhn = neg(qhn);
                v3 Le = light_query(L, qhn, neg(wi));
                float pa = 0.f;
                for (int qq = 0; qq < L->count; ++qq)
                    if (s->ltris[L->first + qq].tri == tri) { pa = s->ltris[L->first + qq].ratio * (1.0f / s->ltris[L->first + qq].area); break; }
                v3 pq = sub(qh, p);
                float lpdf = pa * dot(pq, pq) / fabsf(dot(unit(sub(p, qh)), qhn));
                float w2 = balance(dpdf, lpdf);
                v3 brdf = eval3(mat, bu, bv, bw, wo, wi);
                float cosv = fabsf(dot(wi, pn));
                Lo = add(Lo, mulv(scl(Le, (w2 * cosv) / dpdf), brdf));
            }
        }
    }
    return mk(Lo.x / L->sel_pdf, Lo.y / L->sel_pdf, Lo.z / L->sel_pdf);
}

typedef struct {
    const orc_scene *s; const float *cam; int spp, max_bounce, direct, guide_mode; uint32_t seed; float tmin;
    int tile, res, tiles_x, tiles_y, W, x0, y0, w, h;
    const float *maps, *row_cdf, *col_cdf, *totals; float *out;
} pt_args;

static float map_pdf(const pt_args *a, int slice, v3 local) {      /* lightslice.cc:2352-2377 */
    if (local.z <= 0.0f) return 0.0f;
    float sq[2], l3[3] = {local.x, local.y, local.z};
    orc_direction_to_square(l3, sq);
    int res = a->res;
    int col = (int)(sq[0] * res), row = (int)(sq[1] * res);
    if (col >= res) col = res - 1;
    if (row >= res) row = res - 1;
    float pdf_k = a->maps[((size_t)slice * res + row) * res + col] / a->totals[slice];
    float pus = (float)(res * res) * pdf_k;
    float pud = pus * INV_PI;
    return pud * 0.5f;
}

static float guide_pdf(const pt_args *a, int slice, v3 u, v3 v, v3 w, v3 sn, v3 wi) {
    if (dot(sn, wi) <= 0.0f) return 0.0f;
    if (slice < 0) return INV_2PI;
    return map_pdf(a, slice, w2l(u, v, w, wi));
}

static v3 guide_sample(const pt_args *a, int slice, const float *mat, v3 u, v3 v, v3 w, v3 sn, v3 wo,
                       const uint32_t *r, v3 *wi, float *pdf) {
    if (slice < 0) {
        float phi = TWO_PI * u24(r[0]);
        float theta = xacos(u24(r[1]));
        *wi = unit(sph2w(u, v, w, theta, phi));
    } else {
        int res = a->res;
        float ur = u24(r[0]), uc = u24(r[1]);
        const float *rcdf = a->row_cdf + (size_t)slice * res;
        int row = -1, col = -1;
        for (int i = 0; i < res; ++i) if (ur < rcdf[i]) { row = i; break; }
        if (row >= 0) {
            const float *ccdf = a->col_cdf + ((size_t)slice * res + row) * res;
            for (int j = 0; j < res; ++j) if (uc < ccdf[j]) { col = j; break; }
        }
        if (row < 0 || col < 0) { *wi = mk(0.f, 0.f, 0.f); *pdf = 0.f; return mk(0.f, 0.f, 0.f); }
        float sq[2] = {((float)col + u24(r[2])) / (float)res, ((float)row + u24(r[3])) / (float)res};
        float l3[3];
        orc_square_to_direction(sq, l3);
        *wi = l2w(u, v, w, ld(l3));
    }
    *pdf = guide_pdf(a, slice, u, v, w, sn, *wi);
    return eval3(mat, u, v, w, wo, *wi);
}

/* Camera::shoot_ray(const SensorSample&), camera.cc:67-80 */
static ray_t camera_sample_ray(const float *c, float x, float y) {
    v3 u = ld(c + 3), v = ld(c + 6), w = ld(c + 9);
    float height = c[18];
    float fx = (x - c[12]) / c[14];
    float fy = ((height - y) - c[13]) / c[15];
    v3 d = unit(l2w(u, v, w, mk(fx, fy, -c[16])));
    return ray_ctor(ld(c), d);
}

/* the slice of pixel (x, y) when guiding (its tile, if the tile's map is not empty), else -1 */
static int pt_pixel_slice(const pt_args *a, int x, int y) {
    if (a->guide_mode != 1) return -1;
    int sx = x / a->tile, sy = y / a->tile;
    if (sx > a->tiles_x - 1) sx = a->tiles_x - 1;
    if (sy > a->tiles_y - 1) sy = a->tiles_y - 1;
    int sl = sy * a->tiles_x + sx;
    return a->totals[sl] > 0.0f ? sl : -1;
}

/* One path sample from primary ray `ray`: VirtualPointLight::trace_wellington
 * (vpl.cc:934-1041) unrolled into a throughput product; the numbers of bounce b
 * come from Philox(seed, pixel, smp, b) */
static v3 pt_trace_from(const pt_args *a, ray_t ray, uint32_t pixel, int slice, int smp) {
    const orc_scene *s = a->s;
    v3 L = mk(0.f, 0.f, 0.f), T = mk(1.f, 1.f, 1.f);
    float th = 0, tb = 0, tg = 0;
    int tri = -1;
    bvh_closest(s, &ray, a->tmin, 1e6f, &th, &tb, &tg, &tri, NULL);
    for (int bounce = 1; tri >= 0 && bounce <= a->max_bounce; ++bounce) {
        float rc[ORC_RECV_STRIDE];
        int m, lidx;
        fill_receiver(s, &ray, th, tb, tg, tri, rc, &m, &lidx);
        if (lidx >= 0) {
            if (bounce == 1 && a->direct) L = add(L, mulv(T, light_query(&s->lights[lidx], ld(rc + 6), ld(rc + 9))));
            break;
        }
        if (m < 0) break;
        uint32_t u[12];
        path_rng(a->seed, pixel, (uint32_t)smp, (uint32_t)bounce, u);
        int brdf_sampling = u24(u[0]) < 0.5f;
        if (bounce > 1 || a->direct) L = add(L, mulv(T, next_event(s, rc, m, u, a->tmin)));
        if (bounce >= a->max_bounce) break;
        const float *mat = s->mat + ORC_MAT_STRIDE * m;
        v3 bu = ld(rc + 12), bv = ld(rc + 15), bw = ld(rc + 18);
        v3 sn = ld(rc + 6), wo = ld(rc + 9), p = ld(rc);
        int gslice = bounce == 1 ? slice : -1;
        v3 wi, brdf;
        float pdf;
        if (brdf_sampling) brdf = bsdf_sample(mat, bu, bv, bw, wo, u + 8, &wi, &pdf);
        else brdf = guide_sample(a, gslice, mat, bu, bv, bw, sn, wo, u + 8, &wi, &pdf);
        if (!(pdf > 0.0f)) break;
        ray = ray_ctor(p, wi);
        tri = -1;
        bvh_closest(s, &ray, a->tmin, 1e6f, &th, &tb, &tg, &tri, NULL);
        if (tri < 0) break;
        if (s->surf_light[s->tri_surface[tri]] >= 0) break;
        float po = brdf_sampling ? guide_pdf(a, gslice, bu, bv, bw, sn, wi) : bsdf_pdf(mat, bu, bv, bw, wo, wi);
        float weight = pdf / (pdf + po);
        float cosv = fabsf(dot(wi, sn));
        float sc = ((weight * cosv) / pdf) * 2.0f;
        T = scl(mulv(T, brdf), sc);
    }
    return L;
}

static void pt_range(void *a_, int lo, int hi) {
    pt_args *a = a_;
    for (int t = lo; t < hi; ++t) {
        int x = a->x0 + t % a->w, y = a->y0 + t / a->w;
        uint32_t pixel = (uint32_t)(y * a->W + x);
        int slice = pt_pixel_slice(a, x, y);
        int strata = 1;
        while ((strata + 1) * (strata + 1) <= a->spp) ++strata;
        int stratified = strata * strata == a->spp;
        v3 acc = mk(0.f, 0.f, 0.f);
        for (int smp = 0; smp < a->spp; ++smp) {
            uint32_t r0[12];
            path_rng(a->seed, pixel, (uint32_t)smp, 0u, r0);
            float jx = u24(r0[0]), jy = u24(r0[1]);
            if (stratified) {
                jx = ((float)(smp % strata) + jx) / (float)strata;
                jy = ((float)(smp / strata) + jy) / (float)strata;
            }
            ray_t ray = camera_sample_ray(a->cam, (float)x + jx, (float)y + jy);
            acc = add(acc, pt_trace_from(a, ray, pixel, slice, smp));
        }
        size_t o = (size_t)y * a->W + x;
        a->out[3 * o] = acc.x / (float)a->spp;
        a->out[3 * o + 1] = acc.y / (float)a->spp;
        a->out[3 * o + 2] = acc.z / (float)a->spp;
    }
}

void orc_guided_pt(const orc_scene *s, const float cam[ORC_CAM_STRIDE], int spp, int max_bounce, int direct,
                   int guide_mode, uint32_t seed, float tmin, int tile, int res, const float *maps,
                   const float *row_cdf, const float *col_cdf, const float *totals, int x0, int y0, int w,
                   int h, float *out, int nthreads) {
    int W = (int)cam[17], H = (int)cam[18];
    pt_args a = {s, cam, spp, max_bounce, direct, guide_mode, seed, tmin, tile, res,
                 (W + tile - 1) / tile, (H + tile - 1) / tile, W, x0, y0, w, h, maps, row_cdf, col_cdf, totals, out};
    parallel_for(w * h, nthreads, pt_range, &a);
}

/* the primary ray record of sample smp[i] of pixel (px[i], py[i]) as orc_guided_pt
 * draws it for spp samples per pixel (jitter from Philox(seed, pixel, sample, 0),
 * 2-D strata when spp is a square; Camera::shoot_ray(SensorSample), camera.cc:67-80) */
void orc_camera_sample_rays(const float cam[ORC_CAM_STRIDE], int spp, uint32_t seed, int n, const int *px,
                            const int *py, const int *smp, float *out) {
    int W = (int)cam[17];
    int strata = 1;
    while ((strata + 1) * (strata + 1) <= spp) ++strata;
    int stratified = strata * strata == spp;
    for (int i = 0; i < n; ++i) {
        uint32_t r0[12];
        path_rng(seed, (uint32_t)(py[i] * W + px[i]), (uint32_t)smp[i], 0u, r0);
        float jx = u24(r0[0]), jy = u24(r0[1]);
        if (stratified) {
            jx = ((float)(smp[i] % strata) + jx) / (float)strata;
            jy = ((float)(smp[i] / strata) + jy) / (float)strata;
        }
        ray_t r = camera_sample_ray(cam, (float)px[i] + jx, (float)py[i] + jy);
        st(out + ORC_RAY_STRIDE * i, r.o);
        st(out + ORC_RAY_STRIDE * i + 3, r.d);
    }
}

typedef struct { const pt_args *a; const float *rays; const int *pix; const int *smp; float *out; } ptr_args;

static void ptr_range(void *a_, int lo, int hi) {
    ptr_args *q = a_;
    for (int i = lo; i < hi; ++i) {
        int x = q->pix[2 * i], y = q->pix[2 * i + 1];
        ray_t ray = ray_from_record(q->rays + ORC_RAY_STRIDE * i);
        v3 L = pt_trace_from(q->a, ray, (uint32_t)(y * q->a->W + x), pt_pixel_slice(q->a, x, y), q->smp[i]);
        q->out[3 * i] = L.x; q->out[3 * i + 1] = L.y; q->out[3 * i + 2] = L.z;
    }
}

void orc_guided_pt_rays(const orc_scene *s, int width, int height, int max_bounce, int direct, int guide_mode,
                        uint32_t seed, float tmin, int tile, int res, const float *maps, const float *row_cdf,
                        const float *col_cdf, const float *totals, int n, const float *rays, const int *pix,
                        const int *smp, float *out, int nthreads) {
    pt_args a = {s, NULL, 1, max_bounce, direct, guide_mode, seed, tmin, tile, res,
                 (width + tile - 1) / tile, (height + tile - 1) / tile, width, 0, 0, 0, 0, maps, row_cdf, col_cdf,
                 totals, NULL};
    ptr_args q = {&a, rays, pix, smp, out};
    parallel_for(n, nthreads, ptr_range, &q);
}

/* ======================================================================= */
/* Bounce-1 VPL generation                                                 */
/* ======================================================================= */
/* cos and sin of 2 pi u, u in [0, 1): octant reduction and Taylor polynomials
 * of degree 9 / 10 on [0, pi/4] (the engine's own routine, restated; a few ulp
 * from libm's cos/sin of 2 pi u) */
void orc_sincos_turn(float u, float *s, float *c) {
    const float x = u * 8.0f;
    int o = (int)x;
    if (o > 7) o = 7;
    if (o < 0) o = 0;
    const float r = x - (float)o;
    const float f = (o & 1) ? 1.0f - r : r;
    const float a = f * 0.785398163397448f;
    const float a2 = a * a;
    const float sa =
        a * (1.0f + a2 * (-0.166666667f + a2 * (0.00833333333f + a2 * (-0.000198412698f + a2 * 2.75573192e-6f))));
    const float ca = 1.0f + a2 * (-0.5f + a2 * (0.0416666667f + a2 * (-0.00138888889f +
                                                                     a2 * (2.48015873e-5f + a2 * -2.75573192e-7f))));
    switch (o) {
        case 0: *s = sa; *c = ca; break;
        case 1: *s = ca; *c = sa; break;
        case 2: *s = ca; *c = -sa; break;
        case 3: *s = sa; *c = -ca; break;
        case 4: *s = -sa; *c = -ca; break;
        case 5: *s = -ca; *c = -sa; break;
        case 6: *s = -ca; *c = sa; break;
        default: *s = -sa; *c = ca; break;
    }
}

/* the 8 numbers of light sample `index` (Philox4x32-10 keyed by the seed) */
static void vpl_rng(uint32_t seed, uint64_t index, uint32_t u[8]) {
    for (uint32_t b = 0; b < 2; ++b) {
        uint32_t c[4] = {(uint32_t)index, (uint32_t)(index >> 32), b, 0x5650u};
        orc_philox(c, seed ^ 0x3c6ef372u, 0x4c475650u);
        u[4 * b] = c[0]; u[4 * b + 1] = c[1]; u[4 * b + 2] = c[2]; u[4 * b + 3] = c[3];
    }
}

/* VplGenerator::spawn with max_bounce = 2 over area lights (vpl_generator.cc:127-181):
 * per light sample, Scene::sample_light (strict <), SubpathGenerator::sample_direct_vertex
 * (path.cc:171-220, 305-334) with AreaLight::sample (area_light.cc:26-44: triangle by
 * cumulative area ratio t <= ratio (quad.cc:75-87), uniform point (triangle.cc:298-308),
 * cosine-weighted direction in init_from_nu(shape normal, shape tangent)), the closest hit;
 * a hit on a surface with a material that is not a light stores a bounce-1 VPL with
 * power Le |n_q . wo| / (pdf_light (pdf_q pdf_wo)) and wi = -ray direction.  Samples are
 * taken in order until num_vpls are stored.  Returns 0, or -1 if max_samples ran out. */
int orc_generate_vpls(const orc_scene *s, long long num_vpls, uint32_t seed, long long max_samples, float tmin,
                      float *out_vpl, int *out_mat, long long *out_samples) {
    long long have = 0, i = 0;
    for (; i < max_samples && have < num_vpls; ++i) {
        uint32_t u[8];
        vpl_rng(seed, (uint64_t)i, u);
        const float ul = u24(u[0]);
        int k = -1;
        for (int q = 0; q < s->nlights; ++q)
            if (ul < s->lights[q].sel_cdf) { k = q; break; }
        if (k < 0 || s->lights[k].count <= 0) continue;
        const light_t *L = &s->lights[k];
        const float ut = u24(u[2]);
        int ti = L->first + L->count - 1;
        for (int q = 0; q < L->count; ++q)
            if (ut <= s->ltris[L->first + q].cum) { ti = L->first + q; break; }
        const ltri_t *T = &s->ltris[ti];
        const float tmp = sqrtf(1.0f - u24(u[3]));
        const float bb = 1.0f - tmp;
        const float cc = u24(u[4]) * tmp;
        const float aa = 1.0f - bb - cc;
        const v3 q = add(add(scl(T->p0, aa), scl(T->p1, bb)), scl(T->p2, cc));
        const float pdf_q = T->ratio * (1.0f / T->area);
        float sp, cp;
        orc_sincos_turn(u24(u[5]), &sp, &cp);
        const float sin_t = sqrtf(u24(u[6]));
        const float x = sin_t * cp, y = sin_t * sp;
        const float z = sqrtf(1.f - sin_t * sin_t);
        v3 bu, bv, bw;
        onb_from_nu(s->n[T->tri], s->tan[T->tri], &bu, &bv, &bw);
        const v3 wo = unit(l2w(bu, bv, bw, mk(x, y, z)));
        const float pdf_wo = z * INV_PI;
        if (!(pdf_q > 0.0f && pdf_wo > 0.0f)) continue;
        const float acc = L->sel_pdf * (pdf_q * pdf_wo);
        const float f = fabsf(dot(T->n, wo));
        const v3 power = mk((L->le.x * f) / acc, (L->le.y * f) / acc, (L->le.z * f) / acc);
        const ray_t r = ray_ctor(q, wo);
        float t = 0.f, b = 0.f, g = 0.f;
        int tri = -1;
        bvh_closest(s, &r, tmin, 1e6f, &t, &b, &g, &tri, NULL);
        if (tri < 0) continue;
        float rc[ORC_RECV_STRIDE];
        int m = -1, light = -1;
        fill_receiver(s, &r, t, b, g, tri, rc, &m, &light);
        if (light >= 0 || m < 0) continue;
        float *o = out_vpl + 21 * have;
        for (int e = 0; e < 3; ++e) { o[e] = rc[e]; o[3 + e] = rc[6 + e]; }
        for (int e = 0; e < 9; ++e) o[6 + e] = rc[12 + e];
        st(o + 15, neg(r.d));
        st(o + 18, power);
        out_mat[have] = m;
        ++have;
    }
    *out_samples = i;
    return have == num_vpls ? 0 : -1;
}
