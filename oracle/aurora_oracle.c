/*
 * aurora_oracle.c -- CPU restatement of Aurora's path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): called by tests/, by the smoke check
 * and by the CPU-baseline leg of bench.py.  The product never links it.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).  Contraction is
 * off so that every float expression rounds the way the reference's C++ does
 * on x86-64/SSE, and so the Moller-Trumbore test is bit-identical to the GPU's.
 *
 * Citations are relative to the reference repository root; `src/` is
 * `Aurora/src/`.
 */
#include "aurora_oracle.h"

#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* constants.h:13-23 */
#define CAUSTIC_BLUR 0.2
#define FIREFLY 200
/* at most this many 0.2 steps of the clamp: every finite float is below 200
 * after 53, an infinite channel ends the loop (the reference would spin) */
#define FIREFLY_MAX_STEPS 64
#define ISBLACK 0.0001
#define RAY_BIAS 0.0005

/* ------------------------------------------------------------------------ */
/* vector helpers: src/core/vector.h, math_utils.h:31-129                    */
/* ------------------------------------------------------------------------ */
typedef struct { float x, y, z; } V3;

static inline V3 v3(float x, float y, float z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
static inline V3 v3p(const float *p) { return v3(p[0], p[1], p[2]); }
static inline V3 vadd(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 vsub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 vmul(V3 a, float f) { return v3(a.x * f, a.y * f, a.z * f); }
static inline V3 vneg(V3 a) { return v3(-a.x, -a.y, -a.z); }
/* math_utils.h:31-33 */
static inline float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
/* math_utils.h:49-53 */
static inline V3 cross3(V3 a, V3 b) {
    return v3((a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x));
}
/* vector.h:69-79: zero length reports EPSILON_LENGTH = 0.0001 */
static inline float vlen(V3 a) {
    float l = sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
    if (l == 0.f) return (float)0.0001;
    return l;
}
/* math_utils.h:39-41 with vector.h:116-120 (division multiplies by 1/f) */
static inline V3 vnorm(V3 a) { float inv = 1.f / vlen(a); return vmul(a, inv); }

/* math_utils.h:114-129 */
static inline void coordinate_system(V3 n, V3 *t, V3 *b) {
    if (!n.x && n.y && !n.z) *t = cross3(n, v3(1.f, 0.f, 0.f));
    else *t = cross3(n, v3(0.f, 1.f, 0.f));
    *b = cross3(n, *t);
}
/* math_utils.h:81-90 */
static inline V3 tangent_to_world(V3 dir, V3 n) {
    V3 t, b;
    coordinate_system(n, &t, &b);
    V3 r = vmul(vnorm(t), dir.x);
    r = vadd(r, vmul(vnorm(b), dir.y));
    r = vadd(r, vmul(n, dir.z));
    return r;
}
/* math_utils.h:68-76 */
static inline V3 sample_hemisphere_uniform(float r1, float r2) {
    float z = r1;
    float r = sqrtf(fmax(0.f, 1.f - z * z));
    float phi = 2 * M_PI * r2;
    return v3(r * cosf(phi), r * sinf(phi), z);
}
/* math_utils.h:222-224 */
static inline float schlick(float vdoth, float refl) {
    return refl + powf(1.f - vdoth, 5.f) * (1.f - refl);
}
/* math_utils.h:216-220 */
static inline float fclampf(float x, float lo, float hi) {
    float r = x < lo ? lo : x;
    return r > hi ? hi : r;
}
/* math_utils.h:236-238 */
static inline int ref_floor(float x) { return (int)x - (x < 0 && (int)x != x); }

typedef struct { float r, g, b; } C3;
static inline C3 c3(float r, float g, float b) { C3 c; c.r = r; c.g = g; c.b = b; return c; }
static inline C3 c3p(const float *p) { return c3(p[0], p[1], p[2]); }
static inline C3 cadd(C3 a, C3 b) { return c3(a.r + b.r, a.g + b.g, a.b + b.b); }
static inline C3 cmulc(C3 a, C3 b) { return c3(a.r * b.r, a.g * b.g, a.b * b.b); }
static inline C3 cmulf(C3 a, float f) { return c3(a.r * f, a.g * f, a.b * f); }
/* color.h operator*(double) */
static inline C3 cmuld(C3 a, double d) { return c3((float)(a.r * d), (float)(a.g * d), (float)(a.b * d)); }
/* color.h operator/(float): multiplies by 1/f */
static inline C3 cdivf(C3 a, float f) { float d = 1 / f; return c3(a.r * d, a.g * d, a.b * d); }
/* color.h lum(): 0.2126r + 0.7152g + 0.0722b in double */
static inline float clum(C3 a) { return (float)(0.2126 * a.r + 0.7152 * a.g + 0.0722 * a.b); }
/* color.h isBlack() */
static inline int cblack(C3 a) { return a.r < ISBLACK && a.g < ISBLACK && a.b < ISBLACK; }

/* ------------------------------------------------------------------------ */
/* random numbers                                                            */
/* ------------------------------------------------------------------------ */
static inline uint32_t pcg_hash(uint32_t v) {
    uint32_t s = v * 747796405u + 2891336453u;
    uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}
uint32_t orc_path_key(uint32_t seed, uint32_t pixel, uint32_t sample) {
    return pcg_hash(pcg_hash(pcg_hash(seed) ^ pixel) ^ (sample * 0x9E3779B9u));
}
float orc_slot_uniform(uint32_t key, uint32_t slot) {
    return (float)(pcg_hash(key + slot * 0x85EBCA6Bu) >> 8) * (1.0f / 16777216.0f);
}

/* A path's random numbers: injected slots, the Halton source (scene sampler
 * 1: the sample-th Halton point of the pixel, slots 0 and 1 the camera offsets
 * inside the pixel, slot k < 256 dimension k; oracle_halton.c) or the hash. */
typedef struct {
    const float *slots;
    uint32_t key, key_cam;           /* vertex slots / camera slots 0, 1 */
    int qmc;
    unsigned w, h, x, y, sample_cam, sample_v;
} Rng;
static inline float rng_u(const Rng *r, int slot) {
    if (r->slots) return r->slots[slot];
    if (r->qmc && slot < 256)
        return orc_halton_slot(r->w, r->h, slot < 2 ? r->sample_cam : r->sample_v, r->x, r->y, (unsigned)slot);
    return orc_slot_uniform(slot < 2 ? r->key_cam : r->key, (uint32_t)slot);
}
/* Path `sub` of the camera sample (x, y, sample): camera slots from the
 * sample's stream, vertex slots from the path's (sub 0: the same stream; the
 * Halton source takes its vertex point at sample * ls + sub). */
static inline Rng make_rng_sub(const orc_scene *s, const float *slots, uint32_t seed, int x, int y, uint32_t sample,
                               int sub, int ls) {
    Rng r;
    r.slots = slots;
    r.key_cam = orc_path_key(seed, (uint32_t)(y * s->width + x), sample);
    r.key = sub == 0 ? r.key_cam : pcg_hash(r.key_cam ^ ((uint32_t)sub * 0x68E31DA4u));
    r.qmc = s->sampler == 1;
    r.w = (unsigned)s->width; r.h = (unsigned)s->height;
    r.x = (unsigned)x; r.y = (unsigned)y;
    r.sample_cam = sample;
    r.sample_v = ls > 1 ? sample * (unsigned)ls + (unsigned)sub : sample;
    return r;
}
static inline Rng make_rng(const orc_scene *s, const float *slots, uint32_t seed, int x, int y, uint32_t sample) {
    return make_rng_sub(s, slots, seed, x, y, sample, 0, 1);
}

/* ------------------------------------------------------------------------ */
/* triangle test: src/geometry/renderableTriangle.cpp:33-66 (Moller-Trumbore) */
/* ------------------------------------------------------------------------ */
int orc_intersect_tri(const float *p1, const float *e1p, const float *e2p, const orc_ray *r,
                      float *t_out, float *b1_out, float *b2_out) {
    V3 d = v3p(r->d), o = v3p(r->o), e1 = v3p(e1p), e2 = v3p(e2p), p = v3p(p1);
    V3 s1 = cross3(d, e2);
    float divisor = dot3(s1, e1);
    if (divisor == 0.f) return 0;
    float inv = 1.f / divisor;
    V3 dv = vsub(o, p);
    float b1 = dot3(dv, s1) * inv;
    if (b1 < 0.f || b1 > 1.f) return 0;
    V3 s2 = cross3(dv, e1);
    float b2 = dot3(d, s2) * inv;
    if (b2 < 0.f || (b1 + b2) > 1.f) return 0;
    float t = dot3(e2, s2) * inv;
    if (t < r->tmin || t > r->tmax) return 0;
    *t_out = t; *b1_out = b1; *b2_out = b2;
    return 1;
}

static inline void tri_pe(const orc_scene *s, int prim, float *p, float *e1, float *e2) {
    const int *ix = s->tris + 3 * prim;
    const float *a = s->verts + 3 * ix[0], *b = s->verts + 3 * ix[1], *c = s->verts + 3 * ix[2];
    for (int k = 0; k < 3; ++k) { p[k] = a[k]; e1[k] = b[k] - a[k]; e2[k] = c[k] - a[k]; }
}

/* Closest hit with the lowest primitive id winning exact-t ties; the interval
 * is [tmin, tmax] inclusive as in renderableTriangle.cpp:64. */
static void trace_one_brute(const orc_scene *s, const orc_ray *ray, int any_hit, orc_hit *h) {
    orc_ray r = *ray;
    h->prim = -1; h->t = r.tmax; h->u = h->v = 0.f;
    for (int i = 0; i < s->n_tris; ++i) {
        float p[3], e1[3], e2[3], t, b1, b2;
        tri_pe(s, i, p, e1, e2);
        if (orc_intersect_tri(p, e1, e2, &r, &t, &b1, &b2)) {
            if (h->prim < 0 || t < h->t) {
                h->prim = i; h->t = t; h->u = b1; h->v = b2;
                r.tmax = t;
                if (any_hit) return;
            }
        }
    }
}

void orc_trace_brute(const orc_scene *s, const orc_ray *rays, int n, int any_hit, orc_hit *out) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) trace_one_brute(s, rays + i, any_hit, out + i);
}

/* ------------------------------------------------------------------------ */
/* binary SAH BVH (stand-in for Embree's bvh4.spatialsplit in the CPU path)  */
/* ------------------------------------------------------------------------ */
typedef struct { float lo[3], hi[3]; int left, count; } BNode; /* count>0: leaf, left=first */
struct orc_bvh {
    BNode *nodes; int n_nodes;
    int *prims;
    float *tri; /* per sorted slot: p, e1, e2 (9 floats) */
};

typedef struct { float lo[3], hi[3], c[3]; } PrimInfo;

static void box_empty(float *lo, float *hi) { for (int k = 0; k < 3; ++k) { lo[k] = FLT_MAX; hi[k] = -FLT_MAX; } }
static void box_grow(float *lo, float *hi, const float *plo, const float *phi) {
    for (int k = 0; k < 3; ++k) { if (plo[k] < lo[k]) lo[k] = plo[k]; if (phi[k] > hi[k]) hi[k] = phi[k]; }
}
static float box_area(const float *lo, const float *hi) {
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    if (dx < 0 || dy < 0 || dz < 0) return 0.f;
    return 2.f * (dx * dy + dy * dz + dz * dx);
}

typedef struct { BNode *nodes; int n, cap; PrimInfo *pi; int *prims; int leaf; } Builder;

static int bnode_alloc(Builder *b) {
    if (b->n == b->cap) { b->cap = b->cap ? 2 * b->cap : 1024; b->nodes = (BNode *)realloc(b->nodes, sizeof(BNode) * b->cap); }
    return b->n++;
}

#define NBINS 16
static void build_rec(Builder *b, int ni, int first, int count) {
    float lo[3], hi[3], clo[3], chi[3];
    box_empty(lo, hi); box_empty(clo, chi);
    for (int i = first; i < first + count; ++i) {
        PrimInfo *p = &b->pi[b->prims[i]];
        box_grow(lo, hi, p->lo, p->hi);
        box_grow(clo, chi, p->c, p->c);
    }
    BNode *nd = &b->nodes[ni];
    memcpy(nd->lo, lo, sizeof lo); memcpy(nd->hi, hi, sizeof hi);
    if (count <= b->leaf) { nd->left = first; nd->count = count; return; }
    int best_axis = -1, best_bin = -1; float best_cost = FLT_MAX;
    for (int ax = 0; ax < 3; ++ax) {
        float ext = chi[ax] - clo[ax];
        if (ext <= 0.f) continue;
        int cnt[NBINS] = {0}; float blo[NBINS][3], bhi[NBINS][3];
        for (int k = 0; k < NBINS; ++k) box_empty(blo[k], bhi[k]);
        for (int i = first; i < first + count; ++i) {
            PrimInfo *p = &b->pi[b->prims[i]];
            int k = (int)(NBINS * ((p->c[ax] - clo[ax]) / ext)); if (k >= NBINS) k = NBINS - 1; if (k < 0) k = 0;
            cnt[k]++; box_grow(blo[k], bhi[k], p->lo, p->hi);
        }
        float rlo[NBINS][3], rhi[NBINS][3]; int rc[NBINS];
        float alo[3], ahi[3]; box_empty(alo, ahi); int ac = 0;
        for (int k = NBINS - 1; k > 0; --k) {
            box_grow(alo, ahi, blo[k], bhi[k]); ac += cnt[k];
            memcpy(rlo[k], alo, sizeof alo); memcpy(rhi[k], ahi, sizeof ahi); rc[k] = ac;
        }
        box_empty(alo, ahi); ac = 0;
        for (int k = 0; k < NBINS - 1; ++k) {
            box_grow(alo, ahi, blo[k], bhi[k]); ac += cnt[k];
            if (ac == 0 || rc[k + 1] == 0) continue;
            float cost = ac * box_area(alo, ahi) + rc[k + 1] * box_area(rlo[k + 1], rhi[k + 1]);
            if (cost < best_cost) { best_cost = cost; best_axis = ax; best_bin = k; }
        }
    }
    int mid;
    if (best_axis < 0) {
        mid = first + count / 2; /* all centroids coincide: split by index */
    } else {
        float ext = chi[best_axis] - clo[best_axis];
        int i = first, j = first + count - 1;
        while (i <= j) {
            PrimInfo *p = &b->pi[b->prims[i]];
            int k = (int)(NBINS * ((p->c[best_axis] - clo[best_axis]) / ext)); if (k >= NBINS) k = NBINS - 1; if (k < 0) k = 0;
            if (k <= best_bin) ++i;
            else { int t = b->prims[i]; b->prims[i] = b->prims[j]; b->prims[j] = t; --j; }
        }
        mid = i;
        if (mid == first || mid == first + count) mid = first + count / 2;
    }
    int l = bnode_alloc(b), r = bnode_alloc(b);
    (void)r;
    b->nodes[ni].left = l; b->nodes[ni].count = 0;
    build_rec(b, l, first, mid - first);
    build_rec(b, l + 1, mid, first + count - mid);
}

orc_bvh *orc_bvh_build(const orc_scene *s, int leaf_size) {
    Builder b; memset(&b, 0, sizeof b);
    b.leaf = leaf_size > 0 ? leaf_size : 4;
    b.pi = (PrimInfo *)malloc(sizeof(PrimInfo) * (size_t)(s->n_tris > 0 ? s->n_tris : 1));
    b.prims = (int *)malloc(sizeof(int) * (size_t)(s->n_tris > 0 ? s->n_tris : 1));
    for (int i = 0; i < s->n_tris; ++i) {
        PrimInfo *p = &b.pi[i];
        box_empty(p->lo, p->hi);
        float mag = 0.f;
        for (int c = 0; c < 3; ++c) {
            const float *v = s->verts + 3 * s->tris[3 * i + c];
            box_grow(p->lo, p->hi, v, v);
            for (int k = 0; k < 3; ++k) if (fabsf(v[k]) > mag) mag = fabsf(v[k]);
        }
        /* conservative padding: traversal must never cull a triangle that the
         * Moller-Trumbore test accepts */
        float pad = mag * 1e-5f + 1e-30f;
        for (int k = 0; k < 3; ++k) { p->lo[k] -= pad; p->hi[k] += pad; p->c[k] = 0.5f * (p->lo[k] + p->hi[k]); }
        b.prims[i] = i;
    }
    int root = bnode_alloc(&b);
    if (s->n_tris > 0) build_rec(&b, root, 0, s->n_tris);
    else { box_empty(b.nodes[0].lo, b.nodes[0].hi); b.nodes[0].left = 0; b.nodes[0].count = 0; b.nodes[0].count = -1; }
    orc_bvh *out = (orc_bvh *)malloc(sizeof(orc_bvh));
    out->nodes = b.nodes; out->n_nodes = b.n; out->prims = b.prims;
    out->tri = (float *)malloc(sizeof(float) * 9 * (size_t)(s->n_tris > 0 ? s->n_tris : 1));
    for (int i = 0; i < s->n_tris; ++i) tri_pe(s, b.prims[i], out->tri + 9 * i, out->tri + 9 * i + 3, out->tri + 9 * i + 6);
    free(b.pi);
    return out;
}

void orc_bvh_free(orc_bvh *b) {
    if (!b) return;
    free(b->nodes); free(b->prims); free(b->tri); free(b);
}

static inline int box_hit(const BNode *n, const float *o, const float *inv, float tmin, float tmax, float *tnear) {
    float t0 = tmin, t1 = tmax * 1.0000004f;
    for (int k = 0; k < 3; ++k) {
        float a = (n->lo[k] - o[k]) * inv[k], b = (n->hi[k] - o[k]) * inv[k];
        if (a > b) { float t = a; a = b; b = t; }
        if (a != a) a = -FLT_MAX; /* 0 * inf */
        if (b != b) b = FLT_MAX;
        if (a > t0) t0 = a;
        if (b < t1) t1 = b;
    }
    *tnear = t0;
    return t0 <= t1;
}

static void trace_one_bvh(const orc_scene *s, const orc_bvh *bv, const orc_ray *ray, int any_hit, orc_hit *h) {
    orc_ray r = *ray;
    h->prim = -1; h->t = r.tmax; h->u = h->v = 0.f;
    if (s->n_tris == 0) return;
    float inv[3];
    for (int k = 0; k < 3; ++k) inv[k] = 1.f / r.d[k];
    int stack[128]; int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        int ni = stack[--sp];
        const BNode *n = &bv->nodes[ni];
        float tn;
        if (!box_hit(n, r.o, inv, r.tmin, r.tmax, &tn)) continue;
        if (n->count > 0) {
            for (int i = n->left; i < n->left + n->count; ++i) {
                const float *tr = bv->tri + 9 * i;
                float t, b1, b2;
                if (orc_intersect_tri(tr, tr + 3, tr + 6, &r, &t, &b1, &b2)) {
                    int prim = bv->prims[i];
                    if (h->prim < 0 || t < h->t || (t == h->t && prim < h->prim)) {
                        h->prim = prim; h->t = t; h->u = b1; h->v = b2; r.tmax = t;
                        if (any_hit) return;
                    }
                }
            }
        } else if (n->count == 0) {
            int l = n->left, rr = n->left + 1;
            float tl, tr;
            int hl = box_hit(&bv->nodes[l], r.o, inv, r.tmin, r.tmax, &tl);
            int hr = box_hit(&bv->nodes[rr], r.o, inv, r.tmin, r.tmax, &tr);
            if (hl && hr) {
                if (tl <= tr) { stack[sp++] = rr; stack[sp++] = l; }
                else { stack[sp++] = l; stack[sp++] = rr; }
            } else if (hl) stack[sp++] = l;
            else if (hr) stack[sp++] = rr;
        }
    }
}

void orc_trace_bvh(const orc_scene *s, const orc_bvh *b, const orc_ray *rays, int n, int any_hit, orc_hit *out) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; ++i) trace_one_bvh(s, b, rays + i, any_hit, out + i);
}

static inline void trace(const orc_scene *s, const orc_bvh *b, orc_ray *r, int any_hit, orc_hit *h) {
    if (b) trace_one_bvh(s, b, r, any_hit, h);
    else trace_one_brute(s, r, any_hit, h);
}

/* ------------------------------------------------------------------------ */
/* transforms: src/core/transform.cpp:206-214                                */
/* ------------------------------------------------------------------------ */
static V3 xform_point(const float *m, V3 p) {
    float xp = m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3];
    float yp = m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7];
    float zp = m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11];
    float wp = m[12] * p.x + m[13] * p.y + m[14] * p.z + m[15];
    if (wp == 1.f) return v3(xp, yp, zp);
    float inv = 1.f / wp;
    return v3(xp * inv, yp * inv, zp * inv);
}

/* ------------------------------------------------------------------------ */
/* lights                                                                    */
/* ------------------------------------------------------------------------ */
/* squareLight.cpp:20-52 and light.h:42-45 (setExposure) */
void orc_light_init(orc_light *l) {
    l->intensity = powf(2, l->exposure);
    if (l->type != ORC_LIGHT_AREA) return;
    V3 p1 = xform_point(l->m, v3(-l->sx, -l->sy, 0));
    V3 p2 = xform_point(l->m, v3(-l->sx, l->sy, 0));
    V3 p3 = xform_point(l->m, v3(l->sx, l->sy, 0));
    V3 p4 = xform_point(l->m, v3(l->sx, -l->sy, 0));
    V3 n = vnorm(cross3(vsub(p1, p2), vsub(p2, p3)));
    l->N[0] = n.x; l->N[1] = n.y; l->N[2] = n.z;
    V3 pc[4] = {p1, p2, p3, p4};
    for (int i = 0; i < 4; ++i) { l->pcam[i][0] = pc[i].x; l->pcam[i][1] = pc[i].y; l->pcam[i][2] = pc[i].z; }
    /* m_tri1 = (p1, p2, p4), m_tri2 = (p3, p2, p4); renderableTriangle.cpp:21-30 */
    V3 a[2] = {p1, p3};
    for (int k = 0; k < 2; ++k) {
        V3 e1 = vsub(p2, a[k]), e2 = vsub(p4, a[k]);
        float *t = l->tri[k];
        t[0] = a[k].x; t[1] = a[k].y; t[2] = a[k].z;
        t[3] = e1.x; t[4] = e1.y; t[5] = e1.z;
        t[6] = e2.x; t[7] = e2.y; t[8] = e2.z;
    }
}

typedef struct { orc_ray ray; float pdf; C3 color; } Sample3D;

/* ------------------------------------------------------------------------ */
/* textures and shaders                                                      */
/* ------------------------------------------------------------------------ */
/* FP16 round-to-nearest-even, by value: 11 significant bits in the normal
 * range, steps of 2^-24 below 2^-14, infinity from 65520 on (the rounding of
 * Imf::Rgba's half). */
float orc_half_round(float x) {
    if (isnan(x) || isinf(x)) return x;
    float a = fabsf(x), r;
    if (a < 6.103515625e-05f) {
        r = ldexpf(rintf(ldexpf(a, 24)), -24);
    } else {
        int e;
        float f = frexpf(a, &e);
        r = ldexpf(rintf(ldexpf(f, 11)), e - 11);
        if (r > 65504.f) r = INFINITY;
    }
    return copysignf(r, x);
}

static int to_int_x86(float v) {
    /* cvttss2si: out-of-range and NaN give INT_MIN */
    if (!(v >= -2147483648.f && v < 2147483648.f)) return INT_MIN;
    return (int)v;
}

/* texture2D.cpp:48-66 (filterSize 0).  An index left at the size by the
 * clamps (arguments beyond 1 + 1/size) would read past the row; no caller
 * passes one, and the last texel is used instead. */
static const float *texel(const orc_texture *t, float s, float tt) {
    int x = to_int_x86(floorf(s * t->width));
    int y = to_int_x86(floorf(tt * t->height));
    if (y == t->height) y -= 1;
    if (x == t->width) x -= 1;
    if (x > t->width || x < 0) { x = x < t->width ? x : t->width; x = x > 0 ? x : 0; }
    if (y > t->height || y < 0) { y = y < t->height ? y : t->height; y = y > 0 ? y : 0; }
    if (x >= t->width) x = t->width - 1;
    if (y >= t->height) y = t->height - 1;
    return t->rgba + 4 * ((size_t)y * t->width + x);
}

void orc_texture_read(const orc_texture *t, float s, float tt, float *rgb_out) {
    const float *p = texel(t, s, tt);
    rgb_out[0] = p[0]; rgb_out[1] = p[1]; rgb_out[2] = p[2];
}

/* texture2D.cpp:120-155: alpha <- (r+g+b)/3 stored as half, normalisation =
 * float sum of the unrounded values over the texel count */
void orc_texture_init_pdf(orc_texture *t) {
    float lum = 0;
    for (int i = 0; i < t->height; i++)
        for (int j = 0; j < t->width; j++) {
            float *p = t->rgba + 4 * ((size_t)i * t->width + j);
            float pdf = (p[0] + p[1] + p[2]) / 3.f;
            p[3] = orc_half_round(pdf);
            lum += pdf;
        }
    t->pdf_norm = lum / (t->width * t->height);
}

/* texture2D.cpp:80-103 */
static float texture_pdf(const orc_texture *t, float s, float tt) { return texel(t, s, tt)[3] / t->pdf_norm; }

static V3 xform_vec(const float *m, V3 v) {
    return v3(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
              m[8] * v.x + m[9] * v.y + m[10] * v.z);
}
static float v3_at(V3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

/* Shader<T>::evaluate: constantShader.h, texture2DShader.h:28-31 and 47-50,
 * triplanarTextureShader.h:28-53 (color) and 76-99 (float: red channel) */
static C3 shader_eval(const orc_scene *sc, int idx, V3 P, V3 Ng, float su, float sv, int attr) {
    const orc_shader *sh = &sc->shaders[idx];
    if (sh->type == ORC_SHADER_CONSTANT) return c3(sh->value[0], sh->value[1], sh->value[2]);
    const orc_texture *tex = &sc->textures[sh->texture];
    if (sh->type == ORC_SHADER_TEXTURE2D) {
        const float *p = texel(tex, su - floorf(su), 1.f - (sv - floorf(sv)));
        return c3(p[0], p[1], p[2]);
    }
    const float *c2o = sc->mats[attr].c2o;
    float w[3], wsum = 0.f;
    if (sh->kind == ORC_SHADER_COLOR) {
        V3 N = vnorm(xform_vec(c2o, Ng));
        for (int i = 0; i < 3; ++i) {
            w[i] = fabsf(v3_at(N, i));
            w[i] = (float)(w[i] - 0.25);
            w[i] = (w[i] < 0.f) ? 0.f : w[i];
            wsum += w[i];
        }
    } else {
        for (int i = 0; i < 3; ++i) {
            w[i] = fabsf(v3_at(Ng, i));
            w[i] = (float)(w[i] - 0.2);
            w[i] = (w[i] < 0.f) ? 0.f : w[i];
            wsum += w[i];
        }
    }
    for (int i = 0; i < 3; ++i) w[i] /= wsum;
    V3 Po = xform_point(c2o, P);
    float inv = 1.f / sh->scale;
    Po = v3(Po.x * inv, Po.y * inv, Po.z * inv);
    C3 res = c3(0.f, 0.f, 0.f);
    for (int i = 0; i < 3; ++i) {
        if (w[i] > 0) {
            float s = fabsf(fmodf(v3_at(Po, (i + 1) % 3), 1.f));
            float t = fabsf(fmodf(v3_at(Po, (i + 2) % 3), 1.f));
            const float *p = texel(tex, s, t);
            res = cadd(res, cmulf(c3(p[0], p[1], p[2]), w[i]));
        }
    }
    return res;
}

void orc_shader_eval(const orc_scene *sc, int shader, const float *P, const float *Ng, float s, float t, int attr,
                     float *out) {
    C3 c = shader_eval(sc, shader, v3p(P), v3p(Ng), s, t, attr);
    out[0] = c.r; out[1] = c.g; out[2] = c.b;
}

/* infiniteAreaLight.cpp:61-78 with the map: u from atan2 of the world
 * direction (the division in double), v = (1 - y) / 2 */
static void env_uv(const orc_light *l, V3 d, float *u, float *v) {
    V3 dir = xform_vec(l->c2w, d);
    *v = (float)(0.5 * (1.f - dir.y));
    *u = (float)(atan2f(dir.z, dir.x) / (M_PI * 2.f));
    if (*u < 0.f) *u += 1.f;
}

static C3 light_eval(const orc_scene *sc, const orc_light *l, V3 d) {
    if (l->type == ORC_LIGHT_AREA) return cmulf(c3p(l->color), l->intensity); /* squareLight.h:84-86 */
    if (l->envmap < 0) return cmuld(c3p(l->color), pow(2, l->exposure));      /* infiniteAreaLight.cpp:74-76 */
    float u, v;
    env_uv(l, d, &u, &v);
    const float *p = texel(&sc->textures[l->envmap], u, v);
    return cmuld(c3(p[0], p[1], p[2]), pow(2, l->exposure));
}

void orc_env_eval(const orc_scene *sc, const orc_light *l, const float *dir, float *rgb_out) {
    C3 c = light_eval(sc, l, v3p(dir));
    rgb_out[0] = c.r; rgb_out[1] = c.g; rgb_out[2] = c.b;
}

/* squareLight.h:57-82 / infiniteAreaLight.cpp:35-59 */
static Sample3D light_sample(const orc_scene *sc, const orc_light *l, V3 orig, V3 nn, float x, float y) {
    Sample3D s;
    if (l->type == ORC_LIGHT_AREA) {
        V3 lp = v3((float)((x - 0.5) * l->sx * 2), (float)((y - 0.5) * l->sy * 2), 0.f);
        lp = xform_point(l->m, lp);
        V3 dir = vsub(lp, orig);
        float len = vlen(dir);
        dir = vnorm(dir);
        s.ray.o[0] = orig.x; s.ray.o[1] = orig.y; s.ray.o[2] = orig.z;
        s.ray.d[0] = dir.x; s.ray.d[1] = dir.y; s.ray.d[2] = dir.z;
        s.ray.tmin = (float)RAY_BIAS;
        s.ray.tmax = (float)(len - (RAY_BIAS * 10));
        float d = dot3(v3p(l->N), vneg(dir));
        if (signbit(d)) d = 0.f; /* d *= 1 - signbit(d) */
        s.color = cmulf(c3p(l->color), l->intensity);
        if (d == 0.f) s.pdf = 0.f;
        else s.pdf = (float)((len * len) / (d * l->sx * l->sy * 4.));
    } else {
        V3 dw;
        float pdf;
        if (l->envmap >= 0) {
            float vy = 1.f - 2.f * y;
            float r = sqrtf(fmaxf(0.f, 1.f - vy * vy));
            float phi = (float)(2.f * M_PI * x);
            dw = xform_vec(l->w2c, v3(r * cosf(phi), vy, r * sinf(phi)));
            pdf = texture_pdf(&sc->textures[l->envmap], x, y);
        } else {
            pdf = 2.f;
            dw = tangent_to_world(sample_hemisphere_uniform(x, y), nn);
        }
        dw = vnorm(dw);
        s.ray.o[0] = orig.x; s.ray.o[1] = orig.y; s.ray.o[2] = orig.z;
        s.ray.d[0] = dw.x; s.ray.d[1] = dw.y; s.ray.d[2] = dw.z;
        s.ray.tmin = (float)RAY_BIAS; s.ray.tmax = 1000000.f;
        s.pdf = (float)(pdf / (M_PI * 4.));
        s.color = light_eval(sc, l, dw);
    }
    return s;
}

/* squareLight.h:88-97 (ray maxt shortened on hit) / infiniteAreaLight.cpp:80-106 */
static float light_pdf(const orc_scene *sc, const orc_light *l, orc_ray *ray) {
    if (l->type == ORC_LIGHT_AREA) {
        float t, b1, b2;
        int hit = orc_intersect_tri(l->tri[0], l->tri[0] + 3, l->tri[0] + 6, ray, &t, &b1, &b2);
        if (!hit) hit = orc_intersect_tri(l->tri[1], l->tri[1] + 3, l->tri[1] + 6, ray, &t, &b1, &b2);
        if (!hit) return 0.f;
        ray->tmax = t;
        float d = dot3(v3p(l->N), vnorm(v3p(ray->d)));
        d *= (float)((0.f < d) - (d < 0.f));
        float pdf = (float)((ray->tmax * ray->tmax) / (d * l->sx * l->sy * 4.));
        ray->tmax = (float)(ray->tmax - RAY_BIAS * 10);
        return pdf;
    }
    if (l->envmap < 0) return (float)(1.f / (M_PI * 2.f));
    float u, v;
    env_uv(l, v3p(ray->d), &u, &v);
    return (float)(texture_pdf(&sc->textures[l->envmap], u, v) / (M_PI * 2.f)); /* Hemisphere domain */
}

/* squareLight.h:99-110; infiniteAreaLight.h:29 */
static int light_visible(const orc_light *l, V3 orig, V3 nn) {
    if (l->type != ORC_LIGHT_AREA) return 1;
    for (int i = 0; i < 4; ++i)
        if (dot3(vsub(v3p(l->pcam[i]), orig), nn) > 0.f) return 1;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* BRDFs                                                                     */
/* ------------------------------------------------------------------------ */
typedef struct { int kind; int dist; C3 color; float rough; } Bsdf;

/* cookTorrance.cpp:40-45 (Beckmann: D*cos) and the GGX analogue */
static inline float ct_d(int dist, float rough, float cosa, float tana) {
    const float r2 = fmaxf(0.001f, rough * rough);
    if (dist == ORC_DIST_GGX) {
        float t = r2 + tana * tana;
        double div = M_PI * powf(cosa, 3.f) * t * t;
        return (float)(r2 / div);
    }
    const float e = expf(-(tana * tana) / r2);
    const double divisor = r2 * M_PI * powf(cosa, 3.f);
    return (float)(e / divisor);
}

/* cookTorrance.cpp:125-144 (no 1e-3 floor, no clamp) */
static inline float ct_d_eval(int dist, float rough, float ndoth) {
    const float r2 = rough * rough;
    float tana = tanf(acosf(ndoth));
    if (dist == ORC_DIST_GGX) {
        float t = r2 + tana * tana;
        return (float)((1.f / (M_PI * powf(ndoth, 3.f) * t * t)) * r2);
    }
    return (float)((1.f / (r2 * M_PI * powf(ndoth, 3.f))) * expf(-(tana * tana) / r2));
}

static Sample3D bsdf_sample(const Bsdf *b, V3 vn, V3 nn, float r1, float r2) {
    Sample3D s;
    memset(&s, 0, sizeof s);
    if (b->kind == ORC_BSDF_LAMBERT) {
        /* lambert.h:44-59 */
        V3 dir = sample_hemisphere_uniform(r1, r2);
        V3 w = vnorm(tangent_to_world(dir, nn));
        s.ray.d[0] = w.x; s.ray.d[1] = w.y; s.ray.d[2] = w.z;
        s.ray.tmin = (float)RAY_BIAS; s.ray.tmax = 1000000.f;
        s.pdf = (float)(1.f / M_PI * 0.5);
        s.color = cdivf(cmulf(b->color, 1.f), (float)M_PI);
        return s;
    }
    /* cookTorrance.cpp:47-95 */
    float z;
    if (b->dist == ORC_DIST_GGX) z = cosf(atanf(b->rough * sqrtf((float)(r1 / (1. - r1)))));
    else z = cosf(atanf(b->rough * sqrtf((float)(-log(1. - r1)))));
    float r = sqrtf(fmax(0.f, 1.f - z * z));
    float phi = M_PI * 2 * r2;
    V3 dir = vnorm(v3(r * cosf(phi), r * sinf(phi), z));
    V3 hn = vnorm(tangent_to_world(dir, nn));
    float vdoth = dot3(vn, hn);
    V3 ln = vnorm(vsub(vmul(vmul(hn, 2.f), vdoth), vn));
    float ndoth = dot3(nn, hn);
    C3 res = c3(0.f, 0.f, 0.f);
    float pdf = 1.f;
    if (vdoth > 0.0001) {
        const float tana = (ndoth > 0.999) ? 0.f : tanf(acosf(ndoth));
        float D = ct_d(b->dist, b->rough, ndoth, tana);
        D = fminf(D, 1000.f);
        pdf = D / (4 * vdoth);
        res = cdivf(cmulf(b->color, D), 4 * vdoth * vdoth);
        if (pdf > 100000) { pdf *= 0.01f; res = cmulf(res, 0.01f); }
    }
    s.ray.d[0] = ln.x; s.ray.d[1] = ln.y; s.ray.d[2] = ln.z;
    s.ray.tmin = (float)RAY_BIAS; s.ray.tmax = 10000000.f;
    s.pdf = pdf; s.color = res;
    return s;
}

/* lambert.h:67-80; cookTorrance.cpp:97-117 */
static float bsdf_pdf(const Bsdf *b, V3 ln, V3 vn, V3 nn) {
    if (b->kind == ORC_BSDF_LAMBERT) return (float)(1.f / M_PI * 0.5);
    V3 hn = vnorm(vadd(ln, vn));
    float vdoth = dot3(vn, hn);
    float ndoth = dot3(nn, hn);
    float D = ct_d(b->dist, b->rough, ndoth, tanf(acosf(ndoth)));
    if (isnan(D)) D = 0.f;
    return D / (4 * vdoth);
}

/* lambert.h:67-75; cookTorrance.cpp:125-144 */
static C3 bsdf_eval(const Bsdf *b, V3 ln, V3 vn, V3 nn) {
    if (b->kind == ORC_BSDF_LAMBERT) {
        float d = dot3(ln, nn);
        if (d > 0.f) return cdivf(cmulf(b->color, 1.f), (float)M_PI);
        return c3(0.f, 0.f, 0.f);
    }
    V3 hn = vnorm(vadd(ln, vn));
    const float vdoth = dot3(vn, hn);
    const float vdotn = dot3(vn, nn);
    const float ndoth = dot3(hn, nn);
    const float ldotn = dot3(ln, nn);
    const float D = ct_d_eval(b->dist, b->rough, ndoth);
    if (vdoth <= 0.001 || ldotn <= 0.001 || vdotn <= 0.001 || isnan(D)) return c3(0.f, 0.f, 0.f);
    return cdivf(cmulf(b->color, D), (float)(4. * vdoth * vdoth));
}

static Sample3D delta_sample(const Bsdf *b, V3 vn, V3 nn);

/* kinds ORC_BSDF_MIRROR / ORC_BSDF_TRANSMIT: specMirror.cpp / specTransmit.cpp
 * (evalSampleWorld and pdf are zero; params[3] is the ior) */
void orc_bsdf_eval(int kind, const float *params, const float *V, const float *N, const float *L,
                   float *f_out, float *pdf_out) {
    Bsdf b; b.kind = kind; b.color = c3p(params); b.rough = params[3]; b.dist = (int)params[4];
    if (kind >= ORC_BSDF_MIRROR) { f_out[0] = f_out[1] = f_out[2] = 0.f; *pdf_out = 0.f; return; }
    C3 f = bsdf_eval(&b, v3p(L), v3p(V), v3p(N));
    f_out[0] = f.r; f_out[1] = f.g; f_out[2] = f.b;
    *pdf_out = bsdf_pdf(&b, v3p(L), v3p(V), v3p(N));
}

void orc_bsdf_sample(int kind, const float *params, const float *V, const float *N, float u1, float u2,
                     float *dir_out, float *pdf_out, float *color_out) {
    Bsdf b; b.kind = kind; b.color = c3p(params); b.rough = params[3]; b.dist = (int)params[4];
    Sample3D s = kind >= ORC_BSDF_MIRROR ? delta_sample(&b, v3p(V), v3p(N)) : bsdf_sample(&b, v3p(V), v3p(N), u1, u2);
    memcpy(dir_out, s.ray.d, sizeof(float) * 3);
    *pdf_out = s.pdf;
    color_out[0] = s.color.r; color_out[1] = s.color.g; color_out[2] = s.color.b;
}

/* ------------------------------------------------------------------------ */
/* Kelemen albedo table: kelemenMaterial.cpp:113-178 (definition) and        */
/* getAlbedo 180-223 (lookup).                                               */
/* ------------------------------------------------------------------------ */
void orc_kelemen_table(float reflectance, int dist, int n_samples, uint32_t seed, float *out) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int e = 0; e < ORC_KELEMEN_COS * ORC_KELEMEN_ROUGH; ++e) {
        int i = e / ORC_KELEMEN_ROUGH, r = e % ORC_KELEMEN_ROUGH;
        float rough = fmaxf(0.00001f, (float)r / (ORC_KELEMEN_ROUGH - 1.f));
        uint32_t key = orc_path_key(seed, (uint32_t)e, 0u);
        double albedo = 0;
        for (int j = 0; j < n_samples; ++j) {
            float r1 = orc_slot_uniform(key, 3u * (uint32_t)j);
            float r2 = orc_slot_uniform(key, 3u * (uint32_t)j + 1u);
            float r3 = orc_slot_uniform(key, 3u * (uint32_t)j + 2u);
            float z;
            if (dist == ORC_DIST_GGX) z = cosf(atanf(rough * sqrtf(r1 / (1.f - r1))));
            else z = cosf(atanf(rough * sqrtf(-logf(1.f - r1))));
            float rr = sqrtf(fmaxf(0.f, 1.f - z * z));
            float phi = (float)(M_PI * 2 * r2);
            V3 hn = vnorm(v3(rr * cosf(phi), rr * sinf(phi), z));
            V3 nn = v3(0, 0, 1);
            z = fmaxf(0.0001f, (float)i / (ORC_KELEMEN_COS - 1));
            rr = sqrtf(fmaxf(0.f, 1.f - z * z));
            phi = (float)(r3 * 2 * M_PI);
            V3 vn = vnorm(v3(rr * cosf(phi), rr * sinf(phi), z));
            if (dot3(hn, vn) < 0.f) hn = vneg(hn);
            float vdoth = dot3(vn, hn);
            if (vdoth > 0.000001) {
                V3 ln = vnorm(vsub(vmul(vmul(hn, 2.f), vdoth), vn));
                float hdotl = dot3(hn, ln);
                float F = schlick(hdotl, reflectance);
                double a = F * fmaxf(0.f, dot3(ln, nn)) / vdoth;
                albedo += a;
            }
        }
        out[e] = (float)(albedo / n_samples);
    }
}

/* kelemenMaterial.cpp:180-223.  The reference reads past the table when
 * roughness >= 1 (rB = 32); indices are clamped here. */
float orc_kelemen_albedo(const float *tab, float costheta, float roughness) {
    int i = ref_floor(costheta * (ORC_KELEMEN_COS - 1));
    int ra = ref_floor(roughness * (ORC_KELEMEN_ROUGH - 1));
    if (i < 0) i = 0;
    if (i > ORC_KELEMEN_COS - 1) i = ORC_KELEMEN_COS - 1;
    if (ra < 0) ra = 0;
    if (ra > ORC_KELEMEN_ROUGH - 1) ra = ORC_KELEMEN_ROUGH - 1;
    int rb = ra + 1 > ORC_KELEMEN_ROUGH - 1 ? ra + 1 : ORC_KELEMEN_ROUGH - 1; /* max(): always 31 */
    if (rb > ORC_KELEMEN_ROUGH - 1) rb = ORC_KELEMEN_ROUGH - 1;
    int i1 = i + 1 < ORC_KELEMEN_COS - 1 ? i + 1 : ORC_KELEMEN_COS - 1;
    float va = tab[i * ORC_KELEMEN_ROUGH + ra];
    float vb = tab[i1 * ORC_KELEMEN_ROUGH + ra];
    float mix = costheta * (ORC_KELEMEN_COS - 1) - i;
    float res_a = (float)((1. - mix) * va + mix * vb);
    va = tab[i * ORC_KELEMEN_ROUGH + rb];
    vb = tab[i1 * ORC_KELEMEN_ROUGH + rb];
    float res_b = (float)((1. - mix) * va + mix * vb);
    float rmix = roughness * (ORC_KELEMEN_ROUGH - 1) - ra;
    return (float)((1. - rmix) * res_a + rmix * res_b);
}

/* ------------------------------------------------------------------------ */
/* materials: getBrdf                                                        */
/* ------------------------------------------------------------------------ */
/* What the shaders see of a hit (ShadingGeometry): P, the face-forwarded
 * interpolated normal Ng, the interpolated uv and the attribute. */
typedef struct { V3 P, Ng; float s, t; int attr; } Geo;

/* The FP32 albedo texture: the texel texture2DShader.h:28-31 reads at the
 * hit's uv (u - floor(u), 1 - (v - floor(v))) with Texture2D::read's nearest
 * texel and clamps (texture2D.cpp:48-66), without the FP16 rounding of
 * Imf::Rgba.  Returns the texel index, or -1 when the attribute has none. */
static long long ptex_index(const orc_scene *s, int attr, const Geo *g) {
    if (!s->ptex || !s->ptex_mat || !s->ptex_mat[attr]) return -1;
    float a = g->s - floorf(g->s), b = 1.f - (g->t - floorf(g->t));
    int x = to_int_x86(floorf(a * s->ptex_w)), y = to_int_x86(floorf(b * s->ptex_h));
    if (y == s->ptex_h) y -= 1;
    if (x == s->ptex_w) x -= 1;
    if (x < 0) x = 0;
    if (y < 0) y = 0;
    if (x > s->ptex_w - 1) x = s->ptex_w - 1;
    if (y > s->ptex_h - 1) y = s->ptex_h - 1;
    return (long long)y * s->ptex_w + x;
}

/* the diffuse colour of a Lambert lobe: albedo texture, per-triangle
 * override, or the material's (shader-driven) colour */
static C3 diffuse_color(const orc_scene *s, const orc_material *m, int attr, int prim, const Geo *g);

static C3 mat_color(const orc_scene *s, const orc_material *m, int slot, const float *k, const Geo *g) {
    if (m->shader[slot] >= 0) return shader_eval(s, m->shader[slot], g->P, g->Ng, g->s, g->t, g->attr);
    return c3p(k);
}
static float mat_float(const orc_scene *s, const orc_material *m, int slot, float k, const Geo *g) {
    if (m->shader[slot] >= 0) return shader_eval(s, m->shader[slot], g->P, g->Ng, g->s, g->t, g->attr).r;
    return k;
}

/* Returns 0 for the constant (light) material, 1 otherwise.  Shader-driven
 * parameters are read through ShadingEngine::getColor / getFloat. */
static int get_brdf(const orc_scene *s, int attr, int prim, V3 vn, V3 nn, int matte_path, float u_lobe, const Geo *g,
                    Bsdf *b) {
    const orc_material *m = &s->mats[attr];
    switch (m->type) {
    case ORC_MAT_MATTE: /* matteMaterial.cpp:30-35 */
        b->kind = ORC_BSDF_LAMBERT; b->rough = 0.f; b->dist = 0;
        b->color = diffuse_color(s, m, attr, prim, g);
        return 1;
    case ORC_MAT_METAL: { /* metalMaterial.cpp:32-62 */
        float mix = mat_float(s, m, ORC_SLOT_MIX, m->mix, g);
        C3 refl = mat_color(s, m, ORC_SLOT_SPEC, m->spec, g);
        C3 gain = mat_color(s, m, ORC_SLOT_GAIN, m->gain, g);
        float costheta = dot3(vn, nn);
        C3 sc = c3(schlick(costheta, refl.r), schlick(costheta, refl.g), schlick(costheta, refl.b));
        float rough = (u_lobe > mix) ? mat_float(s, m, ORC_SLOT_ROUGH, m->roughness, g)
                                     : mat_float(s, m, ORC_SLOT_ROUGHB, m->roughnessB, g);
        if (matte_path) rough = (float)(rough * (1. - CAUSTIC_BLUR) + CAUSTIC_BLUR);
        b->kind = ORC_BSDF_CT; b->dist = m->distribution; b->color = cmulc(sc, gain); b->rough = rough;
        return 1;
    }
    case ORC_MAT_KELEMEN: { /* kelemenMaterial.cpp:225-255 */
        C3 spec = mat_color(s, m, ORC_SLOT_SPEC, m->spec, g);
        float rough = mat_float(s, m, ORC_SLOT_ROUGH, m->roughness, g);
        if (matte_path) rough = (float)(rough * (1. - CAUSTIC_BLUR) + CAUSTIC_BLUR);
        if (m->reflectance < 1.f) {
            float albedo = orc_kelemen_albedo(m->table, fclampf(dot3(nn, vn), 0.f, 1.f), rough);
            if (u_lobe > albedo) {
                b->kind = ORC_BSDF_LAMBERT; b->rough = 0.f; b->dist = 0;
                b->color = diffuse_color(s, m, attr, prim, g);
                return 1;
            }
        }
        b->kind = ORC_BSDF_CT; b->dist = m->distribution; b->color = spec; b->rough = rough;
        return 1;
    }
    case ORC_MAT_GLASS: { /* glassMaterial.cpp:33-58 */
        b->rough = m->ior; b->dist = 0;
        if (m->reflectance == 1.f) { b->kind = ORC_BSDF_MIRROR; b->color = c3p(m->spec); return 1; }
        float costheta = (float)(dot3(vn, nn) * 0.5);
        if (costheta < 0.f) costheta = -costheta;
        float albedo = schlick(costheta, m->reflectance);
        if (u_lobe > albedo) { b->kind = ORC_BSDF_TRANSMIT; b->color = c3p(m->transmit); }
        else { b->kind = ORC_BSDF_MIRROR; b->color = c3p(m->spec); }
        return 1;
    }
    default: /* constantMaterial.cpp:19-24: pdf-0 sample terminates the path */
        return 0;
    }
}

static C3 diffuse_color(const orc_scene *s, const orc_material *m, int attr, int prim, const Geo *g) {
    long long ti = ptex_index(s, attr, g);
    if (ti >= 0) return c3p(s->ptex + 3 * ti);
    if (s->tri_albedo) return c3p(s->tri_albedo + 3 * prim);
    return mat_color(s, m, ORC_SLOT_DIFFUSE, m->diffuse, g);
}

/* math_utils.h:91-95 (the coordinateSystem frame, not normalised) */
static V3 world_to_tangent(V3 dir, V3 n) {
    V3 t, b;
    coordinate_system(n, &t, &b);
    return v3(dot3(dir, t), dot3(dir, b), dot3(dir, n));
}

/* specMirror.cpp:30-35, specTransmit.cpp:30-58 (b->rough = ior) */
static Sample3D delta_sample(const Bsdf *b, V3 vn, V3 nn) {
    Sample3D s;
    memset(&s, 0, sizeof s);
    V3 vt = world_to_tangent(vn, nn);
    if (b->kind == ORC_BSDF_MIRROR) {
        V3 r = vnorm(tangent_to_world(v3(-vt.x, -vt.y, vt.z), nn));
        s.ray.d[0] = r.x; s.ray.d[1] = r.y; s.ray.d[2] = r.z;
        s.ray.tmin = (float)RAY_BIAS; s.ray.tmax = 100000000.f;
        s.pdf = 1.f; s.color = b->color;
        return s;
    }
    float costheta = dot3(nn, vn);
    float ei = 1, et = b->rough;
    if (costheta < 0.) { ei = b->rough; et = 1.; }
    float sini2 = fmaxf(0.f, 1 - costheta * costheta);
    float eta = ei / et;
    float sint2 = eta * eta * sini2;
    if (sint2 >= 1.) { s.pdf = 0.f; return s; }
    float cost = sqrtf(fmaxf(0.f, 1 - sint2));
    if (costheta > 0.) cost = -cost;
    V3 t = vnorm(tangent_to_world(v3(eta * -vt.x, eta * -vt.y, cost), nn));
    s.ray.d[0] = t.x; s.ray.d[1] = t.y; s.ray.d[2] = t.z;
    s.ray.tmin = (float)RAY_BIAS; s.ray.tmax = 10000000.f;
    s.pdf = 1.f; s.color = b->color;
    return s;
}

/* integrator.cpp:28-31 */
static inline float power_heuristic(int nf, float fpdf, int ng, float gpdf) {
    float f = nf * fpdf, g = ng * gpdf;
    return (f * f) / (f * f + g * g);
}

static int occluded(const orc_scene *s, const orc_bvh *b, orc_ray *r) {
    orc_hit h;
    trace(s, b, r, 1, &h);
    return h.prim >= 0;
}

/* integrator.cpp:100-130 (with 33-54 and 56-98 inlined) */
static C3 integrate_direct(const orc_scene *s, const orc_bvh *bv, V3 P, V3 vn, V3 nn, const Bsdf *b,
                           const Rng *rng, int vslot, orc_counters *cnt) {
    C3 res = c3(0.f, 0.f, 0.f);
    int idx[64]; int n = 0;
    for (int i = 0; i < s->n_lights && n < 64; ++i)
        if (light_visible(&s->lights[i], P, nn)) idx[n++] = i;
    if (!n) return res;
    int pick = (int)(rng_u(rng, vslot + 1) * n);
    if (pick > n - 1) pick = n - 1;
    const orc_light *l = &s->lights[idx[pick]];
    if (b->kind == ORC_BSDF_LAMBERT) {
        Sample3D ls = light_sample(s, l, P, nn, rng_u(rng, vslot + 2), rng_u(rng, vslot + 3));
        float costheta = dot3(nn, v3p(ls.ray.d));
        if (costheta >= 0. && ls.pdf > 0.) {
            cnt->shadow++;
            if (!occluded(s, bv, &ls.ray)) {
                C3 f = bsdf_eval(b, v3p(ls.ray.d), vn, nn);
                res = cadd(res, cdivf(cmulf(cmulc(ls.color, f), costheta), ls.pdf));
            }
        }
    } else {
        Sample3D ls = light_sample(s, l, P, nn, rng_u(rng, vslot + 2), rng_u(rng, vslot + 3));
        float costheta = dot3(nn, v3p(ls.ray.d));
        if (costheta > 0. && ls.pdf > 0.) {
            float bpdf = bsdf_pdf(b, v3p(ls.ray.d), vn, nn);
            if (bpdf > 0.) {
                cnt->shadow++;
                if (!occluded(s, bv, &ls.ray)) {
                    float w = power_heuristic(1, ls.pdf, 1, bpdf);
                    C3 f = bsdf_eval(b, v3p(ls.ray.d), vn, nn);
                    res = cadd(res, cdivf(cmulf(cmulc(cmulf(ls.color, w), f), costheta), ls.pdf));
                }
            }
        }
        Sample3D bs = bsdf_sample(b, vn, nn, rng_u(rng, vslot + 4), rng_u(rng, vslot + 5));
        bs.ray.o[0] = P.x; bs.ray.o[1] = P.y; bs.ray.o[2] = P.z;
        costheta = dot3(nn, v3p(bs.ray.d));
        if (costheta > 0.) {
            float lpdf = light_pdf(s, l, &bs.ray);
            if (lpdf > 0. && bs.pdf > 0.) {
                cnt->shadow++;
                if (!occluded(s, bv, &bs.ray)) {
                    float w = power_heuristic(1, bs.pdf, 1, lpdf);
                    res = cadd(res, cdivf(cmulf(cmulc(cmulf(bs.color, w), light_eval(s, l, v3p(bs.ray.d))), costheta), bs.pdf));
                }
            }
        }
    }
    return cmulf(res, (float)n);
}

/* embreeAccelerator.cpp:41-94: hit point, interpolated face-forwarded normal
 * and interpolated uv */
static void hit_geometry(const orc_scene *s, const orc_ray *r, const orc_hit *h, Geo *g) {
    float t = h->t, u = h->u, v = h->v;
    float w = (float)(1. - u - v);
    g->P = vadd(v3p(r->o), vmul(v3p(r->d), t));
    const float *n = s->normals + 9 * h->prim;
    V3 nm = v3(w * n[0] + u * n[3] + v * n[6], w * n[1] + u * n[4] + v * n[7], w * n[2] + u * n[5] + v * n[8]);
    if (dot3(nm, v3p(r->d)) > 0.) nm = vmul(nm, -1.f);
    g->Ng = vnorm(nm);
    g->s = g->t = 0.f;
    if (s->uvs) {
        const float *q = s->uvs + 6 * h->prim;
        g->s = w * q[0] + u * q[2] + v * q[4];
        g->t = w * q[1] + u * q[3] + v * q[5];
    }
    g->attr = s->attr[h->prim];
}

/* Material::runNormalShader, material.cpp:19-33: the shading normal */
static V3 shading_normal(const orc_scene *s, const Geo *g) {
    const orc_material *m = &s->mats[g->attr];
    if (m->shader[ORC_SLOT_NORMAL] < 0) return g->Ng;
    C3 c = shader_eval(s, m->shader[ORC_SLOT_NORMAL], g->P, g->Ng, g->s, g->t, g->attr);
    V3 ns = tangent_to_world(v3(c.r, c.g, c.b), g->Ng);
    if (!(ns.x || ns.y || ns.z)) return g->Ng;
    return vnorm(ns);
}

/* camera.cpp:34-62 */
static void camera_ray(const orc_scene *s, int x, int y, float j0, float j1, orc_ray *r) {
    float fovx = s->fov;
    float fovy = ((float)s->height / (float)s->width) * fovx;
    float u = x + j0, v = y + j1;
    V3 d = v3(((2 * u - s->width) / (float)s->width) * fovx, ((-2 * v + s->height) / (float)s->height) * fovy, -1.f);
    d = vnorm(d);
    r->o[0] = r->o[1] = r->o[2] = 0.f;
    r->d[0] = d.x; r->d[1] = d.y; r->d[2] = d.z;
    r->tmin = 0.1f; r->tmax = 10000000.f;
}

enum { RAY_CAMERA, RAY_DIFFUSE, RAY_SPECULAR, RAY_MIRROR };

/* integrator.cpp:176-281: one path from the camera sample's first hit h (ray
 * `ray`); returns Lo after the firefly clamp */
static C3 path_from_hit(const orc_scene *s, const orc_bvh *bv, orc_ray ray, orc_hit h, const orc_ray *cam_ray,
                        const Rng *rng, orc_counters *cnt) {
    const orc_ray cam = *cam_ray;
    C3 Lo = c3(0.f, 0.f, 0.f), T = c3(1.f, 1.f, 1.f);
    float cp = 1.f;
    int matte_path = 0, ray_type = RAY_CAMERA, bounces = 0, true_bounces = 0;
    while (true_bounces < s->maxdepth) {
        int attr = s->attr[h.prim];
        Geo g;
        hit_geometry(s, &ray, &h, &g);
        V3 P = g.P, Nn = shading_normal(s, &g);
        /* integrator.cpp:132-150 */
        V3 vn = vneg(vnorm(v3p(ray.d)));
        if (dot3(vn, Nn) <= 0.) Nn = vnorm(cross3(vn, cross3(Nn, vn)));
        if (ray_type == RAY_DIFFUSE) matte_path = 1;
        if (ray_type == RAY_CAMERA || ray_type == RAY_MIRROR) Lo = cadd(Lo, cmulc(T, c3p(s->mats[attr].emission)));
        int vslot = 2 + ORC_SLOTS_PER_VERTEX * true_bounces;
        Bsdf b;
        int has_brdf = get_brdf(s, attr, h.prim, vn, Nn, matte_path, rng_u(rng, vslot + 0), &g, &b);
        if (!has_brdf) {
            /* the light-pick draw still happens (integrator.cpp:108-116); the
             * constant BRDF adds nothing and its sample has pdf 0 */
            break;
        }
        int delta = b.kind == ORC_BSDF_MIRROR || b.kind == ORC_BSDF_TRANSMIT; /* MirrorBrdf */
        if (bounces < s->maxdepth && !delta) { /* integrateDirect adds nothing for MirrorBrdf */
            C3 direct = integrate_direct(s, bv, P, vn, Nn, &b, rng, vslot, cnt);
            Lo = cadd(Lo, cmulc(direct, T));
        }
        Sample3D smp = delta ? delta_sample(&b, vn, Nn) : bsdf_sample(&b, vn, Nn, rng_u(rng, vslot + 6), rng_u(rng, vslot + 7));
        if (smp.pdf <= 0.f) break;
        if (!delta) {
            float ndl = dot3(Nn, v3p(smp.ray.d));
            T = cmulc(T, cdivf(cmulf(smp.color, ndl), smp.pdf));
            if (bounces > s->mindepth) cp *= clum(smp.color) * ndl / smp.pdf;
            if (cblack(T)) break;
            ray_type = (b.kind == ORC_BSDF_CT) ? RAY_SPECULAR : RAY_DIFFUSE;
        } else { /* integrator.cpp:231-237 */
            T = cmulc(T, smp.color);
            if (bounces > s->mindepth) cp *= clum(smp.color);
            ray_type = RAY_MIRROR;
        }
        if (bounces > s->mindepth) {
            if (rng_u(rng, vslot + 8) > cp) break;
            T = cdivf(T, cp);
        }
        smp.ray.o[0] = P.x; smp.ray.o[1] = P.y; smp.ray.o[2] = P.z;
        ++true_bounces;
        if (!delta) ++bounces;
        ray = smp.ray;
        cnt->ext++;
        trace(s, bv, &ray, 0, &h);
        if (h.prim < 0) {
            /* integrator.cpp:262-271: a mirror ray's miss adds each env light's
             * eval() in the direction of the camera sample */
            if (ray_type == RAY_MIRROR)
                for (int i = 0; i < s->n_lights; ++i)
                    if (s->lights[i].type == ORC_LIGHT_ENV)
                        Lo = cadd(Lo, cmulc(light_eval(s, &s->lights[i], v3p(cam.d)), T));
            break;
        }
        cnt->hit++;
    }
    /* integrator.cpp:277-281 */
    for (int k = 0; k < FIREFLY_MAX_STEPS && (Lo.r > FIREFLY || Lo.g > FIREFLY || Lo.b > FIREFLY); ++k) Lo = cmulf(Lo, 0.2f);
    return Lo;
}

/* integrator.cpp:152-300: the camera sample of pixel (x, y) -- camera ray from
 * rngs[0]'s camera slots, traced once -- and num_samples paths from its first
 * hit, path j drawing its vertex slots from rngs[j]; the result is the sum of
 * Lo_j / num_samples in order, a miss returns the env lights' eval.  The
 * renderer passes num_samples = 1 in pass 0 and LightSamples after it
 * (renderer.cpp:233, 257). */
static void camera_sample(const orc_scene *s, const orc_bvh *bv, int x, int y, const Rng *rngs, int num_samples,
                          float *out, orc_counters *cnt) {
    orc_ray ray, cam;
    camera_ray(s, x, y, rng_u(&rngs[0], 0), rng_u(&rngs[0], 1), &ray);
    cam = ray;
    cnt->paths += num_samples;
    orc_hit h;
    cnt->ext++;
    trace(s, bv, &ray, 0, &h);
    if (h.prim < 0) {
        C3 L = c3(0.f, 0.f, 0.f); float alpha = 0.f;
        for (int i = 0; i < s->n_lights; ++i)
            if (s->lights[i].type == ORC_LIGHT_ENV) { L = cadd(L, light_eval(s, &s->lights[i], v3p(ray.d))); alpha = 1.f; }
        out[0] = L.r; out[1] = L.g; out[2] = L.b; out[3] = alpha;
        return;
    }
    cnt->hit++;
    C3 res = c3(0.f, 0.f, 0.f);
    for (int j = 0; j < num_samples; ++j) {
        C3 Lo = path_from_hit(s, bv, ray, h, &cam, &rngs[j], cnt);
        if (num_samples == 1) res = cadd(res, Lo);
        else res = cadd(res, cdivf(Lo, (float)num_samples));
    }
    out[0] = res.r; out[1] = res.g; out[2] = res.b; out[3] = 1.f;
}

static void path_trace(const orc_scene *s, const orc_bvh *bv, int x, int y, const Rng *rng, float *out, orc_counters *cnt) {
    camera_sample(s, bv, x, y, rng, 1, out, cnt);
}

void orc_path(const orc_scene *s, const orc_bvh *b, int x, int y, uint32_t seed, uint32_t sample,
              const float *slots, float *rgba_out, orc_counters *cnt) {
    Rng r = make_rng(s, slots, seed, x, y, sample);
    path_trace(s, b, x, y, &r, rgba_out, cnt);
}

void orc_render(const orc_scene *s, const orc_bvh *b, uint32_t seed, int first_pass, int n_passes,
                float *accum, orc_counters *cnt, int n_threads) {
    const int tile = 16;
    int tx = (s->width + tile - 1) / tile, ty = (s->height + tile - 1) / tile;
    int64_t c_paths = 0, c_ext = 0, c_hit = 0, c_sh = 0;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#else
    (void)n_threads;
#endif
    for (int p = first_pass; p < first_pass + n_passes; ++p) {
        /* renderer.cpp:234-250: one parallel loop per pass; OpenMP dynamic
         * scheduling over 16x16 tiles stands in for TBB's parallel_for */
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : c_paths, c_ext, c_hit, c_sh)
        for (int t = 0; t < tx * ty; ++t) {
            orc_counters lc = {0, 0, 0, 0};
            int x0 = (t % tx) * tile, y0 = (t / tx) * tile;
            for (int y = y0; y < y0 + tile && y < s->height; ++y)
                for (int x = x0; x < x0 + tile && x < s->width; ++x) {
                    float o[4];
                    /* renderer.cpp:233, 257: one path in pass 0, LightSamples after */
                    const int ls = s->lightsamples > 1 ? s->lightsamples : 1;
                    const int n_sub = p == 0 ? 1 : ls;
                    Rng rs[n_sub];
                    for (int j = 0; j < n_sub; ++j) rs[j] = make_rng_sub(s, NULL, seed, x, y, (uint32_t)p, j, ls);
                    camera_sample(s, b, x, y, rs, n_sub, o, &lc);
                    float *a = accum + 4 * ((size_t)y * s->width + x);
                    a[0] += o[0]; a[1] += o[1]; a[2] += o[2]; a[3] += o[3];
                }
            c_paths += lc.paths; c_ext += lc.ext; c_hit += lc.hit; c_sh += lc.shadow;
        }
    }
    if (cnt) { cnt->paths += c_paths; cnt->ext += c_ext; cnt->hit += c_hit; cnt->shadow += c_sh; }
}

void orc_render_paths(const orc_scene *s, const orc_bvh *b, uint32_t seed, const int *pixel, const int *sample,
                      int n, const float *slots, float *out, orc_counters *cnt, int n_threads) {
    int nslots = 2 + ORC_SLOTS_PER_VERTEX * s->maxdepth;
    int64_t c_paths = 0, c_ext = 0, c_hit = 0, c_sh = 0;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#else
    (void)n_threads;
#endif
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : c_paths, c_ext, c_hit, c_sh)
    for (int i = 0; i < n; ++i) {
        orc_counters lc = {0, 0, 0, 0};
        Rng r = make_rng(s, slots ? slots + (size_t)i * nslots : NULL, seed, pixel[i] % s->width, pixel[i] / s->width,
                         (uint32_t)sample[i]);
        path_trace(s, b, pixel[i] % s->width, pixel[i] / s->width, &r, out + 4 * (size_t)i, &lc);
        c_paths += lc.paths; c_ext += lc.ext; c_hit += lc.hit; c_sh += lc.shadow;
    }
    if (cnt) { cnt->paths += c_paths; cnt->ext += c_ext; cnt->hit += c_hit; cnt->shadow += c_sh; }
}

/* ---- single-function entry points for the closed-form tests ---- */
void orc_camera_ray(const orc_scene *s, int x, int y, float j0, float j1, orc_ray *r) { camera_ray(s, x, y, j0, j1, r); }

void orc_light_sample(const orc_light *l, const float *P, const float *N, float x, float y, orc_ray *ray,
                      float *pdf, float *color) {
    Sample3D s = light_sample(NULL, l, v3p(P), v3p(N), x, y);
    *ray = s.ray;
    *pdf = s.pdf;
    color[0] = s.color.r; color[1] = s.color.g; color[2] = s.color.b;
}

float orc_light_pdf(const orc_light *l, orc_ray *ray) { return light_pdf(NULL, l, ray); }
void orc_light_sample_in(const orc_scene *sc, const orc_light *l, const float *P, const float *N, float x, float y, orc_ray *ray,
                      float *pdf, float *color) {
    Sample3D s = light_sample(sc, l, v3p(P), v3p(N), x, y);
    *ray = s.ray;
    *pdf = s.pdf;
    color[0] = s.color.r; color[1] = s.color.g; color[2] = s.color.b;
}

float orc_light_pdf_in(const orc_scene *sc, const orc_light *l, orc_ray *ray) { return light_pdf(sc, l, ray); }


int orc_light_visible(const orc_light *l, const float *P, const float *N) { return light_visible(l, v3p(P), v3p(N)); }

/* ------------------------------------------------------------------------ */
/* path-replay adjoint (CPU reference gradient)                              */
/*                                                                          */
/* The reference is not differentiable; this is the CPU counterpart of the  */
/* GPU adjoint: gradients of sum(dL/dI . I) w.r.t. per-triangle albedo,     */
/* Kelemen roughness and Kelemen spec weight (speccolor = (s, s, s)), for   */
/* the estimator path_trace() computes with fixed random numbers.  Sampling */
/* pdfs and MIS weights are held fixed; the Kelemen lobe choice adds the    */
/* score term dlogP(choice) * (radiance from this vertex on).               */
/* ------------------------------------------------------------------------ */
static double dlogd_drough(int dist, float rough, float ndoth) {
    double c2 = (double)ndoth * ndoth;
    if (c2 < 1e-12) c2 = 1e-12;
    double t2 = (1.0 - c2) / c2;
    if (t2 < 0) t2 = 0;
    if (dist == ORC_DIST_GGX) return 2.0 / rough - 4.0 * rough / ((double)rough * rough + t2);
    return 2.0 * t2 / ((double)rough * rough * rough) - 2.0 / rough;
}

/* derivative of orc_kelemen_albedo in roughness (linear inside a table cell) */
static double kelemen_dalbedo(const float *tab, float costheta, float roughness) {
    int i = ref_floor(costheta * (ORC_KELEMEN_COS - 1));
    int ra = ref_floor(roughness * (ORC_KELEMEN_ROUGH - 1));
    if (i < 0) i = 0;
    if (i > ORC_KELEMEN_COS - 1) i = ORC_KELEMEN_COS - 1;
    if (ra < 0) ra = 0;
    if (ra > ORC_KELEMEN_ROUGH - 1) ra = ORC_KELEMEN_ROUGH - 1;
    int rb = ORC_KELEMEN_ROUGH - 1;
    int i1 = i + 1 < ORC_KELEMEN_COS - 1 ? i + 1 : ORC_KELEMEN_COS - 1;
    double mix = costheta * (ORC_KELEMEN_COS - 1) - i;
    double a = (1. - mix) * tab[i * ORC_KELEMEN_ROUGH + ra] + mix * tab[i1 * ORC_KELEMEN_ROUGH + ra];
    double b = (1. - mix) * tab[i * ORC_KELEMEN_ROUGH + rb] + mix * tab[i1 * ORC_KELEMEN_ROUGH + rb];
    return (ORC_KELEMEN_ROUGH - 1) * (b - a);
}

typedef struct { double *albedo, *rough, *spec; float *ptex; } GradAcc;

/* One replay.  grad == 0: returns radiance in L (unclamped) and the clamp
 * step count.  grad == 1: accumulates gradients for incoming delta. */
static int path_replay(const orc_scene *s, const orc_bvh *bv, int x, int y, const Rng *rng, int grad, const double *delta,
                       const double *Ltot, double *Lout, GradAcc *g) {
    orc_ray ray;
    camera_ray(s, x, y, rng_u(rng, 0), rng_u(rng, 1), &ray);
    orc_hit h;
    orc_counters cnt = {0, 0, 0, 0};
    trace(s, bv, &ray, 0, &h);
    double L[3] = {0, 0, 0}, R[3];
    if (Ltot) { R[0] = Ltot[0]; R[1] = Ltot[1]; R[2] = Ltot[2]; }
    if (h.prim < 0) { Lout[0] = Lout[1] = Lout[2] = 0; return -1; }
    C3 T = c3(1.f, 1.f, 1.f);
    float cp = 1.f;
    int matte_path = 0, ray_type = RAY_CAMERA, bounces = 0, tb = 0;
    while (tb < s->maxdepth) {
        int attr = s->attr[h.prim], prim = h.prim;
        Geo hg;
        hit_geometry(s, &ray, &h, &hg);
        V3 P = hg.P, Nn = hg.Ng;
        V3 vn = vneg(vnorm(v3p(ray.d)));
        if (dot3(vn, Nn) <= 0.) Nn = vnorm(cross3(vn, cross3(Nn, vn)));
        if (ray_type == RAY_DIFFUSE) matte_path = 1;
        const orc_material *m = &s->mats[attr];
        if (ray_type == RAY_CAMERA) {
            C3 e = cmulc(T, c3p(m->emission));
            L[0] += e.r; L[1] += e.g; L[2] += e.b;
            if (grad) { R[0] -= e.r; R[1] -= e.g; R[2] -= e.b; }
        }
        int vslot = 2 + ORC_SLOTS_PER_VERTEX * tb;
        Bsdf b;
        float ulobe = rng_u(rng, vslot + 0);
        if (!get_brdf(s, attr, prim, vn, Nn, matte_path, ulobe, &hg, &b)) break;
        int kel = m->type == ORC_MAT_KELEMEN, spec = b.kind == ORC_BSDF_CT;
        /* the CT lobe's roughness is a parameter for Kelemen and for metal
         * (roughnessA = roughnessB; metalMaterial.cpp:34-62 picks the lobe by
         * u against mix, independent of roughness) */
        int rlobe = spec && (kel || m->type == ORC_MAT_METAL);
        /* the Lambert lobe's colour is a parameter: an albedo-texture texel
         * (gradients per texel) or the per-triangle override */
        long long pti = b.kind == ORC_BSDF_LAMBERT ? ptex_index(s, attr, &hg) : -1;
        int alb = b.kind == ORC_BSDF_LAMBERT && (pti >= 0 ? g->ptex != NULL : s->tri_albedo != NULL);
        const float *a_val = pti >= 0 ? s->ptex + 3 * pti : (s->tri_albedo ? s->tri_albedo + 3 * prim : NULL);
        double drough = matte_path && (m->type == ORC_MAT_KELEMEN || m->type == ORC_MAT_METAL) ? (1. - CAUSTIC_BLUR) : 1.;
        if (grad && kel && m->reflectance < 1.f) {
            float c = fclampf(dot3(Nn, vn), 0.f, 1.f);
            double A = orc_kelemen_albedo(m->table, c, b.kind == ORC_BSDF_CT ? b.rough :
                                          (float)(matte_path ? m->roughness * (1. - CAUSTIC_BLUR) + CAUSTIC_BLUR : m->roughness));
            float reff = (float)(matte_path ? m->roughness * (1. - CAUSTIC_BLUR) + CAUSTIC_BLUR : m->roughness);
            double dA = kelemen_dalbedo(m->table, c, reff);
            double dlogp = spec ? dA / (A > 1e-12 ? A : 1e-12) : -dA / (1 - A > 1e-12 ? 1 - A : 1e-12);
            g->rough[attr] += (delta[0] * R[0] + delta[1] * R[1] + delta[2] * R[2]) * dlogp * drough;
        }
        /* next-event estimation with separate light / BSDF parts */
        double cl[3] = {0, 0, 0}, cb[3] = {0, 0, 0}, dll = 0, dlb = 0;
        int idx[64], n = 0;
        for (int i = 0; i < s->n_lights && n < 64; ++i)
            if (light_visible(&s->lights[i], P, Nn)) idx[n++] = i;
        if (n) {
            int pick = (int)(rng_u(rng, vslot + 1) * n);
            if (pick > n - 1) pick = n - 1;
            const orc_light *l = &s->lights[idx[pick]];
            Sample3D ls = light_sample(s, l, P, Nn, rng_u(rng, vslot + 2), rng_u(rng, vslot + 3));
            float ct = dot3(Nn, v3p(ls.ray.d));
            if (!spec) {
                if (ct >= 0. && ls.pdf > 0. && !occluded(s, bv, &ls.ray)) {
                    C3 c = cdivf(cmulf(cmulc(ls.color, bsdf_eval(&b, v3p(ls.ray.d), vn, Nn)), ct), ls.pdf);
                    cl[0] = c.r; cl[1] = c.g; cl[2] = c.b;
                }
            } else {
                if (ct > 0. && ls.pdf > 0.) {
                    float bpdf = bsdf_pdf(&b, v3p(ls.ray.d), vn, Nn);
                    if (bpdf > 0. && !occluded(s, bv, &ls.ray)) {
                        float w = power_heuristic(1, ls.pdf, 1, bpdf);
                        C3 c = cdivf(cmulf(cmulc(cmulf(ls.color, w), bsdf_eval(&b, v3p(ls.ray.d), vn, Nn)), ct), ls.pdf);
                        cl[0] = c.r; cl[1] = c.g; cl[2] = c.b;
                        dll = dlogd_drough(b.dist, b.rough, dot3(Nn, vnorm(vadd(v3p(ls.ray.d), vn))));
                    }
                }
                Sample3D bs = bsdf_sample(&b, vn, Nn, rng_u(rng, vslot + 4), rng_u(rng, vslot + 5));
                bs.ray.o[0] = P.x; bs.ray.o[1] = P.y; bs.ray.o[2] = P.z;
                float ct2 = dot3(Nn, v3p(bs.ray.d));
                if (ct2 > 0.) {
                    float lpdf = light_pdf(s, l, &bs.ray);
                    if (lpdf > 0. && bs.pdf > 0. && !occluded(s, bv, &bs.ray)) {
                        float w = power_heuristic(1, bs.pdf, 1, lpdf);
                        C3 c = cdivf(cmulf(cmulc(cmulf(bs.color, w), light_eval(s, l, v3p(bs.ray.d))), ct2), bs.pdf);
                        cb[0] = c.r; cb[1] = c.g; cb[2] = c.b;
                        dlb = dlogd_drough(b.dist, b.rough, dot3(Nn, vnorm(vadd(v3p(bs.ray.d), vn))));
                    }
                }
            }
            double Tv[3] = {T.r, T.g, T.b};
            for (int k = 0; k < 3; ++k) { cl[k] *= n * Tv[k]; cb[k] *= n * Tv[k]; L[k] += cl[k] + cb[k]; }
            if (grad) {
                if (alb) {
                    for (int k = 0; k < 3; ++k) {
                        if (pti >= 0) {
#pragma omp atomic
                            g->ptex[3 * pti + k] += (float)(delta[k] * cl[k] / a_val[k]);
                        } else {
                            g->albedo[3 * prim + k] += delta[k] * cl[k] / a_val[k];
                        }
                    }
                }
                if (rlobe) {
                    double dc = 0, dlr = 0;
                    for (int k = 0; k < 3; ++k) { dc += delta[k] * (cl[k] + cb[k]); dlr += delta[k] * (cl[k] * dll + cb[k] * dlb); }
                    if (kel) g->spec[attr] += dc / m->spec[0];
                    g->rough[attr] += dlr * drough;
                }
                for (int k = 0; k < 3; ++k) R[k] -= cl[k] + cb[k];
            }
        }
        Sample3D smp = bsdf_sample(&b, vn, Nn, rng_u(rng, vslot + 6), rng_u(rng, vslot + 7));
        if (smp.pdf <= 0.f) break;
        float ndl = dot3(Nn, v3p(smp.ray.d));
        T = cmulc(T, cdivf(cmulf(smp.color, ndl), smp.pdf));
        if (bounces > s->mindepth) cp *= clum(smp.color) * ndl / smp.pdf;
        if (cblack(T)) break;
        ray_type = spec ? RAY_SPECULAR : RAY_DIFFUSE;
        if (bounces > s->mindepth) {
            if (rng_u(rng, vslot + 8) > cp) break;
            T = cdivf(T, cp);
        }
        if (grad) {
            double dr = delta[0] * R[0] + delta[1] * R[1] + delta[2] * R[2];
            if (alb) {
                for (int k = 0; k < 3; ++k) {
                    if (pti >= 0) {
#pragma omp atomic
                        g->ptex[3 * pti + k] += (float)(delta[k] * R[k] / a_val[k]);
                    } else {
                        g->albedo[3 * prim + k] += delta[k] * R[k] / a_val[k];
                    }
                }
            }
            if (rlobe) {
                if (kel) g->spec[attr] += dr / m->spec[0];
                g->rough[attr] += dr * dlogd_drough(b.dist, b.rough, dot3(Nn, vnorm(vadd(v3p(smp.ray.d), vn)))) * drough;
            }
        }
        smp.ray.o[0] = P.x; smp.ray.o[1] = P.y; smp.ray.o[2] = P.z;
        ++tb;
        ++bounces;
        ray = smp.ray;
        trace(s, bv, &ray, 0, &h);
        if (h.prim < 0) break;
    }
    (void)cnt;
    Lout[0] = L[0]; Lout[1] = L[1]; Lout[2] = L[2];
    int k = 0;
    float lc[3] = {(float)L[0], (float)L[1], (float)L[2]};
    while (k < FIREFLY_MAX_STEPS && (lc[0] > FIREFLY || lc[1] > FIREFLY || lc[2] > FIREFLY)) {
        lc[0] *= 0.2f; lc[1] *= 0.2f; lc[2] *= 0.2f; ++k;
    }
    return k;
}

void orc_adjoint(const orc_scene *s, const orc_bvh *b, uint32_t seed, int first_pass, int n_passes, const float *dl_dimage,
                 const float *albedo_tri, float *g_albedo_tri, float *g_rough_mat, float *g_spec_mat, float *g_ptex,
                 int n_threads) {
    (void)albedo_tri;
    int wh = s->width * s->height;
    long long n = (long long)wh * n_passes;
    int nt = n_threads > 0 ? n_threads : 1;
#ifdef _OPENMP
    omp_set_num_threads(nt);
#else
    nt = 1;
#endif
    double *acc = (double *)calloc((size_t)nt * (3 * (size_t)s->n_tris + 2 * (size_t)s->n_attrs), sizeof(double));
#pragma omp parallel
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *base = acc + (size_t)tid * (3 * (size_t)s->n_tris + 2 * (size_t)s->n_attrs);
        GradAcc g;
        g.albedo = base;
        g.rough = base + 3 * (size_t)s->n_tris;
        g.spec = g.rough + s->n_attrs;
        g.ptex = g_ptex;  /* summed with atomics (the texture can be large) */
#pragma omp for schedule(dynamic, 256)
        for (long long i = 0; i < n; ++i) {
            int pixel = (int)(i % wh), sample = first_pass + (int)(i / wh);
            /* a pixel without loss gradient adds exactly nothing */
            if (dl_dimage[4 * pixel] == 0.f && dl_dimage[4 * pixel + 1] == 0.f && dl_dimage[4 * pixel + 2] == 0.f) continue;
            Rng r = make_rng(s, NULL, seed, pixel % s->width, pixel / s->width, (uint32_t)sample);
            double L[3];
            int k = path_replay(s, b, pixel % s->width, pixel / s->width, &r, 0, NULL, NULL, L, &g);
            if (k < 0) continue;
            double cf = 1.0 / n_passes;
            float f = 1.f;
            for (int j = 0; j < k; ++j) f *= 0.2f;
            cf *= f;
            double delta[3] = {dl_dimage[4 * pixel] * cf, dl_dimage[4 * pixel + 1] * cf, dl_dimage[4 * pixel + 2] * cf};
            double L2[3];
            path_replay(s, b, pixel % s->width, pixel / s->width, &r, 1, delta, L, L2, &g);
        }
    }
    size_t np = 3 * (size_t)s->n_tris + 2 * (size_t)s->n_attrs;
    for (size_t j = 0; j < np; ++j) {
        double v = 0;
        for (int t = 0; t < nt; ++t) v += acc[(size_t)t * np + j];
        if (j < 3 * (size_t)s->n_tris) { if (g_albedo_tri) g_albedo_tri[j] = (float)v; }
        else if (j < 3 * (size_t)s->n_tris + s->n_attrs) { if (g_rough_mat) g_rough_mat[j - 3 * s->n_tris] = (float)v; }
        else if (g_spec_mat) g_spec_mat[j - 3 * s->n_tris - s->n_attrs] = (float)v;
    }
    free(acc);
}
