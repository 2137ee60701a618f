/*
 * oracle_subdiv.c -- CPU restatement of Loop subdivision
 * (src/geometry/loopSubdivMesh.cpp:167-744) on an indexed mesh, plus the
 * radial value-noise displacement and area-weighted normals used by the
 * displaced-sphere scenes.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the GPU subdivision kernels).
 *
 * Vertices are welded by exact position first.  Per level:
 *   even vertex: interior  (1 - n*beta) v + beta * sum(neighbours), beta = 1/16
 *                if n == 6 else 3/16 (n == 3) or 3/(8n)          (565-613, 734-740)
 *                boundary  3/4 v + 1/8 (b0 + b1)                  (639-692)
 *   odd vertex:  interior  3/8 (a + b) + 1/8 (c + d)              (376-378)
 *                boundary  1/2 (a + b)                             (369)
 *   faces: child k (k < 3) = (even v_k, odd e_k, odd e_{k-1}) in slots (k, k+1, k-1),
 *          child 3 = (odd e_0, odd e_1, odd e_2); child index 4*parent + k (294-305, 340-424)
 * Then limit positions (gamma = 1/(n + 3/(8 beta)); boundary 1/5, 439-458) and
 * limit normals normalize(cross(T, S)) over the ordered one-ring (496-541).
 *
 * Texture coordinates ride on face corners (a seam gives the two faces of an
 * edge different uvs at its ends), per level (174-191, 360-413, 565-613,
 * 639-692): an even corner takes the one-ring rule of its vertex with the
 * position's weights, unless uvs disagree around the vertex (two faces'
 * uvs at it, or two faces' uvs at one ring neighbour): then it keeps its uv;
 * an odd vertex takes 3/8, 3/8, 1/8, 1/8 when both faces agree on the edge's
 * end uvs, else each face the midpoint of its own two (boundary: midpoint).
 * uvs are not pushed to the limit (477-482).
 *
 * Neighbour sums here run in edge-creation order; the GPU walks rings face by
 * face.  Positions therefore agree to float rounding, not bit for bit.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "aurora_oracle.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct { int a, b, f0, f1, l0, l1; } Edge;

/* open-addressing map (min, max) -> edge index */
typedef struct { uint64_t *keys; int *vals; size_t cap; } EMap;

static uint64_t ekey(int a, int b) {
    unsigned lo = (unsigned)(a < b ? a : b), hi = (unsigned)(a < b ? b : a);
    return ((uint64_t)lo << 32) | hi;
}
static size_t ehash(uint64_t k, size_t cap) {
    k ^= k >> 33; k *= 0xff51afd7ed558ccdULL; k ^= k >> 33;
    return (size_t)(k & (cap - 1));
}
static void emap_init(EMap *m, size_t n) {
    size_t cap = 16;
    while (cap < 2 * n + 16) cap <<= 1;
    m->cap = cap;
    m->keys = (uint64_t *)malloc(sizeof(uint64_t) * cap);
    m->vals = (int *)malloc(sizeof(int) * cap);
    for (size_t i = 0; i < cap; ++i) m->keys[i] = UINT64_MAX;
}
static int *emap_slot(EMap *m, uint64_t k, int *found) {
    size_t i = ehash(k, m->cap);
    while (m->keys[i] != UINT64_MAX && m->keys[i] != k) i = (i + 1) & (m->cap - 1);
    *found = m->keys[i] == k;
    m->keys[i] = k;
    return &m->vals[i];
}
static void emap_free(EMap *m) { free(m->keys); free(m->vals); }

/* edges in face order; returns count or -1 on a non-manifold edge */
static int build_edges(int nf, const int *F, Edge *E, int *FE) {
    EMap m;
    emap_init(&m, (size_t)3 * nf);
    int ne = 0;
    for (int f = 0; f < nf; ++f)
        for (int j = 0; j < 3; ++j) {
            int a = F[3 * f + j], b = F[3 * f + (j + 1) % 3], found;
            int *slot = emap_slot(&m, ekey(a, b), &found);
            if (!found) {
                *slot = ne;
                E[ne].a = a; E[ne].b = b; E[ne].f0 = f; E[ne].f1 = -1; E[ne].l0 = j; E[ne].l1 = -1;
                FE[3 * f + j] = ne++;
            } else {
                Edge *e = &E[*slot];
                if (e->f1 >= 0) { emap_free(&m); return -1; }
                e->f1 = f; e->l1 = j;
                FE[3 * f + j] = *slot;
            }
        }
    emap_free(&m);
    return ne;
}

static float loop_beta(int n) { return n == 3 ? 3.f / 16.f : 3.f / (8.f * n); }

/* per-vertex valence, neighbour sum and boundary neighbours */
static void vertex_stats(int nv, int ne, const Edge *E, const float *P, int *val, float *sum, int *bnd, int *bn) {
    for (int v = 0; v < nv; ++v) { val[v] = 0; bnd[v] = 0; bn[2 * v] = bn[2 * v + 1] = -1; sum[3 * v] = sum[3 * v + 1] = sum[3 * v + 2] = 0.f; }
    for (int e = 0; e < ne; ++e) {
        int a = E[e].a, b = E[e].b;
        val[a]++; val[b]++;
        for (int k = 0; k < 3; ++k) { sum[3 * a + k] += P[3 * b + k]; sum[3 * b + k] += P[3 * a + k]; }
        if (E[e].f1 < 0) {
            bnd[a] = bnd[b] = 1;
            if (bn[2 * a] < 0) bn[2 * a] = b; else bn[2 * a + 1] = b;
            if (bn[2 * b] < 0) bn[2 * b] = a; else bn[2 * b + 1] = a;
        }
    }
}

static void ring_rule(const float *P, int v, int n, const float *sum, int boundary, const int *bn, float beta, float *out) {
    for (int k = 0; k < 3; ++k) {
        if (boundary) {
            out[k] = (1.f - 2.f * beta) * P[3 * v + k] + beta * P[3 * bn[2 * v] + k] + beta * P[3 * bn[2 * v + 1] + k];
        } else {
            out[k] = (1.f - n * beta) * P[3 * v + k] + beta * sum[3 * v + k];
        }
    }
}

/* ordered one-ring of v following the reference's face walk; faces_of: CSR */
static int ordered_ring(int v, const int *F, const int *off, const int *inc, int *ring, int *boundary) {
    int nfv = off[v + 1] - off[v];
    if (nfv == 0) return 0;
    /* local helpers */
#define LOCAL(f, vv) (F[3 * (f)] == (vv) ? 0 : (F[3 * (f) + 1] == (vv) ? 1 : 2))
    /* face across edge (v, x) other than f: the incident face containing x
     * in the slot after v (interior walk) */
    int start = inc[off[v]];
    /* detect boundary: walk forward */
    int f = start, n = 0;
    *boundary = 0;
    for (int guard = 0; guard <= nfv; ++guard) {
        int i = LOCAL(f, v);
        int nx = F[3 * f + (i + 1) % 3];
        ring[n++] = nx;
        int g = -1;
        for (int k = off[v]; k < off[v + 1]; ++k) {
            int h = inc[k];
            if (h == f) continue;
            int j = LOCAL(h, v);
            if (F[3 * h + (j + 2) % 3] == nx) { g = h; break; }
        }
        if (g < 0) { *boundary = 1; break; }
        f = g;
        if (f == start) return n;
    }
    /* boundary: forward to the last face, then prevVerts backward */
    f = start;
    for (int guard = 0; guard <= nfv; ++guard) {
        int i = LOCAL(f, v);
        int nx = F[3 * f + (i + 1) % 3], g = -1;
        for (int k = off[v]; k < off[v + 1]; ++k) {
            int h = inc[k];
            if (h == f) continue;
            int j = LOCAL(h, v);
            if (F[3 * h + (j + 2) % 3] == nx) { g = h; break; }
        }
        if (g < 0) break;
        f = g;
    }
    n = 0;
    ring[n++] = F[3 * f + (LOCAL(f, v) + 1) % 3];
    for (int guard = 0; guard <= nfv; ++guard) {
        int i = LOCAL(f, v);
        int pv = F[3 * f + (i + 2) % 3];
        ring[n++] = pv;
        int g = -1;
        for (int k = off[v]; k < off[v + 1]; ++k) {
            int h = inc[k];
            if (h == f) continue;
            int j = LOCAL(h, v);
            if (F[3 * h + (j + 1) % 3] == pv) { g = h; break; }
        }
        if (g < 0) break;
        f = g;
    }
#undef LOCAL
    return n;
}

static void faces_csr(int nv, int nf, const int *F, int **off_out, int **inc_out) {
    int *off = (int *)calloc((size_t)nv + 1, sizeof(int));
    int *inc = (int *)malloc(sizeof(int) * (size_t)3 * (nf > 0 ? nf : 1));
    for (int i = 0; i < 3 * nf; ++i) off[F[i] + 1]++;
    for (int v = 0; v < nv; ++v) off[v + 1] += off[v];
    int *cur = (int *)malloc(sizeof(int) * (size_t)(nv > 0 ? nv : 1));
    for (int v = 0; v < nv; ++v) cur[v] = off[v];
    for (int f = 0; f < nf; ++f)
        for (int j = 0; j < 3; ++j) inc[cur[F[3 * f + j]]++] = f;
    free(cur);
    *off_out = off;
    *inc_out = inc;
}

/* the uv of vertex v in face f */
static const float *fuv(const float *UV, const int *F, int f, int v) {
    int j = F[3 * f] == v ? 0 : (F[3 * f + 1] == v ? 1 : 2);
    return UV + 6 * f + 2 * j;
}

/* One level of corner uvs (UV: 6 per face of F) into UV2 (6 per child face). */
static void uv_level(int nv, int nf, const int *F, int ne, const Edge *E, const int *FE, const int *val, const int *bnd,
                     const float *UV, float *UV2) {
    float *first = (float *)malloc(sizeof(float) * 2 * (size_t)(nv > 0 ? nv : 1));
    float *sum = (float *)calloc(2 * (size_t)(nv > 0 ? nv : 1), sizeof(float));
    float *bsum = (float *)calloc(2 * (size_t)(nv > 0 ? nv : 1), sizeof(float));
    int *seen = (int *)calloc((size_t)(nv > 0 ? nv : 1), sizeof(int));
    int *split = (int *)calloc((size_t)(nv > 0 ? nv : 1), sizeof(int));
    for (int f = 0; f < nf; ++f)
        for (int j = 0; j < 3; ++j) {
            int v = F[3 * f + j];
            const float *u = UV + 6 * f + 2 * j;
            if (!seen[v]) { first[2 * v] = u[0]; first[2 * v + 1] = u[1]; seen[v] = 1; }
            else if (u[0] != first[2 * v] || u[1] != first[2 * v + 1]) split[v] = 1;
        }
    for (int e = 0; e < ne; ++e) {
        int a = E[e].a, b = E[e].b;
        const float *ua = fuv(UV, F, E[e].f0, a), *ub = fuv(UV, F, E[e].f0, b);
        if (E[e].f1 >= 0) {
            const float *ua1 = fuv(UV, F, E[e].f1, a), *ub1 = fuv(UV, F, E[e].f1, b);
            if (ub1[0] != ub[0] || ub1[1] != ub[1]) split[a] = 1;
            if (ua1[0] != ua[0] || ua1[1] != ua[1]) split[b] = 1;
        } else {
            bsum[2 * a] += ub[0]; bsum[2 * a + 1] += ub[1];
            bsum[2 * b] += ua[0]; bsum[2 * b + 1] += ua[1];
        }
        sum[2 * a] += ub[0]; sum[2 * a + 1] += ub[1];
        sum[2 * b] += ua[0]; sum[2 * b + 1] += ua[1];
    }
    for (int f = 0; f < nf; ++f)
        for (int j = 0; j < 3; ++j) {
            int v = F[3 * f + j], j1 = (j + 1) % 3, j2 = (j + 2) % 3, b = F[3 * f + j1];
            const float *own = UV + 6 * f + 2 * j, *ub = UV + 6 * f + 2 * j1;
            float ev[2] = {own[0], own[1]};
            if (val[v] > 0 && !split[v]) {
                int n = val[v];
                float beta = bnd[v] ? 1.f / 8.f : (n == 6 ? 1.f / 16.f : loop_beta(n));
                for (int k = 0; k < 2; ++k)
                    ev[k] = bnd[v] ? (1.f - 2.f * beta) * own[k] + beta * bsum[2 * v + k]
                                   : (1.f - n * beta) * own[k] + beta * sum[2 * v + k];
            }
            float *ce = UV2 + 6 * (4 * f + j) + 2 * j;
            ce[0] = ev[0]; ce[1] = ev[1];
            const Edge *ed = &E[FE[3 * f + j]];
            int g = ed->f0 == f ? ed->f1 : ed->f0;
            float od[2] = {0.5f * (own[0] + ub[0]), 0.5f * (own[1] + ub[1])};
            if (g >= 0) {
                const float *ga = fuv(UV, F, g, v), *gb = fuv(UV, F, g, b);
                if (ga[0] == own[0] && ga[1] == own[1] && gb[0] == ub[0] && gb[1] == ub[1]) {
                    const float *uc = UV + 6 * f + 2 * j2;
                    int d = F[3 * g] != v && F[3 * g] != b ? F[3 * g] : (F[3 * g + 1] != v && F[3 * g + 1] != b ? F[3 * g + 1] : F[3 * g + 2]);
                    const float *ud = fuv(UV, F, g, d);
                    for (int k = 0; k < 2; ++k) od[k] = 3.f / 8.f * (own[k] + ub[k]) + 1.f / 8.f * (uc[k] + ud[k]);
                }
            }
            float *o1 = UV2 + 6 * (4 * f + j) + 2 * j1, *o2 = UV2 + 6 * (4 * f + j1) + 2 * j, *o3 = UV2 + 6 * (4 * f + 3) + 2 * j;
            o1[0] = o2[0] = o3[0] = od[0];
            o1[1] = o2[1] = o3[1] = od[1];
        }
    free(first); free(sum); free(bsum); free(seen); free(split);
}

int orc_loop_subdivide(int nv_in, int nf_in, const float *Pin, const int *Fin, int levels, float *P_out, int *F_out,
                       float *N_out, int nv_cap, int nf_cap, int *nv_out, int *nf_out, const float *UVin, float *UV_out,
                       int push_limit) {
    /* weld by exact position (loopSubdivMesh.cpp:176-193) */
    int *remap = (int *)malloc(sizeof(int) * (size_t)(nv_in > 0 ? nv_in : 1));
    float *P = (float *)malloc(sizeof(float) * 3 * (size_t)(nv_in > 0 ? nv_in : 1));
    int nv = 0;
    for (int i = 0; i < nv_in; ++i) {
        int found = -1;
        for (int j = 0; j < nv; ++j)
            if (P[3 * j] == Pin[3 * i] && P[3 * j + 1] == Pin[3 * i + 1] && P[3 * j + 2] == Pin[3 * i + 2]) { found = j; break; }
        if (found < 0) { memcpy(&P[3 * nv], &Pin[3 * i], sizeof(float) * 3); found = nv++; }
        remap[i] = found;
    }
    int *F = (int *)malloc(sizeof(int) * 3 * (size_t)(nf_in > 0 ? nf_in : 1));
    float *UV = UVin ? (float *)malloc(sizeof(float) * 6 * (size_t)(nf_in > 0 ? nf_in : 1)) : NULL;
    int nf = 0;
    for (int f = 0; f < nf_in; ++f) {
        int a = remap[Fin[3 * f]], b = remap[Fin[3 * f + 1]], c = remap[Fin[3 * f + 2]];
        if (a == b || a == c || b == c) continue; /* degenerate: skipped */
        F[3 * nf] = a; F[3 * nf + 1] = b; F[3 * nf + 2] = c;
        if (UV) memcpy(UV + 6 * nf, UVin + 6 * f, sizeof(float) * 6);
        ++nf;
    }
    free(remap);
    int rc = 0;
    for (int lvl = 0; lvl < levels && rc == 0; ++lvl) {
        Edge *E = (Edge *)malloc(sizeof(Edge) * (size_t)3 * (nf > 0 ? nf : 1));
        int *FE = (int *)malloc(sizeof(int) * (size_t)3 * (nf > 0 ? nf : 1));
        int ne = build_edges(nf, F, E, FE);
        if (ne < 0) { free(E); free(FE); rc = -1; break; }
        int nv2 = nv + ne, nf2 = 4 * nf;
        float *P2 = (float *)malloc(sizeof(float) * 3 * (size_t)nv2);
        int *F2 = (int *)malloc(sizeof(int) * 3 * (size_t)nf2);
        int *val = (int *)malloc(sizeof(int) * (size_t)nv), *bnd = (int *)malloc(sizeof(int) * (size_t)nv);
        int *bn = (int *)malloc(sizeof(int) * 2 * (size_t)nv);
        float *sum = (float *)malloc(sizeof(float) * 3 * (size_t)nv);
        vertex_stats(nv, ne, E, P, val, sum, bnd, bn);
        for (int v = 0; v < nv; ++v) {
            if (val[v] == 0) { memcpy(&P2[3 * v], &P[3 * v], sizeof(float) * 3); continue; }
            float beta = bnd[v] ? 1.f / 8.f : (val[v] == 6 ? 1.f / 16.f : loop_beta(val[v]));
            ring_rule(P, v, val[v], sum, bnd[v], bn, beta, &P2[3 * v]);
        }
        for (int e = 0; e < ne; ++e) {
            const Edge *ed = &E[e];
            float *o = &P2[3 * (nv + e)];
            if (ed->f1 < 0) {
                for (int k = 0; k < 3; ++k) o[k] = 0.5f * (P[3 * ed->a + k] + P[3 * ed->b + k]);
            } else {
                int c = F[3 * ed->f0 + (ed->l0 + 2) % 3], d = F[3 * ed->f1 + (ed->l1 + 2) % 3];
                for (int k = 0; k < 3; ++k)
                    o[k] = 3.f / 8.f * (P[3 * ed->a + k] + P[3 * ed->b + k]) + 1.f / 8.f * (P[3 * c + k] + P[3 * d + k]);
            }
        }
        for (int f = 0; f < nf; ++f) {
            int v[3] = {F[3 * f], F[3 * f + 1], F[3 * f + 2]};
            int odd[3] = {nv + FE[3 * f], nv + FE[3 * f + 1], nv + FE[3 * f + 2]};
            for (int k = 0; k < 3; ++k) {
                int c = 4 * f + k;
                F2[3 * c + k] = v[k];
                F2[3 * c + (k + 1) % 3] = odd[k];
                F2[3 * c + (k + 2) % 3] = odd[(k + 2) % 3];
            }
            for (int k = 0; k < 3; ++k) F2[3 * (4 * f + 3) + k] = odd[k];
        }
        if (UV) {
            float *UV2 = (float *)malloc(sizeof(float) * 6 * (size_t)nf2);
            uv_level(nv, nf, F, ne, E, FE, val, bnd, UV, UV2);
            free(UV);
            UV = UV2;
        }
        free(E); free(FE); free(val); free(bnd); free(bn); free(sum);
        free(P); free(F);
        P = P2; F = F2; nv = nv2; nf = nf2;
    }
    if (rc == 0 && (nv > nv_cap || nf > nf_cap)) rc = -2;
    if (rc == 0) {
        /* limit positions and normals */
        float *PL = (float *)malloc(sizeof(float) * 3 * (size_t)(nv > 0 ? nv : 1));
        if (levels > 0 && push_limit) {
            Edge *E = (Edge *)malloc(sizeof(Edge) * (size_t)3 * (nf > 0 ? nf : 1));
            int *FE = (int *)malloc(sizeof(int) * (size_t)3 * (nf > 0 ? nf : 1));
            int ne = build_edges(nf, F, E, FE);
            int *val = (int *)malloc(sizeof(int) * (size_t)nv), *bnd = (int *)malloc(sizeof(int) * (size_t)nv);
            int *bn = (int *)malloc(sizeof(int) * 2 * (size_t)nv);
            float *sum = (float *)malloc(sizeof(float) * 3 * (size_t)nv);
            vertex_stats(nv, ne, E, P, val, sum, bnd, bn);
            for (int v = 0; v < nv; ++v) {
                if (val[v] == 0) { memcpy(&PL[3 * v], &P[3 * v], sizeof(float) * 3); continue; }
                float g = bnd[v] ? 1.f / 5.f : 1.f / (val[v] + 3.f / (8.f * loop_beta(val[v])));
                ring_rule(P, v, val[v], sum, bnd[v], bn, g, &PL[3 * v]);
            }
            free(E); free(FE); free(val); free(bnd); free(bn); free(sum);
        } else {
            memcpy(PL, P, sizeof(float) * 3 * (size_t)nv);
        }
        int *off, *inc;
        faces_csr(nv, nf, F, &off, &inc);
        int ring[64];
        for (int v = 0; v < nv; ++v) {
            int boundary = 0;
            int n = ordered_ring(v, F, off, inc, ring, &boundary);
            float S[3] = {0, 0, 0}, T[3] = {0, 0, 0};
            const float *p = &PL[3 * v];
            if (n == 0) { N_out[3 * v] = 0.f; N_out[3 * v + 1] = 1.f; N_out[3 * v + 2] = 0.f; continue; }
            if (!boundary) {
                for (int k = 0; k < n; ++k) {
                    float ang = (float)(2.f * M_PI * k / n);
                    for (int j = 0; j < 3; ++j) { S[j] += PL[3 * ring[k] + j] * cosf(ang); T[j] += PL[3 * ring[k] + j] * sinf(ang); }
                }
            } else {
                for (int j = 0; j < 3; ++j) S[j] = PL[3 * ring[n - 1] + j] - PL[3 * ring[0] + j];
                for (int j = 0; j < 3; ++j) {
                    if (n == 2) T[j] = PL[3 * ring[0] + j] + PL[3 * ring[1] + j] - 2 * p[j];
                    else if (n == 3) T[j] = PL[3 * ring[1] + j] - p[j];
                    else if (n == 4) T[j] = -PL[3 * ring[0] + j] + 2 * PL[3 * ring[1] + j] + 2 * PL[3 * ring[2] + j] - PL[3 * ring[3] + j] - 2 * p[j];
                    else {
                        float th = (float)(M_PI / (n - 1));
                        float t = sinf(th) * (PL[3 * ring[0] + j] + PL[3 * ring[n - 1] + j]);
                        for (int k = 1; k < n - 1; ++k) t += (2 * cosf(th) - 2) * sinf(k * th) * PL[3 * ring[k] + j];
                        T[j] = -t;
                    }
                }
            }
            float c[3] = {T[1] * S[2] - T[2] * S[1], T[2] * S[0] - T[0] * S[2], T[0] * S[1] - T[1] * S[0]};
            float l = sqrtf(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
            if (l == 0.f) l = 0.0001f;
            for (int j = 0; j < 3; ++j) N_out[3 * v + j] = c[j] * (1.f / l);
        }
        free(off); free(inc);
        memcpy(P_out, PL, sizeof(float) * 3 * (size_t)nv);
        memcpy(F_out, F, sizeof(int) * 3 * (size_t)nf);
        if (UV && UV_out) memcpy(UV_out, UV, sizeof(float) * 6 * (size_t)nf);
        free(PL);
    }
    *nv_out = nv;
    *nf_out = nf;
    free(P); free(F); free(UV);
    return rc;
}

/* ---- displacement: lattice value noise, fbm, radial offset ---- */
static uint32_t hash32(uint32_t v) {
    uint32_t s = v * 747796405u + 2891336453u;
    uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}
static float lattice(int ix, int iy, int iz, uint32_t seed) {
    uint32_t h = hash32(seed ^ hash32((uint32_t)ix * 73856093u ^ hash32((uint32_t)iy * 19349663u ^ hash32((uint32_t)iz * 83492791u))));
    return (float)(h >> 8) * (2.0f / 16777216.0f) - 1.0f;
}
static float smooth01(float t) { return t * t * (3.f - 2.f * t); }
static float value_noise(float x, float y, float z, uint32_t seed) {
    float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    float tx = smooth01(x - fx), ty = smooth01(y - fy), tz = smooth01(z - fz);
    float c[2][2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int d = 0; d < 2; ++d) c[a][b][d] = lattice(ix + a, iy + b, iz + d, seed);
    float x00 = c[0][0][0] + (c[1][0][0] - c[0][0][0]) * tx;
    float x10 = c[0][1][0] + (c[1][1][0] - c[0][1][0]) * tx;
    float x01 = c[0][0][1] + (c[1][0][1] - c[0][0][1]) * tx;
    float x11 = c[0][1][1] + (c[1][1][1] - c[0][1][1]) * tx;
    float y0 = x00 + (x10 - x00) * ty, y1 = x01 + (x11 - x01) * ty;
    return y0 + (y1 - y0) * tz;
}

/* p <- p * (1 + amplitude * fbm(p)), object space centred on the origin */
void orc_displace_radial(int nv, float *P, float amplitude, float frequency, int octaves, uint32_t seed) {
    for (int v = 0; v < nv; ++v) {
        float *p = &P[3 * v];
        float sum = 0.f, norm = 0.f, amp = 1.f, fr = frequency;
        for (int o = 0; o < octaves; ++o) {
            sum += amp * value_noise(p[0] * fr, p[1] * fr, p[2] * fr, seed + (uint32_t)o * 1013904223u);
            norm += amp;
            amp *= 0.5f;
            fr *= 2.f;
        }
        float n = norm > 0.f ? sum / norm : 0.f;
        float s = 1.f + amplitude * n;
        p[0] *= s; p[1] *= s; p[2] *= s;
    }
}

void orc_vertex_normals(int nv, int nf, const float *P, const int *F, float *N) {
    for (int i = 0; i < 3 * nv; ++i) N[i] = 0.f;
    for (int f = 0; f < nf; ++f) {
        const float *a = &P[3 * F[3 * f]], *b = &P[3 * F[3 * f + 1]], *c = &P[3 * F[3 * f + 2]];
        float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        float n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        for (int k = 0; k < 3; ++k)
            for (int j = 0; j < 3; ++j) N[3 * F[3 * f + k] + j] += n[j];
    }
    for (int v = 0; v < nv; ++v) {
        float *n = &N[3 * v];
        float l = sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        if (l == 0.f) { n[0] = 0.f; n[1] = 1.f; n[2] = 0.f; continue; }
        n[0] /= l; n[1] /= l; n[2] /= l;
    }
}
