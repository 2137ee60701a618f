// subdiv.hip -- Loop subdivision as GPU preprocessing.
//
// Restates `LoopSubdivMesh::LoopSubdivMesh` (src/geometry/loopSubdivMesh.cpp:167-562)
// on an indexed mesh with explicit edge records instead of pointer-linked
// SDVertex/SDFace objects:
//   - vertices are welded by exact position (the reference's std::set<Point>)
//   - per level: even vertices (one-ring rule with Warren's beta, boundary
//     1/8 rule), odd vertices (3/8, 3/8, 1/8, 1/8; boundary 1/2, 1/2), four
//     children per face with child k = 4*parent + k and the vertex order of
//     loopSubdivMesh.cpp:340-424, and the refined edge records
//   - limit positions (gamma rule, boundary 1/5) and limit normals
//     normalize(cross(T, S)) from the ordered one-ring (loopSubdivMesh.cpp:496-541)
// One-rings are walked face to face in the reference's order (SDVertex::oneRing,
// loopSubdivMesh.cpp:615-637), one thread per vertex.
//
// Optional radial value-noise displacement (a scene-file extension) follows,
// with area-weighted vertex normals recomputed from the displaced surface.
#include <map>
#include <utility>

#include "subdiv.h"

namespace aur {

#define AUR_MAX_RING 32

struct MeshDev {
    const float *P;
    const int *F;
    const int *FE;
    const int4 *E;   // (a, b, f0, f1)
    const int2 *EL;  // local edge index in f0, f1
    const int2 *VS;  // start (face, local index) per vertex
};

__device__ inline int other_face(const MeshDev &m, int e, int f) {
    int4 r = m.E[e];
    return r.z == f ? r.w : r.z;
}
__device__ inline int local_of(const MeshDev &m, int f, int v) {
    return m.F[3 * f] == v ? 0 : (m.F[3 * f + 1] == v ? 1 : 2);
}

// Ordered one-ring of vertex v (SDVertex::oneRing).  Returns the valence
// (-1 if it exceeds AUR_MAX_RING), sets *boundary.
__device__ int one_ring(const MeshDev &m, int v, int *ring, int *faces, bool *boundary) {
    int2 s = m.VS[v];
    int f = s.x, i = s.y, n = 0;
    *boundary = false;
    if (f < 0) return 0;
    // interior walk: record nextVert, cross the edge (v, nextVert)
    do {
        if (n >= AUR_MAX_RING) return -1;
        faces[n] = f;
        ring[n++] = m.F[3 * f + (i + 1) % 3];
        int g = other_face(m, m.FE[3 * f + i], f);
        if (g < 0) { *boundary = true; break; }
        i = local_of(m, g, v);
        f = g;
    } while (f != s.x);
    if (!*boundary) return n;
    // boundary: go to the last face forward, then collect prevVerts backward
    f = s.x; i = s.y;
    for (int guard = 0; guard < AUR_MAX_RING; ++guard) {
        int g = other_face(m, m.FE[3 * f + i], f);
        if (g < 0) break;
        i = local_of(m, g, v);
        f = g;
    }
    n = 0;
    ring[n++] = m.F[3 * f + (i + 1) % 3];
    for (;;) {
        if (n >= AUR_MAX_RING) return -1;
        faces[n - 1] = f;
        ring[n++] = m.F[3 * f + (i + 2) % 3];
        int g = other_face(m, m.FE[3 * f + (i + 2) % 3], f);
        if (g < 0) break;
        i = local_of(m, g, v);
        f = g;
    }
    return n;
}

__device__ inline f3 ld3(const float *P, int v) { return mk3(P[3 * v], P[3 * v + 1], P[3 * v + 2]); }
__device__ inline void st3(float *P, int v, f3 p) { P[3 * v] = p.x; P[3 * v + 1] = p.y; P[3 * v + 2] = p.z; }

// loopSubdivMesh.cpp:734-744
__device__ inline float loop_beta(int n) { return n == 3 ? 3.f / 16.f : 3.f / (8.f * n); }
__device__ inline float loop_gamma(int n) { return 1.f / (n + 3.f / (8.f * loop_beta(n))); }

// weightOneRing (loopSubdivMesh.cpp:565-613) / weightBoundary (639-692)
__device__ inline f3 weight_ring(const float *P, int v, const int *ring, int n, bool boundary, float beta) {
    f3 p = ld3(P, v);
    if (boundary) {
        f3 r = mul3(p, 1.f - 2.f * beta);
        r = add3(r, mul3(ld3(P, ring[0]), beta));
        r = add3(r, mul3(ld3(P, ring[n - 1]), beta));
        return r;
    }
    f3 r = mul3(p, 1.f - n * beta);
    for (int k = 0; k < n; ++k) r = add3(r, mul3(ld3(P, ring[k]), beta));
    return r;
}

__global__ void k_even(MeshDev m, int nv, float *__restrict__ Pn, int2 *__restrict__ VSn, int *__restrict__ err) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    int ring[AUR_MAX_RING], faces[AUR_MAX_RING];
    bool boundary;
    int n = one_ring(m, v, ring, faces, &boundary);
    int2 s = m.VS[v];
    if (n < 0) { atomicOr(err, 1); n = 0; }
    if (n == 0 || s.x < 0) {
        st3(Pn, v, ld3(m.P, v));
        VSn[v] = make_int2(-1, 0);
        return;
    }
    float beta;
    if (boundary) beta = 1.f / 8.f;
    else beta = (n == 6) ? 1.f / 16.f : loop_beta(n);
    st3(Pn, v, weight_ring(m.P, v, ring, n, boundary, beta));
    // vert->child->startFace = sf->children[sf->vnum(vert->P)]
    VSn[v] = make_int2(4 * s.x + s.y, s.y);
}

__global__ void k_odd(MeshDev m, int nv, int ne, float *__restrict__ Pn, int2 *__restrict__ VSn) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    int4 r = m.E[e];
    int2 l = m.EL[e];
    f3 a = ld3(m.P, r.x), b = ld3(m.P, r.y);
    f3 p;
    if (r.w < 0) {
        p = mul3(add3(a, b), 0.5f);
    } else {
        f3 c = ld3(m.P, m.F[3 * r.z + (l.x + 2) % 3]);
        f3 d = ld3(m.P, m.F[3 * r.w + (l.y + 2) % 3]);
        p = mul3(add3(a, b), 3.f / 8.f);
        p = add3(p, mul3(add3(c, d), 1.f / 8.f));
    }
    st3(Pn, nv + e, p);
    // vert->startFace = face->children[3]
    VSn[nv + e] = make_int2(4 * r.z + 3, l.x);
}

// ---------------------------------------------------------------------------
// texture coordinates per face corner (loopSubdivMesh.cpp:174-191, 360-413,
// 565-613, 639-692, 477-482): a corner's uv belongs to the face, so a seam
// keeps different uvs on its two sides.  One thread per parent corner (f, j):
//   even child corner: the one-ring rule with the position's weights, unless
//     the ring holds a seam (a face's uv at v differs from this corner's, or
//     the two faces of a ring edge disagree on the neighbour's uv): then the
//     corner keeps its uv;
//   odd vertex of edge (v_j, v_j+1): 3/8, 3/8, 1/8, 1/8 of this face's uvs
//     and the other face's opposite uv when both faces agree on the edge's
//     two uvs, else (and on a boundary) the midpoint of this face's two;
// uvs are not pushed to the limit surface.
// ---------------------------------------------------------------------------
__device__ inline float2 corner_uv(const MeshDev &m, const float2 *UV, int f, int v) {
    return UV[3 * f + local_of(m, f, v)];
}
__device__ inline bool same_uv(float2 a, float2 b) { return a.x == b.x && a.y == b.y; }

__global__ void k_uv(MeshDev m, int nf, const float2 *__restrict__ UV, float2 *__restrict__ UVn) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * nf) return;
    const int f = t / 3, j = t - 3 * f;
    const int v = m.F[3 * f + j];
    const float2 own = UV[3 * f + j];
    // even child corner: child j, slot j
    int ring[AUR_MAX_RING], faces[AUR_MAX_RING];
    bool boundary;
    const int n = one_ring(m, v, ring, faces, &boundary);
    float2 ev = own;
    if (n > 0) {
        const int nfv = boundary ? n - 1 : n;
        bool split = false;
        for (int k = 0; k < nfv; ++k) {
            if (!same_uv(corner_uv(m, UV, faces[k], v), own)) split = true;
            // the edge (v, w) shared by faces k and k+1
            if (boundary) {
                if (k + 1 < nfv && !same_uv(corner_uv(m, UV, faces[k], ring[k + 1]),
                                            corner_uv(m, UV, faces[k + 1], ring[k + 1])))
                    split = true;
            } else {
                const int k1 = (k + 1) % n;
                if (!same_uv(corner_uv(m, UV, faces[k], ring[k]), corner_uv(m, UV, faces[k1], ring[k])))
                    split = true;
            }
        }
        if (!split) {
            if (boundary) {
                const float beta = 1.f / 8.f;
                const float2 r0 = corner_uv(m, UV, faces[0], ring[0]);
                const float2 r1 = corner_uv(m, UV, faces[nfv - 1], ring[n - 1]);
                ev.x = (1.f - 2.f * beta) * own.x + beta * (r0.x + r1.x);
                ev.y = (1.f - 2.f * beta) * own.y + beta * (r0.y + r1.y);
            } else {
                const float beta = (n == 6) ? 1.f / 16.f : loop_beta(n);
                ev.x = (1.f - n * beta) * own.x;
                ev.y = (1.f - n * beta) * own.y;
                for (int k = 0; k < n; ++k) {
                    const float2 r = corner_uv(m, UV, faces[k], ring[k]);
                    ev.x += beta * r.x;
                    ev.y += beta * r.y;
                }
            }
        }
    }
    UVn[3 * (4 * f + j) + j] = ev;
    // odd vertex on the edge (v_j, v_j+1), seen from this face
    const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    const int b = m.F[3 * f + j1];
    const float2 ua = own, ub = UV[3 * f + j1];
    const int g = other_face(m, m.FE[3 * f + j], f);
    float2 od = make_float2(0.5f * (ua.x + ub.x), 0.5f * (ua.y + ub.y));
    if (g >= 0 && same_uv(corner_uv(m, UV, g, v), ua) && same_uv(corner_uv(m, UV, g, b), ub)) {
        const float2 uc = UV[3 * f + j2];
        const int lg = local_of(m, g, v), lb = local_of(m, g, b);
        const float2 ud = UV[3 * g + (3 - lg - lb)];
        od.x = 3.f / 8.f * (ua.x + ub.x) + 1.f / 8.f * (uc.x + ud.x);
        od.y = 3.f / 8.f * (ua.y + ub.y) + 1.f / 8.f * (uc.y + ud.y);
    }
    UVn[3 * (4 * f + j) + j1] = od;
    UVn[3 * (4 * f + j1) + j] = od;
    UVn[3 * (4 * f + 3) + j] = od;
}

__device__ inline int half_id(const MeshDev &m, int e, int at_vertex) {
    return 2 * e + (m.E[e].x == at_vertex ? 0 : 1);
}

__global__ void k_faces(MeshDev m, int nf, int nv, int ne, int *__restrict__ Fn, int *__restrict__ FEn) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    int v[3] = {m.F[3 * f], m.F[3 * f + 1], m.F[3 * f + 2]};
    int e[3] = {m.FE[3 * f], m.FE[3 * f + 1], m.FE[3 * f + 2]};
    int odd[3] = {nv + e[0], nv + e[1], nv + e[2]};
    for (int k = 0; k < 3; ++k) {
        int c = 4 * f + k;
        int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
        Fn[3 * c + k] = v[k];
        Fn[3 * c + k1] = odd[k];
        Fn[3 * c + k2] = odd[k2];
        FEn[3 * c + k] = half_id(m, e[k], v[k]);
        FEn[3 * c + k1] = 2 * ne + 3 * f + k2;
        FEn[3 * c + k2] = half_id(m, e[k2], v[k]);
    }
    int c = 4 * f + 3;
    for (int k = 0; k < 3; ++k) {
        Fn[3 * c + k] = odd[k];
        FEn[3 * c + k] = 2 * ne + 3 * f + k;
    }
}

__global__ void k_edges(MeshDev m, int nf, int nv, int ne, int4 *__restrict__ En, int2 *__restrict__ ELn) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < ne) {
        int e = t;
        int4 r = m.E[e];
        int2 l = m.EL[e];
        int fs[2] = {r.z, r.w}, ls[2] = {l.x, l.y};
        int4 h0 = make_int4(r.x, nv + e, -1, -1), h1 = make_int4(nv + e, r.y, -1, -1);
        int2 l0 = make_int2(-1, -1), l1 = make_int2(-1, -1);
        for (int s = 0; s < 2; ++s) {
            if (fs[s] < 0) continue;
            int f = fs[s], k = ls[s];
            bool a_first = m.F[3 * f + k] == r.x;  // does local edge k start at endpoint a?
            int ca = a_first ? k : (k + 1) % 3, cb = a_first ? (k + 1) % 3 : k;
            if (s == 0) { h0.z = 4 * f + ca; l0.x = k; h1.z = 4 * f + cb; l1.x = k; }
            else { h0.w = 4 * f + ca; l0.y = k; h1.w = 4 * f + cb; l1.y = k; }
        }
        En[2 * e] = h0; ELn[2 * e] = l0;
        En[2 * e + 1] = h1; ELn[2 * e + 1] = l1;
    } else if (t < ne + 3 * nf) {
        int q = t - ne;
        int f = q / 3, j = q % 3;
        int a = nv + m.FE[3 * f + j], b = nv + m.FE[3 * f + (j + 1) % 3];
        En[2 * ne + q] = make_int4(a, b, 4 * f + 3, 4 * f + (j + 1) % 3);
        ELn[2 * ne + q] = make_int2(j, (j + 2) % 3);
    }
}

// loopSubdivMesh.cpp:439-458: push every vertex to the limit surface
__global__ void k_limit(MeshDev m, int nv, float *__restrict__ Pl, int *__restrict__ err) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    int ring[AUR_MAX_RING], faces[AUR_MAX_RING];
    bool boundary;
    int n = one_ring(m, v, ring, faces, &boundary);
    if (n < 0) { atomicOr(err, 1); n = 0; }
    if (n == 0) { st3(Pl, v, ld3(m.P, v)); return; }
    float beta = boundary ? 1.f / 5.f : loop_gamma(n);
    st3(Pl, v, weight_ring(m.P, v, ring, n, boundary, beta));
}

// loopSubdivMesh.cpp:496-541: limit normals from the ordered ring of limit points
__global__ void k_limit_normals(MeshDev m, int nv, float *__restrict__ N) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    int ring[AUR_MAX_RING], faces[AUR_MAX_RING];
    bool boundary;
    int n = one_ring(m, v, ring, faces, &boundary);
    if (n <= 0) { st3(N, v, mk3(0.f, 1.f, 0.f)); return; }
    f3 S = mk3(0.f, 0.f, 0.f), T = mk3(0.f, 0.f, 0.f);
    f3 p = ld3(m.P, v);
    if (!boundary) {
        for (int k = 0; k < n; ++k) {
            float ang = (2.f * AUR_PI_F * k) / n;
            f3 q = ld3(m.P, ring[k]);
            S = add3(S, mul3(q, cosf(ang)));
            T = add3(T, mul3(q, sinf(ang)));
        }
    } else {
        S = sub3(ld3(m.P, ring[n - 1]), ld3(m.P, ring[0]));
        if (n == 2) T = sub3(add3(ld3(m.P, ring[0]), ld3(m.P, ring[1])), mul3(p, 2.f));
        else if (n == 3) T = sub3(ld3(m.P, ring[1]), p);
        else if (n == 4)
            T = add3(add3(add3(add3(mul3(ld3(m.P, ring[0]), -1.f), mul3(ld3(m.P, ring[1]), 2.f)), mul3(ld3(m.P, ring[2]), 2.f)),
                          mul3(ld3(m.P, ring[3]), -1.f)), mul3(p, -2.f));
        else {
            float theta = AUR_PI_F / (float)(n - 1);
            T = mul3(add3(ld3(m.P, ring[0]), ld3(m.P, ring[n - 1])), sinf(theta));
            for (int k = 1; k < n - 1; ++k) {
                float wt = (2.f * cosf(theta) - 2.f) * sinf(k * theta);
                T = add3(T, mul3(ld3(m.P, ring[k]), wt));
            }
            T = neg3(T);
        }
    }
    st3(N, v, norm3(cross3(T, S)));
}

// ---------------------------------------------------------------------------
// radial value-noise displacement (scene-file extension "displace")
// ---------------------------------------------------------------------------
__device__ __host__ inline float lattice_value(int ix, int iy, int iz, uint32_t seed) {
    uint32_t h = pcg_hash(seed ^ pcg_hash((uint32_t)ix * 73856093u ^ pcg_hash((uint32_t)iy * 19349663u ^ pcg_hash((uint32_t)iz * 83492791u))));
    return (float)(h >> 8) * (2.0f / 16777216.0f) - 1.0f;
}
__device__ inline float smooth01(float t) { return t * t * (3.f - 2.f * t); }
__device__ inline float value_noise(f3 p, uint32_t seed) {
    float fx = floorf(p.x), fy = floorf(p.y), fz = floorf(p.z);
    int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    float tx = smooth01(p.x - fx), ty = smooth01(p.y - fy), tz = smooth01(p.z - fz);
    float c[2][2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int d = 0; d < 2; ++d) c[a][b][d] = lattice_value(ix + a, iy + b, iz + d, seed);
    float x00 = c[0][0][0] + (c[1][0][0] - c[0][0][0]) * tx;
    float x10 = c[0][1][0] + (c[1][1][0] - c[0][1][0]) * tx;
    float x01 = c[0][0][1] + (c[1][0][1] - c[0][0][1]) * tx;
    float x11 = c[0][1][1] + (c[1][1][1] - c[0][1][1]) * tx;
    float y0 = x00 + (x10 - x00) * ty;
    float y1 = x01 + (x11 - x01) * ty;
    return y0 + (y1 - y0) * tz;
}

__global__ void k_displace(int nv, float *__restrict__ P, float amplitude, float frequency, int octaves, uint32_t seed) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    f3 p = ld3(P, v);
    float sum = 0.f, norm = 0.f, amp = 1.f, fr = frequency;
    for (int o = 0; o < octaves; ++o) {
        sum += amp * value_noise(mul3(p, fr), seed + (uint32_t)o * 1013904223u);
        norm += amp;
        amp *= 0.5f;
        fr *= 2.f;
    }
    float n = norm > 0.f ? sum / norm : 0.f;
    st3(P, v, mul3(p, 1.f + amplitude * n));
}

// area-weighted vertex normals: sum of face cross products around the ring
__global__ void k_area_normals(MeshDev m, int nv, float *__restrict__ N) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    int ring[AUR_MAX_RING], faces[AUR_MAX_RING];
    bool boundary;
    int n = one_ring(m, v, ring, faces, &boundary);
    int nfaces = boundary ? n - 1 : n;
    f3 acc = mk3(0.f, 0.f, 0.f);
    for (int k = 0; k < nfaces; ++k) {
        int f = faces[k];
        f3 a = ld3(m.P, m.F[3 * f]), b = ld3(m.P, m.F[3 * f + 1]), c = ld3(m.P, m.F[3 * f + 2]);
        acc = add3(acc, cross3(sub3(b, a), sub3(c, a)));
    }
    st3(N, v, n > 0 ? norm3(acc) : mk3(0.f, 1.f, 0.f));
}

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------
struct LevelBufs {
    DevBuf P, F, FE, E, EL, VS;
    void swap(LevelBufs &o) {
        P.swap(o.P); F.swap(o.F); FE.swap(o.FE); E.swap(o.E); EL.swap(o.EL); VS.swap(o.VS);
    }
};

static MeshDev view(const LevelBufs &b) {
    MeshDev m;
    m.P = b.P.as<float>();
    m.F = b.F.as<int>();
    m.FE = b.FE.as<int>();
    m.E = b.E.as<int4>();
    m.EL = b.EL.as<int2>();
    m.VS = b.VS.as<int2>();
    return m;
}

template <class T>
static void h2d(Ctx *c, DevBuf &d, const std::vector<T> &h) {
    d.reserve(sizeof(T) * std::max<size_t>(h.size(), 1));
    if (!h.empty()) HIP_CHECK(hipMemcpyAsync(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, c->stream));
}

static inline unsigned grid_for(size_t n) { return (unsigned)((n + 255) / 256); }

LoopMesh loop_subdivide(Ctx *c, const std::vector<float> &Pin, const std::vector<int> &Fin, int levels, const Displace &disp,
                        const std::vector<float> *UVin) {
    if (levels < 0 || levels > 10) throw std::invalid_argument("subdivision levels must be in 0..10");
    const int nv0 = (int)(Pin.size() / 3);
    // weld by exact position (loopSubdivMesh.cpp:176-193)
    std::map<std::tuple<float, float, float>, int> weld;
    std::vector<int> remap(nv0);
    std::vector<float> P;
    for (int i = 0; i < nv0; ++i) {
        auto key = std::make_tuple(Pin[3 * i], Pin[3 * i + 1], Pin[3 * i + 2]);
        auto it = weld.find(key);
        if (it == weld.end()) {
            int id = (int)(P.size() / 3);
            weld[key] = id;
            remap[i] = id;
            P.insert(P.end(), &Pin[3 * i], &Pin[3 * i] + 3);
        } else {
            remap[i] = it->second;
        }
    }
    int nv = (int)(P.size() / 3);
    // faces, skipping degenerate ones (loopSubdivMesh.cpp:200-219); corner
    // uvs (6 per input face) follow the faces kept
    const bool has_uv = UVin != nullptr;
    if (has_uv && UVin->size() != 2 * Fin.size()) throw std::invalid_argument("corner uvs do not match the faces");
    std::vector<int> F;
    std::vector<float> FUV;
    for (size_t f = 0; f + 2 < Fin.size(); f += 3) {
        int a = Fin[f], b = Fin[f + 1], d = Fin[f + 2];
        if (a < 0 || b < 0 || d < 0 || a >= nv0 || b >= nv0 || d >= nv0) throw std::invalid_argument("face index out of range");
        a = remap[a]; b = remap[b]; d = remap[d];
        if (a == b || a == d || b == d) continue;
        F.push_back(a); F.push_back(b); F.push_back(d);
        if (has_uv) FUV.insert(FUV.end(), UVin->begin() + 2 * f, UVin->begin() + 2 * f + 6);
    }
    int nf = (int)(F.size() / 3);
    // edges in face order; neighbour pointers (loopSubdivMesh.cpp:223-253)
    std::map<std::pair<int, int>, int> emap;
    std::vector<int> E, EL, FE(3 * (size_t)nf);
    for (int f = 0; f < nf; ++f) {
        for (int j = 0; j < 3; ++j) {
            int a = F[3 * f + j], b = F[3 * f + (j + 1) % 3];
            auto key = std::make_pair(std::min(a, b), std::max(a, b));
            auto it = emap.find(key);
            if (it == emap.end()) {
                int id = (int)(E.size() / 4);
                emap[key] = id;
                E.insert(E.end(), {a, b, f, -1});
                EL.insert(EL.end(), {j, -1});
                FE[3 * f + j] = id;
            } else {
                int id = it->second;
                if (E[4 * id + 3] >= 0) throw std::invalid_argument("non-manifold edge in subdivision mesh");
                E[4 * id + 3] = f;
                EL[2 * id + 1] = j;
                FE[3 * f + j] = id;
            }
        }
    }
    int ne = (int)(E.size() / 4);
    // start faces: the last face that references the vertex (loopSubdivMesh.cpp:209-215)
    std::vector<int> VS(2 * (size_t)nv, -1);
    for (int f = 0; f < nf; ++f)
        for (int j = 0; j < 3; ++j) { VS[2 * F[3 * f + j]] = f; VS[2 * F[3 * f + j] + 1] = j; }
    for (int v = 0; v < nv; ++v) if (VS[2 * v] < 0) VS[2 * v + 1] = 0;

    LevelBufs cur, nxt;
    DevBuf uv_cur, uv_nxt;
    if (has_uv) h2d(c, uv_cur, FUV);
    h2d(c, cur.P, P);
    h2d(c, cur.F, F);
    h2d(c, cur.FE, FE);
    h2d(c, cur.E, E);
    h2d(c, cur.EL, EL);
    h2d(c, cur.VS, VS);
    DevBuf err;
    err.reserve(sizeof(int));
    fill_zero(c, err.p, sizeof(int));
    for (int lvl = 0; lvl < levels; ++lvl) {
        size_t nv2 = (size_t)nv + ne, nf2 = 4 * (size_t)nf, ne2 = 2 * (size_t)ne + 3 * (size_t)nf;
        if (nv2 > 0x7fffffff / 3 || nf2 > 0x7fffffff / 3 || ne2 > 0x7fffffff / 4) throw CapacityError("subdivided mesh too large");
        nxt.P.reserve(sizeof(float) * 3 * nv2);
        nxt.VS.reserve(sizeof(int2) * nv2);
        nxt.F.reserve(sizeof(int) * 3 * nf2);
        nxt.FE.reserve(sizeof(int) * 3 * nf2);
        nxt.E.reserve(sizeof(int4) * ne2);
        nxt.EL.reserve(sizeof(int2) * ne2);
        MeshDev m = view(cur);
        if (nv) hipLaunchKernelGGL(k_even, dim3(grid_for(nv)), dim3(256), 0, c->stream, m, nv, nxt.P.as<float>(), nxt.VS.as<int2>(), err.as<int>());
        if (ne) hipLaunchKernelGGL(k_odd, dim3(grid_for(ne)), dim3(256), 0, c->stream, m, nv, ne, nxt.P.as<float>(), nxt.VS.as<int2>());
        if (nf) hipLaunchKernelGGL(k_faces, dim3(grid_for(nf)), dim3(256), 0, c->stream, m, nf, nv, ne, nxt.F.as<int>(), nxt.FE.as<int>());
        if (ne + 3 * nf) hipLaunchKernelGGL(k_edges, dim3(grid_for(ne + 3 * (size_t)nf)), dim3(256), 0, c->stream, m, nf, nv, ne, nxt.E.as<int4>(), nxt.EL.as<int2>());
        if (has_uv && nf) {
            uv_nxt.reserve(sizeof(float2) * 3 * nf2);
            hipLaunchKernelGGL(k_uv, dim3(grid_for(3 * (size_t)nf)), dim3(256), 0, c->stream, m, nf, uv_cur.as<float2>(),
                               uv_nxt.as<float2>());
            uv_cur.swap(uv_nxt);
        }
        HIP_CHECK(hipGetLastError());
        cur.swap(nxt);
        nv = (int)nv2; nf = (int)nf2; ne = (int)ne2;
    }
    DevBuf Pl, N;
    Pl.reserve(sizeof(float) * 3 * std::max(nv, 1));
    N.reserve(sizeof(float) * 3 * std::max(nv, 1));
    MeshDev m = view(cur);
    if (levels > 0) {
        if (nv) hipLaunchKernelGGL(k_limit, dim3(grid_for(nv)), dim3(256), 0, c->stream, m, nv, Pl.as<float>(), err.as<int>());
    } else {
        HIP_CHECK(hipMemcpyAsync(Pl.p, cur.P.p, sizeof(float) * 3 * nv, hipMemcpyDeviceToDevice, c->stream));
    }
    MeshDev ml = m;
    ml.P = Pl.as<float>();
    if (disp.on) {
        if (nv) hipLaunchKernelGGL(k_displace, dim3(grid_for(nv)), dim3(256), 0, c->stream, nv, Pl.as<float>(), disp.amplitude,
                                   disp.frequency, disp.octaves, disp.seed);
        if (nv) hipLaunchKernelGGL(k_area_normals, dim3(grid_for(nv)), dim3(256), 0, c->stream, ml, nv, N.as<float>());
    } else if (nv) {
        hipLaunchKernelGGL(k_limit_normals, dim3(grid_for(nv)), dim3(256), 0, c->stream, ml, nv, N.as<float>());
    }
    HIP_CHECK(hipGetLastError());
    LoopMesh out;
    out.P.resize(3 * (size_t)nv);
    out.N.resize(3 * (size_t)nv);
    out.F.resize(3 * (size_t)nf);
    int herr = 0;
    if (nv) {
        HIP_CHECK(hipMemcpyAsync(out.P.data(), Pl.p, sizeof(float) * out.P.size(), hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipMemcpyAsync(out.N.data(), N.p, sizeof(float) * out.N.size(), hipMemcpyDeviceToHost, c->stream));
    }
    if (nf) HIP_CHECK(hipMemcpyAsync(out.F.data(), cur.F.p, sizeof(int) * out.F.size(), hipMemcpyDeviceToHost, c->stream));
    if (has_uv) {
        out.UV.resize(6 * (size_t)nf);
        if (nf) HIP_CHECK(hipMemcpyAsync(out.UV.data(), uv_cur.p, sizeof(float) * out.UV.size(), hipMemcpyDeviceToHost, c->stream));
    }
    HIP_CHECK(hipMemcpyAsync(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    if (herr) throw CapacityError("vertex valence exceeds the one-ring limit of 32");
    return out;
}

SubdivResult subdivide_object(Ctx *c, const HostObject &o) {
    size_t ncorner = o.P.size() / 3;
    std::vector<int> F(ncorner);
    for (size_t i = 0; i < ncorner; ++i) F[i] = (int)i;
    const bool uv = !o.UV.empty();
    if (uv && o.UV.size() != 2 * ncorner) throw std::invalid_argument("corner uvs do not match the corners");
    LoopMesh lm = loop_subdivide(c, o.P, F, o.subdlevels, o.displace, uv ? &o.UV : nullptr);
    SubdivResult r;
    r.corner_uvs = uv ? lm.UV : std::vector<float>(6 * (lm.F.size() / 3), 0.f);
    size_t nv = lm.P.size() / 3, nf = lm.F.size() / 3;
    r.verts.resize(3 * nv);
    r.tris = lm.F;
    r.corner_normals.resize(9 * nf);
    // TriangleMesh::makeEmbree: points and normals through objectToCamera
    for (size_t v = 0; v < nv; ++v) transform_point(o.o2c, &lm.P[3 * v], &r.verts[3 * v]);
    std::vector<float> Nc(3 * nv);
    for (size_t v = 0; v < nv; ++v) transform_vector(o.o2c, &lm.N[3 * v], &Nc[3 * v]);
    for (size_t f = 0; f < nf; ++f)
        for (int k = 0; k < 3; ++k) {
            int v = lm.F[3 * f + k];
            for (int j = 0; j < 3; ++j) r.corner_normals[9 * f + 3 * k + j] = Nc[3 * (size_t)v + j];
        }
    return r;
}

}  // namespace aur
