// traverse.h -- wide-BVH ray traversal (replaces Embree's intersect/occluded,
// src/geometry/embreeAccelerator.cpp:41-100).
//
// One ray per lane.  A lane's current item is an inner node or a leaf, and
// one step of the loop fetches and processes one item (trav_step).  The
// traversal stack lives in LDS (STACK_LDS entries per lane, lane-minor so a
// wave's pushes hit distinct banks) and spills to a per-thread global area
// beyond that depth.  Closest hit keeps the smallest t
// and, on an exact tie, the smallest triangle id -- the same rule as the CPU
// brute-force oracle, so hit ids compare bit-exactly.
#pragma once

#include "context.h"
#include "dev_common.h"

namespace aur {

// 20 entries x 4 B x 64 lanes = 5 KB of LDS per wave: 32 waves fit in a CU's
// 160 KB, so LDS does not cap the traversal kernels' occupancy
#ifndef AUR_STACK_LDS
#define AUR_STACK_LDS 20
#endif

// Minimum waves per SIMD the traversal kernels are compiled for: 8 (64 VGPRs)
// for 4-wide nodes; 8-wide nodes need more registers than that without
// spilling, so they are left to the compiler
#define AUR_TRAV_WAVES(W) ((W) == 4 ? 8 : 1)

// A lane whose ray is done waits for this many lanes of its wave before they
// fetch new rays together (traverse_stream)
#ifndef AUR_REFILL_LANES
#define AUR_REFILL_LANES 32
#endif

struct TravHit { float t, u, v; int prim; };

template <int W>
struct NodeRegs {
    float lox[W], hix[W], loy[W], hiy[W], loz[W], hiz[W];
    int child[W];  // inner child: node index; leaf: ~(first << 4 | count - 1)
};

// Stack: top STACK_LDS entries per lane in LDS, deeper entries in global
// memory.  Only the lane's LDS address and depth are per-lane values; the
// overflow address is derived from them when needed, so the stack holds two
// VGPRs across the traversal loop.
// LDS-qualified pointers: a pop from LDS stays a ds_read instead of being
// merged with the overflow path's global load into one flat load (which
// waits on both counters)
typedef __attribute__((address_space(3))) int lds_int;

struct TravStack {
#ifdef AUR_TRAV_STATS
    // measurement builds only (-DAUR_TRAV_STATS): nodes visited, leaves and
    // triangles tested by this lane
    unsigned n_visits = 0, n_leaves = 0, n_tris = 0;
#endif
    lds_int *lds_base;  // the workgroup's stack array
    lds_int *lds;       // this lane's entry 0: entry k is lds[k * stride]
    int stride;     // lanes per workgroup
    int *ovf;       // the workgroup's overflow area: lane t owns ovf[t * ovf_per_thread ...]
    int ovf_per_thread;
    int sp;
    __device__ inline void init(int *s_stack, int lanes, int *ovf_all, int per_thread) {
        lds_base = (lds_int *)s_stack;
        lds = lds_base + threadIdx.x;
        stride = lanes;
        ovf = ovf_all + (size_t)blockIdx.x * lanes * per_thread;
        ovf_per_thread = per_thread;
    }
    __device__ inline int *ovf_entry(int k) const {
        int lane = (int)(lds - lds_base);
        // an empty asm makes `lane` opaque here, so the 64-bit overflow
        // address is not hoisted out of the traversal loop into two more
        // registers: it is rebuilt on this rarely taken path instead
        asm volatile("" : "+v"(lane));
        return ovf + (size_t)lane * ovf_per_thread + (k - AUR_STACK_LDS);
    }
    __device__ inline void push(int v) {
        if (sp < AUR_STACK_LDS) lds[sp * stride] = v;
        else *ovf_entry(sp) = v;
        ++sp;
    }
    __device__ inline int pop() {
        --sp;
        return sp < AUR_STACK_LDS ? lds[sp * stride] : *ovf_entry(sp);
    }
};

// Per-ray constants of the slab test and the triangle test.
struct TravRay {
    f3 o, d;
    float ix, iy, iz;   // reciprocal direction (zero components nudged so no 0*inf appears)
    float ox, oy, oz;   // origin times reciprocal direction
    float tmin;
};

__device__ inline void trav_setup(TravRay &r, f3 o, f3 d, float tmin, float tmax, TravHit &h) {
    r.o = o; r.d = d; r.tmin = tmin;
    float dx = fabsf(d.x) < 1e-20f ? copysignf(1e-20f, d.x) : d.x;
    float dy = fabsf(d.y) < 1e-20f ? copysignf(1e-20f, d.y) : d.y;
    float dz = fabsf(d.z) < 1e-20f ? copysignf(1e-20f, d.z) : d.z;
    r.ix = 1.f / dx; r.iy = 1.f / dy; r.iz = 1.f / dz;
    r.ox = o.x * r.ix; r.oy = o.y * r.iy; r.oz = o.z * r.iz;
    h.t = tmax; h.u = 0.f; h.v = 0.f; h.prim = -1;
}

__device__ inline void trav_begin(TravRay &r, f3 o, f3 d, float tmin, float tmax, TravStack &st, int &node,
                                  TravHit &h, int root = 0) {
    trav_setup(r, o, d, tmin, tmax, h);
    st.sp = 0;
    node = root;
}

// Ascending sort of W keys by a min/max network (Batcher's odd-even merge
// sort: 5 comparators for 4 keys, 19 for 8).
template <int W>
__device__ inline void sort_keys(unsigned (&k)[W]) {
    auto ce = [&](int i, int j) {
        const unsigned lo = min(k[i], k[j]), hi = max(k[i], k[j]);
        k[i] = lo;
        k[j] = hi;
    };
    if constexpr (W == 4) {
        ce(0, 1); ce(2, 3); ce(0, 2); ce(1, 3); ce(1, 2);
    } else {
        ce(0, 1); ce(2, 3); ce(4, 5); ce(6, 7);
        ce(0, 2); ce(1, 3); ce(4, 6); ce(5, 7);
        ce(1, 2); ce(5, 6);
        ce(0, 4); ce(1, 5); ce(2, 6); ce(3, 7);
        ce(2, 4); ce(3, 5);
        ce(1, 2); ce(3, 4); ce(5, 6);
    }
}

// The child code in slot s of a node (a select tree over the W codes, so
// nothing is dynamically indexed).
template <int W>
__device__ inline int child_of(const int (&ch)[W], unsigned s) {
    if constexpr (W == 4) {
        const int lo = (s & 1u) ? ch[1] : ch[0], hi = (s & 1u) ? ch[3] : ch[2];
        return (s & 2u) ? hi : lo;
    } else {
        const int a = (s & 1u) ? ch[1] : ch[0], b = (s & 1u) ? ch[3] : ch[2];
        const int c = (s & 1u) ? ch[5] : ch[4], d = (s & 1u) ? ch[7] : ch[6];
        const int ab = (s & 2u) ? b : a, cd = (s & 2u) ? d : c;
        return (s & 4u) ? cd : ab;
    }
}

// Visits an inner node: tests its child boxes and moves `node` to the next
// item, pushing the other children that the ray enters.  Leaves are items
// like inner nodes: they are tested in a later step of their own.  Returns
// true when the ray is done (nothing entered and the stack empty).
template <int W, bool ANY>
__device__ inline bool visit_inner(const NodeRegs<W> &nr, const TravRay &r, TravStack &st, int &node,
                                   const TravHit &h) {
    // exact-t ties must still be visited, so the far bound is widened slightly
    const float tfar_lim = h.t * 1.0000004f;
    // The entered children are visited in the order of 32-bit keys: the
    // entry distance's bits (a non-negative float orders as an unsigned
    // integer) with the slot in the low bits, so the sort moves one register
    // per child.  Closest hit orders inner children and leaves alike front to
    // back; an any-hit query takes the entered leaves first (the query ends
    // at its first accepted triangle), then the inner children front to
    // back.  Missed children get ~0u.
    constexpr unsigned SLOT = W == 4 ? 3u : 7u;
    unsigned key[W];
#pragma unroll
    for (int k = 0; k < W; ++k) {
        float x0 = fmaf(nr.lox[k], r.ix, -r.ox), x1 = fmaf(nr.hix[k], r.ix, -r.ox);
        float y0 = fmaf(nr.loy[k], r.iy, -r.oy), y1 = fmaf(nr.hiy[k], r.iy, -r.oy);
        float z0 = fmaf(nr.loz[k], r.iz, -r.oz), z1 = fmaf(nr.hiz[k], r.iz, -r.oz);
        float a = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), r.tmin));
        float b = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tfar_lim));
        // a >= tmin >= 0: clearing the sign bit only turns -0 into +0; an
        // any-hit query takes its entered leaves first (key = slot)
        const unsigned dist = (ANY && nr.child[k] < 0) ? 0u : (__float_as_uint(a) & 0x7fffffffu & ~SLOT);
        key[k] = a <= b ? (dist | (unsigned)k) : ~0u;
    }
#ifdef AUR_TRAV_STATS
    st.n_visits += 1;
#endif
    sort_keys<W>(key);
    if (key[0] != ~0u) {
#pragma unroll
        for (int k = W - 1; k >= 1; --k)
            if (key[k] != ~0u) st.push(child_of<W>(nr.child, key[k] & SLOT));
        node = child_of<W>(nr.child, key[0] & SLOT);
        return false;
    }
    if (st.sp == 0) return true;
    node = st.pop();
    return false;
}

// Takes one triangle test's result into the hit record; returns true when an
// any-hit query is done.  Closest hit keeps the smallest t and, on an exact
// tie, the smallest triangle id, so the result does not depend on the order
// in which triangles are tested.
template <bool ANY>
__device__ inline bool take_tri(const float4 &a, const float4 &b, const float4 &c, const TravRay &r, TravHit &h) {
    float t, b1, b2;
    if (!intersect_tri(r.o, r.d, mk3(a.x, a.y, a.z), mk3(b.x, b.y, b.z), mk3(c.x, c.y, c.z), r.tmin, h.t, &t, &b1, &b2))
        return false;
    const int prim = __float_as_int(a.w);
    if (ANY) { h.t = t; h.u = b1; h.v = b2; h.prim = prim; return true; }
    if (h.prim < 0 || t < h.t || prim < h.prim) { h.t = t; h.u = b1; h.v = b2; h.prim = prim; }
    return false;
}

// One step of a ray's traversal: the lane's current item is an inner node
// (node >= 0) or a leaf (~(first << 4 | count - 1)).  Either way the step
// fetches six float4 rows into the same registers -- a node's first six
// box rows, or the records of the leaf's next two triangles (48 B each) --
// so every step costs one round trip to memory whichever kind of item each
// lane holds, and a wave's lanes at nodes and at leaves progress together.
// A leaf of more than two triangles stays the lane's item, shortened by two,
// for the next step.  Returns true when the ray is done.
template <int W, bool ANY>
__device__ inline bool trav_step(const WNode<W> *__restrict__ nodes, const float4 *__restrict__ tris,
                                 const TravRay &r, TravStack &st, int &node, TravHit &h) {
    constexpr int Q = W / 4;
    const bool inner = node >= 0;
    const int enc = ~node;
    const int first = enc >> 4, cnt = (enc & 15) + 1;
    const float4 *pn = reinterpret_cast<const float4 *>(nodes + (inner ? node : 0));
    const float4 *p0 = inner ? pn : tris + 3 * first;
    const float4 *p1 = inner ? pn + 3 : tris + 3 * (cnt >= 2 ? first + 1 : first);
    // rows 0-5 from p0/p1; the rest of a node (box rows beyond the sixth of
    // 8-wide nodes, the child codes) from pn, which a leaf lane points at the
    // root so that its loads stay in bounds
    float4 rows[6 * Q];
    int4 crow[Q];
    rows[0] = p0[0]; rows[1] = p0[1]; rows[2] = p0[2];
    rows[3] = p1[0]; rows[4] = p1[1]; rows[5] = p1[2];
#pragma unroll
    for (int j = 6; j < 6 * Q; ++j) rows[j] = pn[j];
#pragma unroll
    for (int q = 0; q < Q; ++q) crow[q] = reinterpret_cast<const int4 *>(pn)[6 * Q + q];
    // every fetched word is used here, before the branch on the item's kind:
    // the loads issue together as full 16-byte loads (the compiler would
    // otherwise narrow rows to the words the leaf test reads and sink the
    // child codes' load below the box tests, a second round trip per node)
#pragma unroll
    for (int j = 0; j < 6 * Q; ++j) asm volatile("" ::"v"(rows[j].x), "v"(rows[j].y), "v"(rows[j].z), "v"(rows[j].w));
#pragma unroll
    for (int q = 0; q < Q; ++q) asm volatile("" ::"v"(crow[q].x), "v"(crow[q].y), "v"(crow[q].z), "v"(crow[q].w));
    if (inner) {
        NodeRegs<W> nr;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const float4 a = rows[0 * Q + q], b = rows[1 * Q + q], c = rows[2 * Q + q];
            const float4 d = rows[3 * Q + q], e = rows[4 * Q + q], f = rows[5 * Q + q];
            const int4 g = crow[q];
            nr.lox[4 * q + 0] = a.x; nr.lox[4 * q + 1] = a.y; nr.lox[4 * q + 2] = a.z; nr.lox[4 * q + 3] = a.w;
            nr.hix[4 * q + 0] = b.x; nr.hix[4 * q + 1] = b.y; nr.hix[4 * q + 2] = b.z; nr.hix[4 * q + 3] = b.w;
            nr.loy[4 * q + 0] = c.x; nr.loy[4 * q + 1] = c.y; nr.loy[4 * q + 2] = c.z; nr.loy[4 * q + 3] = c.w;
            nr.hiy[4 * q + 0] = d.x; nr.hiy[4 * q + 1] = d.y; nr.hiy[4 * q + 2] = d.z; nr.hiy[4 * q + 3] = d.w;
            nr.loz[4 * q + 0] = e.x; nr.loz[4 * q + 1] = e.y; nr.loz[4 * q + 2] = e.z; nr.loz[4 * q + 3] = e.w;
            nr.hiz[4 * q + 0] = f.x; nr.hiz[4 * q + 1] = f.y; nr.hiz[4 * q + 2] = f.z; nr.hiz[4 * q + 3] = f.w;
            nr.child[4 * q + 0] = g.x; nr.child[4 * q + 1] = g.y; nr.child[4 * q + 2] = g.z; nr.child[4 * q + 3] = g.w;
        }
        return visit_inner<W, ANY>(nr, r, st, node, h);
    }
    const float4 *row = rows;
#ifdef AUR_TRAV_STATS
    st.n_leaves += 1;
    st.n_tris += cnt >= 2 ? 2u : 1u;
#endif
    if (take_tri<ANY>(row[0], row[1], row[2], r, h)) return true;
    if (cnt >= 2 && take_tri<ANY>(row[3], row[4], row[5], r, h)) return true;
    if (cnt > 2) {
        node = ~(((first + 2) << 4) | (cnt - 3));
        return false;
    }
    if (st.sp == 0) return true;
    node = st.pop();
    return false;
}

// One whole ray, from node `root` (a batched launch's forest of BVHs).
template <int W, bool ANY>
__device__ inline void traverse(const WNode<W> *__restrict__ nodes, const float4 *__restrict__ tris, f3 o, f3 d,
                                float tmin, float tmax, TravStack &st, TravHit &h, int root = 0) {
    TravRay r;
    int node;
    trav_begin(r, o, d, tmin, tmax, st, node, h, root);
    while (!trav_step<W, ANY>(nodes, tris, r, st, node, h)) {
    }
}

// Persistent traversal of rays first, first + stride, ... below n with lanes
// refilled independently: a lane whose ray is done waits (masked) until at
// least `refill` lanes of its wave are waiting or no lane is busy, then the
// waiting lanes fetch their next rays together.  The wave keeps stepping
// while any lane has a ray, so one long ray no longer holds 63 idle lanes,
// and the fetch cost (a ray load and three divisions) is paid once per group
// of lanes instead of once per lane.  `load(i, o, d, tmin, tmax)` reads ray i;
// `done(i, h)` consumes its result.  With ROOTED, `load(i, o, d, tmin, tmax,
// root)` also names the ray's root node (a batched launch's forest of
// BVHs).  The per-ray results are those of `traverse`: each ray's visit
// order does not depend on its neighbours.
template <int W, bool ANY, bool ROOTED = false, class Load, class Done>
__device__ inline void traverse_stream(const WNode<W> *__restrict__ nodes, const float4 *__restrict__ tris,
                                       TravStack &st, int first, int n, int stride, int refill, Load load,
                                       Done done) {
    TravRay r;
    TravHit h;
    trav_setup(r, mk3(0.f, 0.f, 0.f), mk3(1.f, 1.f, 1.f), 0.f, 0.f, h);
    st.sp = 0;
    int node = 0;
    int i = first;
    bool busy = false;
    for (;;) {
        const unsigned long long want = __ballot(!busy && i < n);
        const unsigned long long working = __ballot(busy);
        if (!want && !working) break;
        if (want && (!working || __popcll(want) >= (unsigned)refill)) {
            // wave-uniform branch; the new rays enter the lane state through
            // selects, so the state never merges at a divergent join
            const bool take = !busy && i < n;
            f3 o = mk3(0.f, 0.f, 0.f), d = mk3(1.f, 1.f, 1.f);
            float tmin = 0.f, tmax = 0.f;
            int root = 0;
            if (take) {
                if constexpr (ROOTED) load(i, o, d, tmin, tmax, root);
                else load(i, o, d, tmin, tmax);
            }
            TravRay nr;
            TravHit nh;
            trav_setup(nr, o, d, tmin, tmax, nh);
            r.o.x = take ? nr.o.x : r.o.x; r.o.y = take ? nr.o.y : r.o.y; r.o.z = take ? nr.o.z : r.o.z;
            r.d.x = take ? nr.d.x : r.d.x; r.d.y = take ? nr.d.y : r.d.y; r.d.z = take ? nr.d.z : r.d.z;
            r.ix = take ? nr.ix : r.ix; r.iy = take ? nr.iy : r.iy; r.iz = take ? nr.iz : r.iz;
            r.ox = take ? nr.ox : r.ox; r.oy = take ? nr.oy : r.oy; r.oz = take ? nr.oz : r.oz;
            r.tmin = take ? nr.tmin : r.tmin;
            h.t = take ? nh.t : h.t; h.u = take ? nh.u : h.u; h.v = take ? nh.v : h.v;
            h.prim = take ? nh.prim : h.prim;
            node = take ? root : node;
            st.sp = take ? 0 : st.sp;
            busy = busy || take;
        }
        if (busy && trav_step<W, ANY>(nodes, tris, r, st, node, h)) {
            done(i, h);
            busy = false;
            i += stride;
        }
    }
}

}  // namespace aur
