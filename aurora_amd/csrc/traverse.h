// traverse.h -- wide-BVH ray traversal (replaces Embree's intersect/occluded,
// src/geometry/embreeAccelerator.cpp:41-100).
//
// One ray per lane.  The traversal stack lives in LDS (STACK_LDS entries per
// lane, lane-minor so a wave's pushes hit distinct banks) and spills to a
// per-thread global area beyond that depth.  Closest hit keeps the smallest t
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
#define AUR_REFILL_LANES 8

struct TravHit { float t, u, v; int prim; };

template <int W>
struct NodeRegs {
    float lox[W], hix[W], loy[W], hiy[W], loz[W], hiz[W];
    int child[W];
    int base;  // index of the first inner child: inner slot k is node base + k
};

template <int W>
__device__ inline void load_node(const WNode<W> *__restrict__ nodes, int ni, NodeRegs<W> &r) {
    const float4 *p = reinterpret_cast<const float4 *>(nodes + ni);
    constexpr int Q = W / 4;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        float4 a = p[0 * Q + q], b = p[1 * Q + q], c = p[2 * Q + q], d = p[3 * Q + q], e = p[4 * Q + q], f = p[5 * Q + q];
        int4 g = reinterpret_cast<const int4 *>(p)[6 * Q + q];
        r.lox[4 * q + 0] = a.x; r.lox[4 * q + 1] = a.y; r.lox[4 * q + 2] = a.z; r.lox[4 * q + 3] = a.w;
        r.hix[4 * q + 0] = b.x; r.hix[4 * q + 1] = b.y; r.hix[4 * q + 2] = b.z; r.hix[4 * q + 3] = b.w;
        r.loy[4 * q + 0] = c.x; r.loy[4 * q + 1] = c.y; r.loy[4 * q + 2] = c.z; r.loy[4 * q + 3] = c.w;
        r.hiy[4 * q + 0] = d.x; r.hiy[4 * q + 1] = d.y; r.hiy[4 * q + 2] = d.z; r.hiy[4 * q + 3] = d.w;
        r.loz[4 * q + 0] = e.x; r.loz[4 * q + 1] = e.y; r.loz[4 * q + 2] = e.z; r.loz[4 * q + 3] = e.w;
        r.hiz[4 * q + 0] = f.x; r.hiz[4 * q + 1] = f.y; r.hiz[4 * q + 2] = f.z; r.hiz[4 * q + 3] = f.w;
        r.child[4 * q + 0] = g.x; r.child[4 * q + 1] = g.y; r.child[4 * q + 2] = g.z; r.child[4 * q + 3] = g.w;
    }
    r.base = nodes[ni].pad[0];
}

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

// Tests the triangles of one leaf; returns true when an any-hit query is done.
template <bool ANY>
__device__ inline bool test_leaf(const float4 *__restrict__ tris, int code, f3 o, f3 d, float tmin, TravHit &h) {
    int enc = ~code;
    int first = enc >> 4, cnt = (enc & 15) + 1;
    for (int i = first; i < first + cnt; ++i) {
        float4 a = tris[3 * i], b = tris[3 * i + 1], c = tris[3 * i + 2];
        float t, b1, b2;
        if (intersect_tri(o, d, mk3(a.x, a.y, a.z), mk3(b.x, b.y, b.z), mk3(c.x, c.y, c.z), tmin, h.t, &t, &b1, &b2)) {
            int prim = __float_as_int(a.w);
            if (ANY) { h.t = t; h.u = b1; h.v = b2; h.prim = prim; return true; }
            if (h.prim < 0 || t < h.t || prim < h.prim) { h.t = t; h.u = b1; h.v = b2; h.prim = prim; }
        }
    }
    return false;
}

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

// Visits one node: tests its child boxes, tests the leaves they let through,
// pushes the inner children to visit later and moves `node` to the next one.
// Returns true when the ray is done (stack empty, or an any-hit query hit).
template <int W, bool ANY>
__device__ inline bool trav_visit(const NodeRegs<W> &nr, const float4 *__restrict__ tris, const TravRay &r,
                                  TravStack &st, int &node, TravHit &h) {
    const float inf = __int_as_float(0x7f800000);
    // exact-t ties must still be visited, so the far bound is widened slightly
    const float tfar_lim = h.t * 1.0000004f;
    // entry distance per inner child (inf: missed, or a leaf: leaves are
    // collected in a mask and tested after the box tests); every index
    // below is a compile-time constant, so the lists stay in VGPRs
    // Closest hit orders the inner children by 32-bit keys: the entry
    // distance's bits (a non-negative float orders as an unsigned integer)
    // with the slot in the low bits, so the sort moves one register per
    // child and the child is base + slot.  Missed children get ~0u.
    constexpr unsigned SLOT = W == 4 ? 3u : 7u;
    float tn[W];
    unsigned key[W];
    int ch[W];
    unsigned leaves = 0;
#pragma unroll
    for (int k = 0; k < W; ++k) {
        float x0 = fmaf(nr.lox[k], r.ix, -r.ox), x1 = fmaf(nr.hix[k], r.ix, -r.ox);
        float y0 = fmaf(nr.loy[k], r.iy, -r.oy), y1 = fmaf(nr.hiy[k], r.iy, -r.oy);
        float z0 = fmaf(nr.loz[k], r.iz, -r.oz), z1 = fmaf(nr.hiz[k], r.iz, -r.oz);
        float a = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), r.tmin));
        float b = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tfar_lim));
        int c = nr.child[k];
        bool hit = a <= b;
        if (ANY) tn[k] = (hit && c >= 0) ? a : inf;
        // a >= tmin >= 0: clearing the sign bit only turns -0 into +0
        else key[k] = (hit && c >= 0) ? ((__float_as_uint(a) & 0x7fffffffu & ~SLOT) | (unsigned)k) : ~0u;
        ch[k] = c;
        if (hit && c < 0) leaves |= 1u << k;
    }
    // the leaves this node's boxes let through, one per pass of the loop:
    // lanes with leaves in different slots test them together
#ifdef AUR_TRAV_STATS
    st.n_visits += 1;
#endif
    while (leaves) {
        const int k = __builtin_ctz(leaves);
        leaves &= leaves - 1;
        int c = ch[0];
#pragma unroll
        for (int j = 1; j < W; ++j) c = (k == j) ? ch[j] : c;
#ifdef AUR_TRAV_STATS
        st.n_leaves += 1;
        st.n_tris += (unsigned)((~c) & 15) + 1u;
#endif
        if (test_leaf<ANY>(tris, c, r.o, r.d, r.tmin, h)) return true;
    }
    if (ANY) {
        // any hit: children in slot order
        const float tcut = h.t * 1.0000004f;
        int next = -1;
#pragma unroll
        for (int k = W - 1; k >= 0; --k) {
            if (tn[k] <= tcut) {
                if (next >= 0) st.push(next);
                next = ch[k];
            }
        }
        if (next >= 0) { node = next; return false; }
    } else {
        sort_keys<W>(key);
        // children entered beyond the (possibly shortened) hit distance are
        // skipped; the slot bits and the cleared low distance bits only ever
        // admit more children (ties and a few ulps), never fewer
        const unsigned kcut = __float_as_uint(fmaxf(h.t * 1.0000004f, 0.f)) | SLOT;
        if (key[0] <= kcut) {
#pragma unroll
            for (int k = W - 1; k >= 1; --k)
                if (key[k] <= kcut) st.push(nr.base + (int)(key[k] & SLOT));
            node = nr.base + (int)(key[0] & SLOT);
            return false;
        }
    }
    if (st.sp == 0) return true;
    node = st.pop();
    return false;
}

template <int W, bool ANY>
__device__ inline bool trav_step(const WNode<W> *__restrict__ nodes, const float4 *__restrict__ tris,
                                 const TravRay &r, TravStack &st, int &node, TravHit &h) {
    NodeRegs<W> nr;
    load_node<W>(nodes, node, nr);
    return trav_visit<W, ANY>(nr, tris, r, st, node, h);
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
