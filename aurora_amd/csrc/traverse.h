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

#ifndef AUR_STACK_LDS
#define AUR_STACK_LDS 24
#endif

struct TravHit { float t, u, v; int prim; };

template <int W>
struct NodeRegs {
    float lox[W], hix[W], loy[W], hiy[W], loz[W], hiz[W];
    int child[W];
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
}

// Stack: top STACK_LDS entries per lane in LDS, deeper entries in global memory.
struct TravStack {
    int *lds;       // base for this lane: lds[k * stride]
    int stride;
    int *ovf;       // this thread's overflow area
    int sp;
    __device__ inline void push(int v) {
        if (sp < AUR_STACK_LDS) lds[sp * stride] = v;
        else ovf[sp - AUR_STACK_LDS] = v;
        ++sp;
    }
    __device__ inline int pop() {
        --sp;
        return sp < AUR_STACK_LDS ? lds[sp * stride] : ovf[sp - AUR_STACK_LDS];
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

template <int W, bool ANY>
__device__ inline void traverse(const WNode<W> *__restrict__ nodes, const float4 *__restrict__ tris, f3 o, f3 d,
                                float tmin, float tmax, TravStack &st, TravHit &h) {
    h.t = tmax; h.u = 0.f; h.v = 0.f; h.prim = -1;
    // reciprocal direction for the slab test (zero components nudged so no 0*inf appears)
    float dx = fabsf(d.x) < 1e-20f ? copysignf(1e-20f, d.x) : d.x;
    float dy = fabsf(d.y) < 1e-20f ? copysignf(1e-20f, d.y) : d.y;
    float dz = fabsf(d.z) < 1e-20f ? copysignf(1e-20f, d.z) : d.z;
    const float ix = 1.f / dx, iy = 1.f / dy, iz = 1.f / dz;
    const float ox = o.x * ix, oy = o.y * iy, oz = o.z * iz;
    st.sp = 0;
    int node = 0;
    const float inf = __int_as_float(0x7f800000);
    for (;;) {
        NodeRegs<W> r;
        load_node<W>(nodes, node, r);
        // exact-t ties must still be visited, so the far bound is widened slightly
        const float tfar_lim = h.t * 1.0000004f;
        // entry distance per inner child (inf: missed, or a leaf: leaves are
        // collected in a mask and tested after the box tests); every index
        // below is a compile-time constant, so the lists stay in VGPRs
        float tn[W];
        int ch[W];
        unsigned leaves = 0;
#pragma unroll
        for (int k = 0; k < W; ++k) {
            float x0 = fmaf(r.lox[k], ix, -ox), x1 = fmaf(r.hix[k], ix, -ox);
            float y0 = fmaf(r.loy[k], iy, -oy), y1 = fmaf(r.hiy[k], iy, -oy);
            float z0 = fmaf(r.loz[k], iz, -oz), z1 = fmaf(r.hiz[k], iz, -oz);
            float a = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), tmin));
            float b = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tfar_lim));
            int c = r.child[k];
            bool hit = a <= b;
            tn[k] = (hit && c >= 0) ? a : inf;
            ch[k] = c;
            if (hit && c < 0) leaves |= 1u << k;
        }
        // the leaves this node's boxes let through, one per pass of the loop:
        // lanes with leaves in different slots test them together
        while (leaves) {
            const int k = __builtin_ctz(leaves);
            leaves &= leaves - 1;
            int c = ch[0];
#pragma unroll
            for (int j = 1; j < W; ++j) c = (k == j) ? ch[j] : c;
            if (test_leaf<ANY>(tris, c, o, d, tmin, h)) return;
        }
        if (!ANY) {
            // sorting network on (entry distance, child): nearest first
#pragma unroll
            for (int i = 0; i < W; ++i) {
#pragma unroll
                for (int j = (i & 1); j + 1 < W; j += 2) {
                    bool sw = tn[j] > tn[j + 1];
                    float ta = sw ? tn[j + 1] : tn[j], tb = sw ? tn[j] : tn[j + 1];
                    int ca = sw ? ch[j + 1] : ch[j], cb = sw ? ch[j] : ch[j + 1];
                    tn[j] = ta; tn[j + 1] = tb; ch[j] = ca; ch[j + 1] = cb;
                }
            }
        }
        // children entered beyond the (possibly shortened) hit distance are skipped
        const float tcut = h.t * 1.0000004f;
        int next = -1;
#pragma unroll
        for (int k = W - 1; k >= 0; --k) {
            if (tn[k] <= tcut) {
                if (next >= 0) st.push(next);
                next = ch[k];
            }
        }
        if (next >= 0) { node = next; continue; }
        if (st.sp == 0) break;
        node = st.pop();
    }
}

}  // namespace aur
