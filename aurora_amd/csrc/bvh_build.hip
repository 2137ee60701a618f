// bvh_build.hip -- GPU-built wide BVH (replaces Embree's bvh4.spatialsplit
// builder, src/geometry/embreeAccelerator.cpp:19-35).
//
// Pipeline, all on the device:
//   1. per-triangle padded bounds and centroid bounds (k_prim_bounds)
//   2. 63-bit Morton codes of the centroids (k_morton)
//   3. radix sort of (code, triangle) pairs (hipCUB)
//   4. binary tree by parallel locally-ordered clustering over the sorted
//      triangles (PLOC, k_ploc_*; Meister & Bittner 2018): the Morton order
//      only seeds the clusters, merges pick the smallest union surface area
//      among 16 neighbours on either side, inner boxes come from the merges
//   5. leaves renumbered in the tree's left-to-right order (k_ploc_place,
//      k_ploc_permute), so every subtree covers a contiguous triangle range
//   6. SAH-optimal collapse into W-wide nodes: a bottom-up dynamic program
//      over the binary tree (k_collapse_cost) decides which subtrees become
//      leaves of up to max_leaf triangles (contiguous ranges of the sorted
//      triangles) and how each wide node's W slots are filled; the nodes are
//      then written top-down, level by level (k_collapse); a node's inner
//      children take its first slots and consecutive node indices from pad[0]
//   7. triangle records in sorted order (k_reorder): (v0 | prim), e1, e2 so a
//      leaf is read with contiguous 16-byte loads
//
// Triangle boxes are padded by 1e-5 of the largest coordinate so traversal
// never culls a triangle the Moller-Trumbore test would accept.
#include <hipcub/hipcub.hpp>

#include "context.h"

namespace aur {

static constexpr unsigned LEAF_BIT = 0x80000000u;

__device__ inline int f2o(float f) { int i = __float_as_int(f); return i >= 0 ? i : i ^ 0x7fffffff; }
__device__ inline float o2f(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7fffffff); }

__global__ void k_init_bounds(int *cb) {
    if (threadIdx.x < 3) { cb[threadIdx.x] = f2o(3.0e38f); cb[3 + threadIdx.x] = f2o(-3.0e38f); }
}

__global__ void k_prim_bounds(const float *__restrict__ verts, const int *__restrict__ tris, int n,
                              float4 *__restrict__ lo, float4 *__restrict__ hi, int *__restrict__ cb) {
    __shared__ float s_min[3][256], s_max[3][256];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    float cmin[3] = {3.0e38f, 3.0e38f, 3.0e38f}, cmax[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
    if (i < n) {
        float l[3] = {3.0e38f, 3.0e38f, 3.0e38f}, h[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
        float mag = 0.f;
        for (int c = 0; c < 3; ++c) {
            const float *v = verts + 3 * tris[3 * i + c];
            for (int k = 0; k < 3; ++k) {
                l[k] = fminf(l[k], v[k]);
                h[k] = fmaxf(h[k], v[k]);
                mag = fmaxf(mag, fabsf(v[k]));
            }
        }
        float pad = mag * 1e-5f + 1e-30f;
        for (int k = 0; k < 3; ++k) {
            l[k] -= pad; h[k] += pad;
            float c = 0.5f * (l[k] + h[k]);
            cmin[k] = c; cmax[k] = c;
        }
        lo[i] = make_float4(l[0], l[1], l[2], 0.f);
        hi[i] = make_float4(h[0], h[1], h[2], 0.f);
    }
    for (int k = 0; k < 3; ++k) { s_min[k][threadIdx.x] = cmin[k]; s_max[k][threadIdx.x] = cmax[k]; }
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int k = 0; k < 3; ++k) {
                s_min[k][threadIdx.x] = fminf(s_min[k][threadIdx.x], s_min[k][threadIdx.x + s]);
                s_max[k][threadIdx.x] = fmaxf(s_max[k][threadIdx.x], s_max[k][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) { atomicMin(&cb[k], f2o(s_min[k][0])); atomicMax(&cb[3 + k], f2o(s_max[k][0])); }
}

__device__ inline unsigned long long spread21(unsigned x) {
    unsigned long long v = x & 0x1fffffu;
    v = (v | (v << 32)) & 0x1f00000000ffffULL;
    v = (v | (v << 16)) & 0x1f0000ff0000ffULL;
    v = (v | (v << 8)) & 0x100f00f00f00f00fULL;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ULL;
    v = (v | (v << 2)) & 0x1249249249249249ULL;
    return v;
}

__global__ void k_morton(const float4 *__restrict__ lo, const float4 *__restrict__ hi, int n, const int *__restrict__ cb,
                         unsigned long long *__restrict__ keys, int *__restrict__ vals) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 l = lo[i], h = hi[i];
    float c[3] = {0.5f * (l.x + h.x), 0.5f * (l.y + h.y), 0.5f * (l.z + h.z)};
    unsigned q[3];
    for (int k = 0; k < 3; ++k) {
        float mn = o2f(cb[k]), mx = o2f(cb[3 + k]);
        float ext = mx - mn;
        float t = ext > 0.f ? (c[k] - mn) / ext : 0.f;
        t = fminf(fmaxf(t, 0.f), 1.f);
        q[k] = (unsigned)(t * 2097151.f);
    }
    keys[i] = (spread21(q[0]) << 2) | (spread21(q[1]) << 1) | spread21(q[2]);
    vals[i] = i;
}

__device__ inline void child_box(int code, const float4 *plo, const float4 *phi, const int *vals, const float4 *nlo,
                                 const float4 *nhi, float4 *l, float4 *h) {
    if (code & LEAF_BIT) {
        int p = vals[code & ~LEAF_BIT];
        *l = plo[p]; *h = phi[p];
    } else {
        *l = nlo[code]; *h = nhi[code];
    }
}

// Collapse decisions of the binary tree (k_collapse_cost): per inner node,
// 4-bit fields of dec[x]: field 0 = 1 when the subtree is best kept as one
// leaf; field i - 1 (i = 2..W) = how many of i slots the left child takes
// when x's children are spread over i slots of a wide node (0: i - 1 slots
// are as good); field W = that split for x's own wide node (all W slots).
struct CollapseIn {
    const int *left, *right, *first, *last, *vals;
    const float4 *plo, *phi, *nlo, *nhi;
    const float *S;                  // per inner node, W costs: subtree in at most 1..W slots
    const unsigned long long *dec;   // per inner node, the decisions above
    int max_leaf;
};

// SAH weights of the collapse: a wide node visited, a leaf visited (its
// triangle records fetched) and a triangle tested, per unit of surface area
// relative to the parent.
static constexpr float SAH_NODE = 1.0f;
static constexpr float SAH_LEAF = 1.0f;
static constexpr float SAH_TRI = 0.25f;

__device__ inline float half_area(float4 l, float4 h) {
    float dx = h.x - l.x, dy = h.y - l.y, dz = h.z - l.z;
    return dx * dy + dy * dz + dz * dx;
}

__device__ inline int code_count(const CollapseIn &in, int code) {
    if (code & LEAF_BIT) return 1;
    return in.last[code] - in.first[code] + 1;
}
__device__ inline int code_first(const CollapseIn &in, int code) {
    if (code & LEAF_BIT) return code & ~LEAF_BIT;
    return in.first[code];
}

// A child code kept whole in one slot: a leaf when it is one triangle or the
// collapse chose a leaf for it, else a wide node of its own.
__device__ inline bool dp_leaf(const CollapseIn &in, int code) {
    return (code & LEAF_BIT) || (in.dec[code] & 1ull);
}

template <int W>
__device__ inline void slot_costs(const CollapseIn &in, int c, float (&s)[W]) {
    if (c & LEAF_BIT) {
        float4 l, h;
        child_box(c, in.plo, in.phi, in.vals, in.nlo, in.nhi, &l, &h);
        const float v = half_area(l, h) * (SAH_LEAF + SAH_TRI);
#pragma unroll
        for (int i = 0; i < W; ++i) s[i] = v;
    } else {
#pragma unroll
        for (int i = 0; i < W; ++i) s[i] = in.S[(size_t)c * W + i];
    }
}

// SAH-optimal collapse of the binary tree into W-wide nodes (the dynamic
// program of Ylitie, Karras and Laine 2017, "Efficient Incoherent Ray
// Traversal on GPUs Through Compressed Wide BVHs", with leaves of up to
// max_leaf triangles), bottom-up: the nodes one PLOC iteration made
// [base, base + k) only have children made by earlier iterations.
//   S(x, 1) = min(leaf cost, A(x) * SAH_NODE + D(x, W))
//   S(x, i) = min(S(x, i - 1), D(x, i)),  D(x, j) = min_q S(l, q) + S(r, j - q)
template <int W>
__global__ void k_collapse_cost(int base, int k, CollapseIn in, float *__restrict__ S,
                                unsigned long long *__restrict__ dec) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= k) return;
    const int x = base + u;
    float sa[W], sb[W];
    slot_costs<W>(in, in.left[x], sa);
    slot_costs<W>(in, in.right[x], sb);
    const float inf = __int_as_float(0x7f800000);
    float D[W + 1];
    int kd[W + 1];
#pragma unroll
    for (int j = 2; j <= W; ++j) {
        float best = inf;
        int bk = 1;
#pragma unroll
        for (int q = 1; q < j; ++q) {
            const float v = sa[q - 1] + sb[j - q - 1];
            if (v < best) { best = v; bk = q; }
        }
        D[j] = best;
        kd[j] = bk;
    }
    const float A = half_area(in.nlo[x], in.nhi[x]);
    const int cnt = in.last[x] - in.first[x] + 1;
    const float c_node = A * SAH_NODE + D[W];
    const float c_leaf = cnt <= in.max_leaf ? A * (SAH_LEAF + SAH_TRI * (float)cnt) : inf;
    unsigned long long d = 0;
    float s = fminf(c_leaf, c_node);
    if (c_leaf <= c_node) d |= 1ull;
    S[(size_t)x * W] = s;
#pragma unroll
    for (int i = 2; i <= W; ++i) {
        if (D[i] < s) {
            s = D[i];
            d |= (unsigned long long)kd[i] << (4 * (i - 1));
        }
        S[(size_t)x * W + i - 1] = s;
    }
    d |= (unsigned long long)kd[W] << (4 * W);
    dec[x] = d;
}

template <int W>
__global__ void k_collapse(CollapseIn in, const int2 *__restrict__ tasks, const int *__restrict__ n_tasks,
                           int2 *__restrict__ next, int *__restrict__ n_next, int *__restrict__ node_count,
                           WNode<W> *__restrict__ nodes, int *__restrict__ max_depth, int depth) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *n_tasks) return;
    int2 task = tasks[i];  // (binary code, wide index)
    int ch[W];
    int nc = 0;
    int code = task.x;
    if (!dp_leaf(in, code)) {
        // the W slots of x's wide node, as the collapse decided: a (subtree,
        // slots) stack, left subtrees first
        int sy[W], si[W], sp = 0;
        const int kx = (int)((in.dec[code] >> (4 * W)) & 15ull);
        sy[sp] = in.right[code]; si[sp++] = W - kx;
        sy[sp] = in.left[code]; si[sp++] = kx;
        while (sp > 0) {
            --sp;
            const int y = sy[sp], i = si[sp];
            if ((y & LEAF_BIT) || i == 1) { ch[nc++] = y; continue; }
            const int q = (int)((in.dec[y] >> (4 * (i - 1))) & 15ull);
            if (q == 0) {
                sy[sp] = y; si[sp++] = i - 1;
            } else {
                sy[sp] = in.right[y]; si[sp++] = i - q;
                sy[sp] = in.left[y]; si[sp++] = q;
            }
        }
    } else {
        // only the root: a tree of one leaf
        ch[0] = code;
        nc = 1;
    }
    // inner children first, in one block of consecutive node indices
    // (pad[0] = the block's first index), so traversal names an inner child
    // by its slot: child = pad[0] + slot
    int n_inner = 0;
    {
        int ord[W];
        for (int k = 0; k < nc; ++k)
            if (!dp_leaf(in, ch[k])) ord[n_inner++] = ch[k];
        int j = n_inner;
        for (int k = 0; k < nc; ++k)
            if (dp_leaf(in, ch[k])) ord[j++] = ch[k];
        for (int k = 0; k < nc; ++k) ch[k] = ord[k];
    }
    const int base = n_inner ? atomicAdd(node_count, n_inner) : 0;
    const int q0 = n_inner ? atomicAdd(n_next, n_inner) : 0;
    WNode<W> nd;
    for (int k = 0; k < W; ++k) {
        if (k >= nc) {
            nd.lox[k] = nd.hix[k] = nd.loy[k] = nd.hiy[k] = nd.loz[k] = nd.hiz[k] = __int_as_float(0x7f800000);
            nd.child[k] = 0x7fffffff;
            nd.pad[k] = 0;
            continue;
        }
        int c = ch[k];
        float4 l, h;
        child_box(c, in.plo, in.phi, in.vals, in.nlo, in.nhi, &l, &h);
        nd.lox[k] = l.x; nd.loy[k] = l.y; nd.loz[k] = l.z;
        nd.hix[k] = h.x; nd.hiy[k] = h.y; nd.hiz[k] = h.z;
        nd.pad[k] = 0;
        int cnt = code_count(in, c);
        if (k < n_inner) {
            next[q0 + k] = make_int2(c, base + k);
            nd.child[k] = base + k;
        } else {
            nd.child[k] = ~((code_first(in, c) << 4) | (cnt - 1));
        }
    }
    nd.pad[0] = base;
    nodes[task.y] = nd;
    atomicMax(max_depth, depth + 1);
}

// ---------------------------------------------------------------------------
// PLOC: parallel locally-ordered clustering (Meister & Bittner 2018,
// "Parallel Locally-Ordered Clustering for Bounding Volume Hierarchy
// Construction").  Clusters start as the Morton-sorted triangles; every
// iteration each cluster finds, among the R clusters on either side of it in
// the current order, the one whose union box has the smallest surface area,
// mutual nearest neighbours merge into a new inner node (which takes the
// lower position), and the survivors are compacted in order.  Pair ties are
// broken on (area, lower index, higher index), so the globally smallest pair
// is always mutual and every iteration merges.  Node and cluster positions
// come from one 64-bit exclusive scan of (merges << 32 | survivors): the tree
// is the same on every run.
// ---------------------------------------------------------------------------
static constexpr int PLOC_R = 16;   // neighbours searched on either side
static constexpr int PLOC_B = 256;

__global__ void k_ploc_init(int n, const int *__restrict__ vals, const float4 *__restrict__ plo,
                            const float4 *__restrict__ phi, int *__restrict__ cid, float4 *__restrict__ clo,
                            float4 *__restrict__ chi) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    int p = vals[s];
    cid[s] = (int)(s | LEAF_BIT);
    clo[s] = plo[p];
    chi[s] = phi[p];
}

template <int R>
__global__ void __launch_bounds__(PLOC_B) k_ploc_nn(int m, const float4 *__restrict__ clo, const float4 *__restrict__ chi,
                                                    int *__restrict__ nn) {
    __shared__ float4 slo[PLOC_B + 2 * R], shi[PLOC_B + 2 * R];
    const int base = blockIdx.x * PLOC_B;
    for (int t = threadIdx.x; t < PLOC_B + 2 * R; t += PLOC_B) {
        const int g = base - R + t;
        if (g >= 0 && g < m) { slo[t] = clo[g]; shi[t] = chi[g]; }
    }
    __syncthreads();
    const int i = base + threadIdx.x;
    if (i >= m) return;
    const float4 l = slo[threadIdx.x + R], h = shi[threadIdx.x + R];
    float best = __int_as_float(0x7f800000);
    int bj = -1;
    // candidates in increasing index order: with a strict '<' the first of
    // equal areas wins, which is the (area, lower, higher) order of pairs
    for (int k = -R; k <= R; ++k) {
        const int j = i + k;
        if (k == 0 || j < 0 || j >= m) continue;
        const float4 l2 = slo[threadIdx.x + R + k], h2 = shi[threadIdx.x + R + k];
        const float dx = fmaxf(h.x, h2.x) - fminf(l.x, l2.x);
        const float dy = fmaxf(h.y, h2.y) - fminf(l.y, l2.y);
        const float dz = fmaxf(h.z, h2.z) - fminf(l.z, l2.z);
        const float a = dx * dy + dy * dz + dz * dx;
        if (a < best) { best = a; bj = j; }
    }
    nn[i] = bj;
}

__global__ void k_ploc_flags(int m, const int *__restrict__ nn, unsigned long long *__restrict__ flags) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int j = nn[i];
    const bool mutual = j >= 0 && nn[j] == i;
    const unsigned long long merge = (mutual && i < j) ? 1ull : 0ull;
    const unsigned long long keep = (mutual && i > j) ? 0ull : 1ull;
    flags[i] = (merge << 32) | keep;
}

struct PlocTree {
    int *left, *right, *count;
    float4 *nlo, *nhi;
};

__global__ void k_ploc_merge(int m, int node_base, const int *__restrict__ nn, const unsigned long long *__restrict__ flags,
                             const unsigned long long *__restrict__ scan, const int *__restrict__ cid,
                             const float4 *__restrict__ clo, const float4 *__restrict__ chi, int *__restrict__ cid2,
                             float4 *__restrict__ clo2, float4 *__restrict__ chi2, PlocTree t, int *__restrict__ totals) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const unsigned long long sc = scan[i];
    if (i == m - 1) {
        const unsigned long long tot = sc + flags[i];
        totals[0] = (int)(tot & 0xffffffffu);  // survivors
        totals[1] = (int)(tot >> 32);          // merges
    }
    const unsigned long long f = flags[i];
    if (!(f & 1ull)) return;  // merged into a lower cluster
    const int out = (int)(sc & 0xffffffffu);
    if (f >> 32) {
        const int j = nn[i];
        const int node = node_base + (int)(sc >> 32);
        const int a = cid[i], b = cid[j];
        const float4 l1 = clo[i], h1 = chi[i], l2 = clo[j], h2 = chi[j];
        const float4 l = make_float4(fminf(l1.x, l2.x), fminf(l1.y, l2.y), fminf(l1.z, l2.z), 0.f);
        const float4 h = make_float4(fmaxf(h1.x, h2.x), fmaxf(h1.y, h2.y), fmaxf(h1.z, h2.z), 0.f);
        t.left[node] = a;
        t.right[node] = b;
        t.count[node] = ((a & LEAF_BIT) ? 1 : t.count[a]) + ((b & LEAF_BIT) ? 1 : t.count[b]);
        t.nlo[node] = l;
        t.nhi[node] = h;
        cid2[out] = node;
        clo2[out] = l;
        chi2[out] = h;
    } else {
        cid2[out] = cid[i];
        clo2[out] = clo[i];
        chi2[out] = chi[i];
    }
}

__global__ void k_ploc_root(int root, int *__restrict__ first) { first[root] = 0; }

// Leaf order = the tree's left-to-right order, so every subtree covers a
// contiguous range: nodes made in one iteration place their children (made
// in earlier iterations), iterations from the last to the first.
__global__ void k_ploc_place(int base, int k, PlocTree t, int *__restrict__ first, int *__restrict__ last,
                             int *__restrict__ pos) {
    int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= k) return;
    const int node = base + u;
    const int f = first[node];
    const int l = t.left[node], r = t.right[node];
    const int cl = (l & LEAF_BIT) ? 1 : t.count[l];
    if (l & LEAF_BIT) { pos[l & ~LEAF_BIT] = f; t.left[node] = (int)(f | LEAF_BIT); } else first[l] = f;
    if (r & LEAF_BIT) { pos[r & ~LEAF_BIT] = f + cl; t.right[node] = (int)((f + cl) | LEAF_BIT); } else first[r] = f + cl;
    last[node] = f + t.count[node] - 1;
}

__global__ void k_ploc_permute(int n, const int *__restrict__ pos, const int *__restrict__ vals, int *__restrict__ out) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    out[pos[s]] = vals[s];
}

__global__ void k_reorder(const float *__restrict__ verts, const int *__restrict__ tris, const int *__restrict__ vals, int n,
                          float4 *__restrict__ out) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    int p = vals[s];
    const float *a = verts + 3 * tris[3 * p], *b = verts + 3 * tris[3 * p + 1], *c = verts + 3 * tris[3 * p + 2];
    out[3 * s] = make_float4(a[0], a[1], a[2], __int_as_float(p));
    out[3 * s + 1] = make_float4(b[0] - a[0], b[1] - a[1], b[2] - a[2], 0.f);
    out[3 * s + 2] = make_float4(c[0] - a[0], c[1] - a[1], c[2] - a[2], 0.f);
}

__global__ void k_seed_task(int2 *tasks, int *n_tasks, int code) {
    tasks[0] = make_int2(code, 0);
    *n_tasks = 1;
}

__global__ void k_fill_u32(unsigned *p, size_t n, unsigned v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

void fill_zero(Ctx *c, void *p, size_t bytes) {
    size_t words = (bytes + 3) / 4;
    if (!words) return;
    unsigned blocks = (unsigned)std::min<size_t>((words + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, c->stream, (unsigned *)p, words, 0u);
    HIP_CHECK(hipGetLastError());
}

template <int W>
static void collapse_levels(Ctx *c, const CollapseIn &in, int root_code, int max_nodes) {
    Bvh &b = c->bvh;
    b.nodes.reserve(sizeof(WNode<W>) * (size_t)std::max(max_nodes, 1));
    DevBuf tq[2], scalars;
    tq[0].reserve(sizeof(int2) * (size_t)std::max(max_nodes, 1));
    tq[1].reserve(sizeof(int2) * (size_t)std::max(max_nodes, 1));
    scalars.reserve(sizeof(int) * 8);
    int *sc = scalars.as<int>();  // [0] n_tasks A, [1] n_tasks B, [2] node_count, [3] max_depth
    fill_zero(c, sc, sizeof(int) * 8);
    hipLaunchKernelGGL(k_seed_task, dim3(1), dim3(1), 0, c->stream, tq[0].as<int2>(), sc + 0, root_code);
    int one = 1;
    HIP_CHECK(hipMemcpyAsync(sc + 2, &one, sizeof(int), hipMemcpyHostToDevice, c->stream));
    int cur = 0, n_tasks = 1, depth = 0;
    while (n_tasks > 0) {
        fill_zero(c, sc + (1 - cur), sizeof(int));
        hipLaunchKernelGGL(k_collapse<W>, dim3((n_tasks + 127) / 128), dim3(128), 0, c->stream, in, tq[cur].as<int2>(),
                           sc + cur, tq[1 - cur].as<int2>(), sc + (1 - cur), sc + 2, b.nodes.as<WNode<W>>(), sc + 3, depth);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpyAsync(&n_tasks, sc + (1 - cur), sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
        cur = 1 - cur;
        ++depth;
    }
    int hs[4];
    HIP_CHECK(hipMemcpyAsync(hs, sc, sizeof hs, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    b.n_nodes = hs[2];
    b.depth = hs[3];
}

void build_bvh(Ctx *c, int width, int max_leaf) {
    if (width != 4 && width != 8) throw std::runtime_error("BVH width must be 4 or 8");
    if (max_leaf < 1 || max_leaf > 16) throw std::runtime_error("max_leaf must be in 1..16");
    const int n = c->n_tris;
    if (n >= (1 << 27)) throw CapacityError("too many triangles for the leaf encoding (max 2^27)");
    Bvh &b = c->bvh;
    b.width = width;
    b.max_leaf = max_leaf;
    b.tri_sorted.reserve(sizeof(float4) * 3 * (size_t)std::max(n, 1));
    if (n == 0) {
        // a single node with no children: every ray misses
        size_t nb = width == 4 ? sizeof(WNode<4>) : sizeof(WNode<8>);
        b.nodes.reserve(nb);
        std::vector<unsigned> host(nb / 4, 0x7f800000u);
        HIP_CHECK(hipMemcpyAsync(b.nodes.p, host.data(), nb, hipMemcpyHostToDevice, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
        b.n_nodes = 1;
        b.depth = 1;
        return;
    }
    const float *verts = c->d_verts.as<float>();
    const int *tris = c->d_tris.as<int>();
    DevBuf plo, phi, cb, keys[2], vals[2], temp;
    plo.reserve(sizeof(float4) * n);
    phi.reserve(sizeof(float4) * n);
    cb.reserve(sizeof(int) * 8);
    keys[0].reserve(sizeof(unsigned long long) * n);
    keys[1].reserve(sizeof(unsigned long long) * n);
    vals[0].reserve(sizeof(int) * n);
    vals[1].reserve(sizeof(int) * n);
    const int bs = 256, grid = (n + bs - 1) / bs;
    hipLaunchKernelGGL(k_init_bounds, dim3(1), dim3(64), 0, c->stream, cb.as<int>());
    hipLaunchKernelGGL(k_prim_bounds, dim3(grid), dim3(bs), 0, c->stream, verts, tris, n, plo.as<float4>(), phi.as<float4>(), cb.as<int>());
    hipLaunchKernelGGL(k_morton, dim3(grid), dim3(bs), 0, c->stream, plo.as<float4>(), phi.as<float4>(), n, cb.as<int>(),
                       keys[0].as<unsigned long long>(), vals[0].as<int>());
    HIP_CHECK(hipGetLastError());
    size_t temp_bytes = 0;
    HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, keys[0].as<unsigned long long>(),
                                                 keys[1].as<unsigned long long>(), vals[0].as<int>(), vals[1].as<int>(), n,
                                                 0, 63, c->stream));
    temp.reserve(temp_bytes);
    HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(temp.p, temp_bytes, keys[0].as<unsigned long long>(),
                                                 keys[1].as<unsigned long long>(), vals[0].as<int>(), vals[1].as<int>(), n,
                                                 0, 63, c->stream));
    const int *svals = vals[1].as<int>();

    int ni = std::max(n - 1, 1);
    DevBuf left, right, first, last, nlo, nhi;
    left.reserve(sizeof(int) * ni);
    right.reserve(sizeof(int) * ni);
    first.reserve(sizeof(int) * ni);
    last.reserve(sizeof(int) * ni);
    nlo.reserve(sizeof(float4) * ni);
    nhi.reserve(sizeof(float4) * ni);
    int root_code;
    DevBuf ploc_vals;
    std::vector<std::pair<int, int>> made;  // (first node, nodes) per PLOC iteration
    if (n >= 2) {
        // clusters (double-buffered), nearest neighbours, flags and their scan
        DevBuf cid[2], clo[2], chi[2], nn, flags, scan, count, pos, totals, stemp;
        for (int k = 0; k < 2; ++k) {
            cid[k].reserve(sizeof(int) * n);
            clo[k].reserve(sizeof(float4) * n);
            chi[k].reserve(sizeof(float4) * n);
        }
        nn.reserve(sizeof(int) * n);
        flags.reserve(sizeof(unsigned long long) * n);
        scan.reserve(sizeof(unsigned long long) * n);
        count.reserve(sizeof(int) * ni);
        pos.reserve(sizeof(int) * n);
        totals.reserve(sizeof(int) * 2);
        ploc_vals.reserve(sizeof(int) * n);
        size_t st_bytes = 0;
        HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, st_bytes, flags.as<unsigned long long>(),
                                                   scan.as<unsigned long long>(), n, c->stream));
        stemp.reserve(st_bytes);
        PlocTree t;
        t.left = left.as<int>(); t.right = right.as<int>(); t.count = count.as<int>();
        t.nlo = nlo.as<float4>(); t.nhi = nhi.as<float4>();
        hipLaunchKernelGGL(k_ploc_init, dim3(grid), dim3(bs), 0, c->stream, n, svals, plo.as<float4>(), phi.as<float4>(),
                           cid[0].as<int>(), clo[0].as<float4>(), chi[0].as<float4>());
        int m = n, node_base = 0, cur = 0;
        while (m > 1) {
            const int g = (m + PLOC_B - 1) / PLOC_B;
            hipLaunchKernelGGL(k_ploc_nn<PLOC_R>, dim3(g), dim3(PLOC_B), 0, c->stream, m, clo[cur].as<float4>(),
                               chi[cur].as<float4>(), nn.as<int>());
            hipLaunchKernelGGL(k_ploc_flags, dim3(g), dim3(PLOC_B), 0, c->stream, m, nn.as<int>(),
                               flags.as<unsigned long long>());
            HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(stemp.p, st_bytes, flags.as<unsigned long long>(),
                                                       scan.as<unsigned long long>(), m, c->stream));
            hipLaunchKernelGGL(k_ploc_merge, dim3(g), dim3(PLOC_B), 0, c->stream, m, node_base, nn.as<int>(),
                               flags.as<unsigned long long>(), scan.as<unsigned long long>(), cid[cur].as<int>(),
                               clo[cur].as<float4>(), chi[cur].as<float4>(), cid[1 - cur].as<int>(),
                               clo[1 - cur].as<float4>(), chi[1 - cur].as<float4>(), t, totals.as<int>());
            HIP_CHECK(hipGetLastError());
            int tot[2];
            HIP_CHECK(hipMemcpyAsync(tot, totals.p, sizeof tot, hipMemcpyDeviceToHost, c->stream));
            HIP_CHECK(hipStreamSynchronize(c->stream));
            if (tot[1] <= 0 || tot[0] != m - tot[1]) throw std::runtime_error("PLOC made no progress");
            made.emplace_back(node_base, tot[1]);
            node_base += tot[1];
            m = tot[0];
            cur = 1 - cur;
        }
        if (node_base != n - 1) throw std::runtime_error("PLOC built a wrong node count");
        root_code = n - 2;
        hipLaunchKernelGGL(k_ploc_root, dim3(1), dim3(1), 0, c->stream, root_code, first.as<int>());
        for (auto it = made.rbegin(); it != made.rend(); ++it)
            hipLaunchKernelGGL(k_ploc_place, dim3((it->second + bs - 1) / bs), dim3(bs), 0, c->stream, it->first,
                               it->second, t, first.as<int>(), last.as<int>(), pos.as<int>());
        hipLaunchKernelGGL(k_ploc_permute, dim3(grid), dim3(bs), 0, c->stream, n, pos.as<int>(), svals,
                           ploc_vals.as<int>());
        HIP_CHECK(hipGetLastError());
        svals = ploc_vals.as<int>();
    } else {
        root_code = (int)(0u | LEAF_BIT);
    }
    CollapseIn in;
    in.left = left.as<int>(); in.right = right.as<int>(); in.first = first.as<int>(); in.last = last.as<int>();
    in.vals = svals; in.plo = plo.as<float4>(); in.phi = phi.as<float4>(); in.nlo = nlo.as<float4>(); in.nhi = nhi.as<float4>();
    in.max_leaf = max_leaf;
    in.S = nullptr;
    in.dec = nullptr;
    DevBuf cost, dec;
    if (n >= 2) {
        cost.reserve(sizeof(float) * (size_t)width * ni);
        dec.reserve(sizeof(unsigned long long) * (size_t)ni);
        in.S = cost.as<float>();
        in.dec = dec.as<unsigned long long>();
        for (const auto &it : made) {
            const dim3 g((it.second + bs - 1) / bs);
            if (width == 4)
                hipLaunchKernelGGL(k_collapse_cost<4>, g, dim3(bs), 0, c->stream, it.first, it.second, in,
                                   cost.as<float>(), dec.as<unsigned long long>());
            else
                hipLaunchKernelGGL(k_collapse_cost<8>, g, dim3(bs), 0, c->stream, it.first, it.second, in,
                                   cost.as<float>(), dec.as<unsigned long long>());
        }
        HIP_CHECK(hipGetLastError());
    }
    if (width == 4) collapse_levels<4>(c, in, root_code, ni + 1);
    else collapse_levels<8>(c, in, root_code, ni + 1);
    hipLaunchKernelGGL(k_reorder, dim3(grid), dim3(bs), 0, c->stream, verts, tris, svals, n, b.tri_sorted.as<float4>());
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(c->stream));
}

}  // namespace aur
