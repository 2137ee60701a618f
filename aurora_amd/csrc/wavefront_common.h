// wavefront_common.h -- pieces of the wavefront tracer shared by the
// single-scene launch (wavefront.hip) and the batched persistent launch
// (batch.hip): queue layout, wave64 compaction helpers, the camera ray and
// the per-vertex shading step.
#pragma once

#include "context.h"
#include "shade_common.h"
#include "traverse.h"

namespace aur {

// Path-state streams (rays, hits, queue state, shadow records, outputs) are
// written once and read once, a bounce later, gigabytes per bounce: they move
// with non-temporal loads and stores, a hint that they need not stay in the
// caches the BVH nodes and triangles share (the 1080p frame 1852 -> 1867
// Mpaths/s, the shade kernel 0.4 ms faster).
typedef float aur_v4f __attribute__((ext_vector_type(4)));
typedef int aur_v4i __attribute__((ext_vector_type(4)));
__device__ inline float4 ld_s(const float4 *p) {
    const aur_v4f v = __builtin_nontemporal_load(reinterpret_cast<const aur_v4f *>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ inline int4 ld_s(const int4 *p) {
    const aur_v4i v = __builtin_nontemporal_load(reinterpret_cast<const aur_v4i *>(p));
    return make_int4(v.x, v.y, v.z, v.w);
}
__device__ inline void st_s(float4 *p, const float4 v) {
    aur_v4f w = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(w, reinterpret_cast<aur_v4f *>(p));
}
__device__ inline void st_s(int4 *p, const int4 v) {
    aur_v4i w = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(w, reinterpret_cast<aur_v4i *>(p));
}

// Instantiates a traversal kernel launch for the BVH width and workgroup size
// chosen at run time; CALL(w, b) expands to the launch.
#define AUR_TRAV_DISPATCH(width, block, CALL)                               \
    do {                                                                    \
        const int w__ = (width), b__ = (block);                             \
        if (w__ == 8) {                                                     \
            if (b__ == 64) CALL(8, 64);                                     \
            else if (b__ == 256) CALL(8, 256);                              \
            else CALL(8, 128);                                              \
        } else {                                                            \
            if (b__ == 64) CALL(4, 64);                                     \
            else if (b__ == 256) CALL(4, 256);                              \
            else CALL(4, 128);                                              \
        }                                                                   \
    } while (0)

int trace_blocks(const Ctx *c);
size_t trace_threads(const Ctx *c);

enum { RT_CAMERA = 0, RT_DIFFUSE = 1, RT_SPECULAR = 2, RT_MIRROR = 3 };
// Path flags: bits 0-1 ray type, 2 matte path, 3 last ray missed, 8-15
// vertex index, 16-23 bounces, 24 no firefly clamp.  The shadow rays a path
// has traced are counted in Lb.w (a float, exact to 2^24), which the batched
// launch sums per job when the path finishes.
enum { F_MATTE = 1 << 2, F_MISS = 1 << 3, F_NOCLAMP = 1 << 24 };
// rays one finished path traced, from its flags: extension rays (camera ray
// included) and those that hit
__device__ inline int path_ext_rays(int flags) { return ((flags >> 8) & 0xff) + 1; }
__device__ inline int path_hit_rays(int flags) { return path_ext_rays(flags) - ((flags & F_MISS) ? 1 : 0); }
enum { C_Q0 = 0, C_Q1 = 1, C_SHADOW = 2, C_FINISH = 3 };
enum { S_PATHS = 0, S_EXT = 1, S_HIT = 2, S_SHADOW = 3 };

// Path state lives with the queue: entry i of extension queue q holds the
// ray, the throughput T (rgb, continue probability), the two radiance
// accumulators La (emission and light-sample NEE, alpha in w) and Lb
// (BSDF-sample NEE; shadow rays traced in w), and a meta record (flags, id0,
// id1, id2).  Shading copies
// a continuing path's state to its slot in the next queue and a finished
// path's to the finish list, so every access is coalesced; shadow rays carry
// the position their contribution goes to.
//   single-scene launch: id0 = path index (pixel and sample follow from it)
//   batched launch:      id0 = local pixel index in the pass, id1 = pass,
//                        id2 = job
struct WfView {
    float4 *ray[2];
    float4 *T[2], *La[2], *Lb[2];
    int4 *meta[2];
    float4 *hit;
    float4 *sray;       // 3 per shadow ray: (o, job bits), (d, tmax), (contribution, target)
    float4 *finLa, *finLb;
    int4 *finMeta;
    float4 *out;
    int *counts;
    unsigned long long *stats;
    unsigned long long *partials;  // per traversal workgroup: hit rays (summed after the run)
    int *ovf;                      // traversal stack overflow: extension rays (and the batched kernels)
    int *sovf;                     // and shadow rays, which may run beside them
    int ovf_per_thread;
    PathSource src;
    UDiv div_per;       // src.ls * width * height (passes after the first)
    uint32_t seed;
    int nslots;
};

// shadow-ray target: position in the next queue, or in the finish list
// (bit 30), accumulator La or Lb (bit 31)
#define SH_TO_FINISH 0x40000000
#define SH_TO_LB ((int)0x80000000)
#define SH_POS_MASK 0x3fffffff

__device__ inline void shadow_add(const WfView &wv, int nq, int tag, float4 c) {
    const int pos = tag & SH_POS_MASK;
    float4 *acc;
    if (tag & SH_TO_FINISH) acc = (tag & SH_TO_LB) ? wv.finLb : wv.finLa;
    else acc = (tag & SH_TO_LB) ? wv.Lb[nq] : wv.La[nq];
    float4 v = acc[pos];
    v.x += c.x; v.y += c.y; v.z += c.z;
    acc[pos] = v;
}

WfView wf_view(Ctx *c, const PathSource &src);
WfView wf_view_on(Ctx *c, Wavefront &w, const PathSource &src);

// ---------------------------------------------------------------------------
// wave-level helpers (wave64)
// ---------------------------------------------------------------------------
__device__ inline int wave_reserve(int *counter, bool pred) {
    unsigned long long mask = __ballot(pred);
    int lane = __lane_id();
    int leader = __ffsll((long long)mask) - 1;
    int base = 0;
    if (mask && lane == leader) base = atomicAdd(counter, __popcll(mask));
    base = __shfl(base, leader < 0 ? 0 : leader);
    return base + __popcll(mask & ((1ull << lane) - 1ull));
}

__device__ inline void wave_count(unsigned long long *counter, bool pred) {
    unsigned long long mask = __ballot(pred);
    int lane = __lane_id();
    int leader = __ffsll((long long)mask) - 1;
    if (mask && lane == leader) atomicAdd(counter, (unsigned long long)__popcll(mask));
}

// Counts predicate lanes per job: the wave's first job takes one atomic for
// all its lanes (the usual case: a wave holds consecutive paths of one job),
// lanes of other jobs add one each.
__device__ inline void wave_count_job(unsigned long long *cnt, int stride, int job, bool pred) {
    unsigned long long mask = __ballot(pred);
    if (!mask) return;
    int lane = __lane_id();
    int leader = __ffsll((long long)mask) - 1;
    int j0 = __shfl(job, leader);
    unsigned long long same = __ballot(pred && job == j0);
    if (lane == leader) atomicAdd(&cnt[(size_t)j0 * stride], (unsigned long long)__popcll(same));
    if (pred && job != j0) atomicAdd(&cnt[(size_t)job * stride], 1ull);
}

// camera.cpp:34-62: pixel (x, y) jittered by two uniforms, pinhole at the
// origin looking down -z; the ray spans [0.1, 1e7].
__device__ inline f3 camera_dir(const SceneView &sv, int pixel, float r0, float r1) {
    const int y = udiv(pixel, sv.div_w), x = pixel - y * sv.width;
    float fovx = sv.fov;
    float fovy = ((float)sv.height / (float)sv.width) * fovx;
    float u = x + r0, v = y + r1;
    f3 d = mk3(((2.f * u - sv.width) / (float)sv.width) * fovx, ((-2.f * v + sv.height) / (float)sv.height) * fovy, -1.f);
    return norm3(d);
}

// The camera ray of a path as k_generate stores it: origin 0, tmin 0.1 and
// tmax 1e7 (camera.cpp:58-59), direction through the path's pixel.
template <class Rng>
__device__ inline void camera_ray_of(const SceneView &sv, const Rng &rng, float4 *a, float4 *b) {
    const f3 d = camera_dir(sv, rng.pixel, rng.u(0), rng.u(1));
    *a = make_float4(0.f, 0.f, 0.f, 0.1f);
    *b = make_float4(d.x, d.y, d.z, 10000000.f);
}

// integrator.cpp:277-281: the firefly clamp, skipped for camera rays that saw
// the environment directly.
__device__ inline f3 firefly_clamp(f3 L, int flags) {
    if (!(flags & F_NOCLAMP))
        for (int k = 0; k < AUR_FIREFLY_MAX_STEPS && (L.x > AUR_FIREFLY || L.y > AUR_FIREFLY || L.z > AUR_FIREFLY); ++k)
            L = mul3(L, 0.2f);
    return L;
}

// ---------------------------------------------------------------------------
// shade: one path vertex (integrator.cpp:180-274) for queue entry i of queue
// q.  shade_compute fills a ShadeOut; shade_emit then reserves places in the
// next queue, the finish list and the shadow list once per workgroup (a
// wave ballot per list, the waves' totals combined in LDS, one atomic per
// list) and writes them.  make_rng(meta) gives the path's random-number
// source; shadow rays record `job` for the batched launch.
// ---------------------------------------------------------------------------
struct ShadowOut { bool valid; f3 o, d; float tmin, tmax; f3 c; };

struct ShadeOut {
    bool cont, finish;
    ShadowOut sh[2];
    f3 P, ndir;
    float ntmax;
    float4 T4, La, Lb;
    int4 meta;
    int job;
};

// What a path whose extension ray missed sees (flags: its flags, updated).
// get_rb() gives the ray's direction record and get_T() its throughput; each
// is called only on the branch that needs it (a camera miss under an
// environment map, a mirror miss under an environment light).
template <bool TEX, class MakeRng, class GetRb, class GetT>
__device__ inline void miss_radiance(const SceneView &sv, const int4 &meta, MakeRng make_rng, GetRb get_rb, GetT get_T,
                                     float4 &La, int &flags) {
    const int rtype = flags & 3;
    if (rtype == RT_CAMERA) {  // integrator.cpp:287-297: camera miss sees the environment
        f3 e = mk3(sv.env[0], sv.env[1], sv.env[2]);
        if constexpr (TEX) {
            if (sv.has_env) {
                const float4 rb = get_rb();
                e = env_miss(sv, mk3(rb.x, rb.y, rb.z));
            } else {
                e = mk3(0.f, 0.f, 0.f);
            }
        }
        La = make_float4(e.x, e.y, e.z, sv.has_env ? 1.f : 0.f);
        flags |= F_NOCLAMP;
    } else if (rtype == RT_MIRROR && sv.has_env) {
        // integrator.cpp:262-271: a mirror ray that misses adds every env
        // light's eval() in the direction of the path's camera ray
        const float4 T4 = get_T();
        f3 cd = mk3(0.f, 0.f, 0.f);
        if constexpr (TEX) {
            const auto rng = make_rng(meta);
            cd = camera_dir(sv, rng.pixel, rng.u(0), rng.u(1));
        }
        for (int li = 0; li < sv.n_lights; ++li) {
            const DLight &l = sv.lights[li];
            if (l.type != LIGHT_ENV) continue;
            f3 e = mk3(l.Le[0], l.Le[1], l.Le[2]);
            if constexpr (TEX) e = env_eval(sv, l, cd);
            La.x += e.x * T4.x; La.y += e.y * T4.y; La.z += e.z * T4.z;
        }
    }
}

// TEX selects the code for shaders, normal maps, environment maps and the
// Halton sample source; the host uses it only for scenes that need any.
// CAM: the camera vertex of a hash-sampled render, whose ray and state are
// not stored (path i, its camera ray recomputed as camera_ray_of builds it
// for k_extend, unit throughput, no radiance yet).
template <bool TEX, bool CAM = false, class MakeRng>
__device__ inline void shade_compute(const SceneView &sv, const WfView &wv, int q, int i, MakeRng make_rng, int job,
                                     ShadeOut &out) {
    int4 meta;
    float4 ra, rb, La, Lb, T4;
    if constexpr (CAM) {
        meta = make_int4(RT_CAMERA, i, 0, 0);
        camera_ray_of(sv, make_rng(meta), &ra, &rb);
        La = Lb = make_float4(0.f, 0.f, 0.f, 0.f);
        T4 = make_float4(1.f, 1.f, 1.f, 1.f);
    } else {
        meta = ld_s(&wv.meta[q][i]);
        ra = ld_s(&wv.ray[q][2 * i]);
        rb = ld_s(&wv.ray[q][2 * i + 1]);
        La = ld_s(&wv.La[q][i]);
        Lb = ld_s(&wv.Lb[q][i]);
        T4 = ld_s(&wv.T[q][i]);
    }
    const float4 hv = ld_s(&wv.hit[i]);
    const int prim = __float_as_int(hv.w);
    int flags = meta.x;
    const int rtype = flags & 3;
    const int tb = (flags >> 8) & 0xff;
    bool cont = false, finish = false;
    ShadowOut sh[2];
    sh[0].valid = sh[1].valid = false;
    f3 P = mk3(0.f, 0.f, 0.f), ndir = mk3(0.f, 0.f, 0.f);
    float ntmax = 0.f;
    if (prim < 0) {
        finish = true;
        miss_radiance<TEX>(sv, meta, make_rng, [&] { return rb; }, [&] { return T4; }, La, flags);
    } else if (tb >= sv.maxdepth) {
        finish = true;  // the extension ray after the last vertex is traced but not shaded
    } else {
        const auto rng = make_rng(meta);
        const int vslot = 2 + AUR_SLOTS_PER_VERTEX * tb;
        // hit record: embreeAccelerator.cpp:111-136
        const f3 o = mk3(ra.x, ra.y, ra.z), d = mk3(rb.x, rb.y, rb.z);
        const float t = hv.x, u = hv.y, v = hv.z;
        const float w = (float)(1.0 - (double)u - (double)v);  // `float w = 1. - u - v` (double)
        P = add3(o, mul3(d, t));
        const float4 s0 = sv.shade[3 * prim], s1 = sv.shade[3 * prim + 1], s2 = sv.shade[3 * prim + 2];
        const int attr = __float_as_int(s0.w);
        f3 nm = mk3(w * s0.x + u * s1.x + v * s2.x, w * s0.y + u * s1.y + v * s2.y, w * s0.z + u * s1.z + v * s2.z);
        if (dot3(nm, d) > 0.f) nm = mul3(nm, -1.f);
        f3 nn = norm3(nm);
        const DMat m = sv.mats[attr];
        ShadeGeo g;
        g.P = P; g.Ng = nn; g.attr = attr; g.s = 0.f; g.t = 0.f;
        if (TEX && m.texmask) {
            if (sv.uv) {
                const float2 t0 = sv.uv[3 * prim], t1 = sv.uv[3 * prim + 1], t2 = sv.uv[3 * prim + 2];
                g.s = w * t0.x + u * t1.x + v * t2.x;
                g.t = w * t0.y + u * t1.y + v * t2.y;
            }
            if (m.texmask & (1 << SLOT_NORMAL)) nn = normal_map(sv, m, g);
        }
        // integrator.cpp:132-150
        const f3 vn = neg3(norm3(d));
        if (dot3(vn, nn) <= 0.f) nn = norm3(cross3(vn, cross3(nn, vn)));
        bool matte = (flags & F_MATTE) != 0;
        if (rtype == RT_DIFFUSE) matte = true;
        f3 T = mk3(T4.x, T4.y, T4.z);
        float cp = T4.w;
        if (rtype == RT_CAMERA || rtype == RT_MIRROR) {  // integrator.cpp:190-193
            float4 &la = La;
            la.x += T.x * m.emission[0]; la.y += T.y * m.emission[1]; la.z += T.z * m.emission[2];
            la.w = 1.f;
        }
        if (m.type == MAT_LIGHT) {
            finish = true;  // ConstantMaterial: getSample has pdf 0 (constant.cpp:23-25)
        } else {
            Bsdf b = get_brdf<TEX>(sv, m, prim, vn, nn, matte, rng.u(vslot + 0), g);
            const bool delta = b.kind >= BSDF_MIRROR;
            // integrateDirect, integrator.cpp:100-130 (nothing for MirrorBrdf)
            int nvis = 0;
            const int li = delta ? -1 : pick_visible_light(sv.lights, sv.n_lights, P, nn, rng.u(vslot + 1), &nvis);
            if (li >= 0) {
                const DLight &l = sv.lights[li];
                const float scale = (float)nvis;
                LightSample ls = TEX ? light_sample_sv(sv, l, P, nn, rng.u(vslot + 2), rng.u(vslot + 3))
                                     : light_sample(l, P, nn, rng.u(vslot + 2), rng.u(vslot + 3));
                float ct = dot3(nn, ls.dir);
                if (b.kind == BSDF_LAMBERT) {
                    // integrateDirectLight, integrator.cpp:33-54
                    if (ct >= 0.f && ls.pdf > 0.f) {
                        f3 f = bsdf_eval(b, ls.dir, vn, nn);
                        f3 c = div3(mul3(mul3v(ls.color, f), ct), ls.pdf);
                        sh[0].valid = true; sh[0].o = P; sh[0].d = ls.dir; sh[0].tmin = AUR_RAY_BIAS; sh[0].tmax = ls.tmax;
                        sh[0].c = mul3v(mul3(c, scale), T);
                    }
                } else {
                    // integrateDirectMIS, integrator.cpp:56-98
                    if (ct > 0.f && ls.pdf > 0.f) {
                        float bpdf = bsdf_pdf(b, ls.dir, vn, nn);
                        if (bpdf > 0.f) {
                            float wgt = power_heuristic(ls.pdf, bpdf);
                            f3 f = bsdf_eval(b, ls.dir, vn, nn);
                            f3 c = div3(mul3(mul3v(mul3(ls.color, wgt), f), ct), ls.pdf);
                            sh[0].valid = true; sh[0].o = P; sh[0].d = ls.dir; sh[0].tmin = AUR_RAY_BIAS; sh[0].tmax = ls.tmax;
                            sh[0].c = mul3v(mul3(c, scale), T);
                        }
                    }
                    BsdfSample bs = bsdf_sample(b, vn, nn, rng.u(vslot + 4), rng.u(vslot + 5));
                    float ct2 = dot3(nn, bs.dir);
                    if (ct2 > 0.f) {
                        float tmax = bs.tmax;
                        float lpdf = TEX ? light_pdf_sv(sv, l, P, bs.dir, AUR_RAY_BIAS, &tmax)
                                         : light_pdf(l, P, bs.dir, AUR_RAY_BIAS, &tmax);
                        if (lpdf > 0.f && bs.pdf > 0.f) {
                            float wgt = power_heuristic(bs.pdf, lpdf);
                            const f3 le = TEX ? light_le_sv(sv, l, bs.dir) : mk3(l.Le[0], l.Le[1], l.Le[2]);
                            f3 c = div3(mul3(mul3v(mul3(bs.color, wgt), le), ct2), bs.pdf);
                            sh[1].valid = true; sh[1].o = P; sh[1].d = bs.dir; sh[1].tmin = AUR_RAY_BIAS; sh[1].tmax = tmax;
                            sh[1].c = mul3v(mul3(c, scale), T);
                        }
                    }
                }
            }
            // continuation, integrator.cpp:208-257
            BsdfSample s = delta ? delta_sample(b, vn, nn) : bsdf_sample(b, vn, nn, rng.u(vslot + 6), rng.u(vslot + 7));
            if (s.pdf <= 0.f) {
                finish = true;
            } else {
                const int bounces = (flags >> 16) & 0xff;
                bool black = false;
                if (!delta) {
                    float ndl = dot3(nn, s.dir);
                    T = mul3v(T, div3(mul3(s.color, ndl), s.pdf));
                    if (bounces > sv.mindepth) cp *= lum3(s.color) * ndl / s.pdf;
                    black = black3(T);
                } else {
                    // integrator.cpp:231-237: no cosine, no pdf, no black test
                    T = mul3v(T, s.color);
                    if (bounces > sv.mindepth) cp *= lum3(s.color);
                }
                if (black) {
                    finish = true;
                } else {
                    int nrt = delta ? RT_MIRROR : (b.kind == BSDF_CT) ? RT_SPECULAR : RT_DIFFUSE;
                    bool killed = false;
                    if (bounces > sv.mindepth) {
                        if (rng.u(vslot + 8) > cp) killed = true;
                        else T = div3(T, cp);
                    }
                    if (killed) {
                        finish = true;
                    } else {
                        cont = true;
                        ndir = s.dir;
                        ntmax = s.tmax;
                        // mirror vertices count as true bounces only
                        flags = nrt | (matte ? F_MATTE : 0) | ((tb + 1) << 8) | ((bounces + (delta ? 0 : 1)) << 16);
                        T4 = make_float4(T.x, T.y, T.z, cp);
                    }
                }
            }
        }
    }
    if (prim < 0) flags |= F_MISS;
    Lb.w += (float)((sh[0].valid ? 1 : 0) + (sh[1].valid ? 1 : 0));
    out.cont = cont;
    out.finish = finish;
    out.sh[0] = sh[0];
    out.sh[1] = sh[1];
    out.P = P;
    out.ndir = ndir;
    out.ntmax = ntmax;
    out.T4 = T4;
    out.La = La;
    out.Lb = Lb;
    out.meta = make_int4(flags, meta.y, meta.z, meta.w);
    out.job = job;
}

// LDS scratch of shade_emit: per-wave counts, then per-wave bases, for the
// next queue, the finish list and the shadow list.
#define SHADE_MAX_WAVES 16
struct EmitScratch {
    int cnt[SHADE_MAX_WAVES][3];
    int base[SHADE_MAX_WAVES][3];
};

// Every thread of the workgroup calls this once per loop round (valid = it
// shaded an entry this round).  With DIRECT, a path that finishes here with
// no shadow ray of this vertex pending (a miss, or the vertex after the last
// one) has its final radiance already: `direct(o)` writes its output in
// place of a finish-list record (48 B written, read back by the finish
// kernel).  Paths whose shadow rays still add to them go to the finish list.
template <bool DIRECT, class Direct>
__device__ inline void shade_emit(const WfView &wv, int q, bool valid, const ShadeOut &o, EmitScratch &es,
                                  Direct direct) {
    const int nq = 1 - q;
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
    const unsigned long long lt = (1ull << lane) - 1ull;
    const bool pd = DIRECT && valid && o.finish && !o.sh[0].valid && !o.sh[1].valid;
    const bool pc = valid && o.cont, pf = valid && o.finish && !pd;
    const bool p0 = valid && o.sh[0].valid, p1 = valid && o.sh[1].valid;
    const unsigned long long mc = __ballot(pc), mf = __ballot(pf), m0 = __ballot(p0), m1 = __ballot(p1);
    if (lane == 0) {
        es.cnt[wave][0] = __popcll(mc);
        es.cnt[wave][1] = __popcll(mf);
        es.cnt[wave][2] = __popcll(m0) + __popcll(m1);
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int k = threadIdx.x;
        const int nw = (blockDim.x + 63) >> 6;
        int tot = 0;
        for (int w = 0; w < nw; ++w) tot += es.cnt[w][k];
        int *ctr = k == 0 ? &wv.counts[nq] : (k == 1 ? &wv.counts[C_FINISH] : &wv.counts[C_SHADOW]);
        int b = tot ? atomicAdd(ctr, tot) : 0;
        for (int w = 0; w < nw; ++w) {
            es.base[w][k] = b;
            b += es.cnt[w][k];
        }
    }
    __syncthreads();
    const int slot = es.base[wave][0] + __popcll(mc & lt);
    const int fs = es.base[wave][1] + __popcll(mf & lt);
    // a wave's light-sample rays first, then its BSDF-sampled rays: each
    // group's records are written contiguously (they add into different
    // accumulators, La and Lb, so their order changes nothing)
    const int ss0 = es.base[wave][2] + __popcll(m0 & lt);
    const int ss1 = es.base[wave][2] + __popcll(m0) + __popcll(m1 & lt);
    if (pc) {
        st_s(&wv.ray[nq][2 * slot], make_float4(o.P.x, o.P.y, o.P.z, AUR_RAY_BIAS));
        st_s(&wv.ray[nq][2 * slot + 1], make_float4(o.ndir.x, o.ndir.y, o.ndir.z, o.ntmax));
        st_s(&wv.T[nq][slot], o.T4);
        st_s(&wv.La[nq][slot], o.La);
        st_s(&wv.Lb[nq][slot], o.Lb);
        st_s(&wv.meta[nq][slot], o.meta);
    }
    if (pf) {
        wv.finLa[fs] = o.La;
        wv.finLb[fs] = o.Lb;
        wv.finMeta[fs] = o.meta;
    }
    if (pd) direct(o);
    const int target = o.cont ? slot : (fs | SH_TO_FINISH);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const bool pk = k ? p1 : p0;
        if (pk) {
            const int s = k ? ss1 : ss0;
            const ShadowOut &h = o.sh[k];
            st_s(&wv.sray[3 * s], make_float4(h.o.x, h.o.y, h.o.z, __int_as_float(o.job)));
            st_s(&wv.sray[3 * s + 1], make_float4(h.d.x, h.d.y, h.d.z, h.tmax));
            st_s(&wv.sray[3 * s + 2], make_float4(h.c.x, h.c.y, h.c.z, __int_as_float(target | (k ? SH_TO_LB : 0))));
        }
    }
}

}  // namespace aur
