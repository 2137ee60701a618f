// adjoint.hip -- path-replay adjoint of the forward integrator.
//
// Gradients of L = sum_pixels dL/dI . I with respect to per-triangle diffuse
// albedo, per-material roughness (Kelemen, and metal with roughnessA =
// roughnessB as aur_set_material_params sets them) and per-material Kelemen
// specular weight,
// for the estimator `Integrator::integrateCameraSample` computes
// (integrator.cpp:152-300) with fixed random numbers.
//
// Path replay (Vicini, Speierer, Jakob 2021, "Path Replay Backpropagation"):
// a path is first replayed to get its radiance L; a second replay walks the
// same vertices (same counter-based random numbers, same rays) carrying the
// remaining radiance R_j = L - sum_{k<j} D_k, where D_k is the direct-lighting
// contribution added at vertex k.  At vertex j with BSDF value f_j:
//   * the continuation weight f_j cos / pdf multiplies every later
//     contribution:   dL += (d log f_j(w_out) / dtheta) * R_{j+1}
//   * the next-event terms use f_j at the light-sample and BSDF-sample
//     directions:     dL += sum_s (d log f_j(w_s) / dtheta) * D_{j,s}
//   * the Kelemen lobe choice (probability A(cos, roughness), chosen without
//     reweighting, kelemenMaterial.cpp:236-253) is a discrete choice whose
//     probability depends on roughness: the score term
//     (d log P(choice) / dtheta) * R_j covers it.
// Sampling pdfs are detached (the sampled directions are treated as fixed),
// which keeps the estimator of the derivative unbiased; MIS weights are
// detached too, which is unbiased because they sum to one per direction.
// The firefly clamp factor 0.2^k is applied to the incoming gradient.
//
// The first replay is the wavefront renderer itself, recording each path's
// radiance and clamp steps (run_paths with PathSource::adj_rec); the second
// runs one thread per path (k_adj_grad); no per-vertex state is stored.  When
// a forward render of the same passes with the same inputs kept those records
// (aur_keep_path_records: 16 bytes per path in HBM), the first replay is
// skipped.  Gradients of small scenes are first summed per workgroup in LDS
// and then added to HBM with one atomic per parameter per workgroup.
#include <hipcub/hipcub.hpp>

#include <cstring>

#include "shade_common.h"
#include "traverse.h"

namespace aur {

#define AUR_ADJ_LDS_PARAMS 6144

struct AdjParams {
    const float4 *dldi;   // per pixel dL/dI (rgb, alpha ignored)
    float inv_passes;
    int first_pass;
    uint32_t seed;
    int n_paths;
    float *g_albedo;      // 3 * n_tris (or null)
    float *g_rough;       // n_attrs
    float *g_spec;        // n_attrs
    float *g_ptex;        // 3 * texels of the albedo texture (or null)
    int n_tris, n_attrs;
    int use_lds;
    unsigned long long *stats;  // paths, ext rays, shadow rays
};

struct GradSink {
    float *albedo, *rough, *spec;  // null when not requested
    float *ptex;                   // albedo texture texels (global memory), or null
    __device__ inline void add(float *p, float v) {
        if (v != 0.f) atomicAdd(p, v);
    }
};

template <int W>
__device__ inline bool occluded(const WNode<W> *nodes, const float4 *tris, f3 o, f3 d, float tmin, float tmax, TravStack &st) {
    TravHit h;
    traverse<W, true>(nodes, tris, o, d, tmin, tmax, st, h);
    return h.prim >= 0;
}

// One replay of a path.  Without GRAD it returns the unclamped radiance and the
// number of 0.2 clamp steps; with GRAD it accumulates gradients given delta
// (incoming dL/dL_path, clamp factor included) and the path's radiance.
template <int W, bool GRAD>
__device__ void replay(const SceneView &sv, const WNode<W> *nodes, TravStack &st, int pixel, const PathRngT<false> &rng,
                       f3 *L_out, int *clamp_steps, bool *camera_miss, f3 delta, f3 L_total, GradSink &g,
                       unsigned long long *n_ext, unsigned long long *n_shadow) {
    const int x = pixel % sv.width, y = pixel / sv.width;
    const float fovx = sv.fov, fovy = ((float)sv.height / (float)sv.width) * fovx;
    const float uu = x + rng.u(0), vv = y + rng.u(1);
    f3 d = norm3(mk3(((2.f * uu - sv.width) / (float)sv.width) * fovx, ((-2.f * vv + sv.height) / (float)sv.height) * fovy, -1.f));
    f3 o = mk3(0.f, 0.f, 0.f);
    TravHit h;
    traverse<W, false>(nodes, sv.tris, o, d, 0.1f, 10000000.f, st, h);
    (*n_ext)++;
    *camera_miss = h.prim < 0;
    f3 L = mk3(0.f, 0.f, 0.f);
    if (h.prim < 0) {
        *L_out = mk3(sv.env[0], sv.env[1], sv.env[2]);
        *clamp_steps = 0;
        return;
    }
    f3 T = mk3(1.f, 1.f, 1.f);
    float cp = 1.f;
    bool matte = false;
    int rtype = 0, tb = 0, bounces = 0;
    f3 R = L_total;
    while (tb < sv.maxdepth) {
        const float t = h.t, u = h.u, v = h.v;
        const float w = (float)(1.0 - (double)u - (double)v);
        const f3 P = add3(o, mul3(d, t));
        const int prim = h.prim;
        const float4 s0 = sv.shade[3 * prim], s1 = sv.shade[3 * prim + 1], s2 = sv.shade[3 * prim + 2];
        const int attr = __float_as_int(s0.w);
        f3 nm = mk3(w * s0.x + u * s1.x + v * s2.x, w * s0.y + u * s1.y + v * s2.y, w * s0.z + u * s1.z + v * s2.z);
        if (dot3(nm, d) > 0.f) nm = mul3(nm, -1.f);
        f3 nn = norm3(nm);
        const f3 vn = neg3(norm3(d));
        if (dot3(vn, nn) <= 0.f) nn = norm3(cross3(vn, cross3(nn, vn)));
        if (rtype == 1) matte = true;
        const DMat m = sv.mats[attr];
        if (rtype == 0) {
            f3 e = mul3v(T, mk3(m.emission[0], m.emission[1], m.emission[2]));
            L = add3(L, e);
            if (GRAD) R = sub3(R, e);
        }
        if (m.type == MAT_LIGHT) break;
        const int vslot = 2 + AUR_SLOTS_PER_VERTEX * tb;
        // shaders are refused by the adjoint entry point; the albedo texture
        // reads the interpolated uv
        ShadeGeo sg;
        sg.P = P; sg.Ng = nn; sg.s = 0.f; sg.t = 0.f; sg.attr = attr;
        const bool ptex_mat = (m.texmask & TEXMASK_PTEX) != 0 && sv.ptex != nullptr;
        if (ptex_mat && sv.uv) {
            const float2 t0 = sv.uv[3 * prim], t1 = sv.uv[3 * prim + 1], t2 = sv.uv[3 * prim + 2];
            sg.s = w * t0.x + u * t1.x + v * t2.x;
            sg.t = w * t0.y + u * t1.y + v * t2.y;
        }
        BrdfChoice ch = get_brdf_ex<false, true>(sv, m, prim, vn, nn, matte, rng.u(vslot + 0), sg);
        const Bsdf &b = ch.b;
        const bool spec_lobe = b.kind == BSDF_CT;
        const bool kel = m.type == MAT_KELEMEN;
        // the roughness of the Cook-Torrance lobe is a parameter: Kelemen, and
        // metal (metalMaterial.cpp:34-62: the lobe picked by u against mix,
        // which does not depend on roughness, so there is no score term)
        const bool rough_lobe = spec_lobe && (kel || m.type == MAT_METAL);
        // the diffuse albedo of this vertex is a parameter when its texel or
        // its triangle's override is being differentiated
        const float *a_val = nullptr;
        float *a_grad = nullptr;
        if (GRAD && b.kind == BSDF_LAMBERT) {
            if (ptex_mat) {
                if (g.ptex) {
                    const size_t ti = 3 * (size_t)ptex_texel(sv, sg.s, sg.t);
                    a_val = sv.ptex + ti;
                    a_grad = g.ptex + ti;
                }
            } else if (sv.albedo && g.albedo) {
                a_val = sv.albedo + 3 * prim;
                a_grad = g.albedo + 3 * prim;
            }
        }
        const bool alb = a_grad != nullptr;
        if (GRAD && kel && ch.kelemen_choice) {
            float dlogp = spec_lobe ? ch.dA / fmaxf(ch.A, 1e-12f) : -ch.dA / fmaxf(1.f - ch.A, 1e-12f);
            g.add(&g.rough[attr], dot3(delta, R) * dlogp * ch.drough);
        }
        // next-event estimation, integrator.cpp:100-130
        f3 cl = mk3(0.f, 0.f, 0.f), cb = mk3(0.f, 0.f, 0.f);
        float dl_l = 0.f, dl_b = 0.f;
        int nvis = 0;
        const int li = pick_visible_light(sv.lights, sv.n_lights, P, nn, rng.u(vslot + 1), &nvis);
        if (li >= 0) {
            const DLight &l = sv.lights[li];
            LightSample ls = light_sample(l, P, nn, rng.u(vslot + 2), rng.u(vslot + 3));
            float ct = dot3(nn, ls.dir);
            if (!spec_lobe) {
                if (ct >= 0.f && ls.pdf > 0.f) {
                    (*n_shadow)++;
                    if (!occluded<W>(nodes, sv.tris, P, ls.dir, AUR_RAY_BIAS, ls.tmax, st))
                        cl = div3(mul3(mul3v(ls.color, bsdf_eval(b, ls.dir, vn, nn)), ct), ls.pdf);
                }
            } else {
                if (ct > 0.f && ls.pdf > 0.f) {
                    float bpdf = bsdf_pdf(b, ls.dir, vn, nn);
                    if (bpdf > 0.f) {
                        (*n_shadow)++;
                        if (!occluded<W>(nodes, sv.tris, P, ls.dir, AUR_RAY_BIAS, ls.tmax, st)) {
                            float wgt = power_heuristic(ls.pdf, bpdf);
                            cl = div3(mul3(mul3v(mul3(ls.color, wgt), bsdf_eval(b, ls.dir, vn, nn)), ct), ls.pdf);
                            if (GRAD) dl_l = ct_dlogd_drough(b.dist, b.rough, dot3(nn, norm3(add3(ls.dir, vn))));
                        }
                    }
                }
                BsdfSample bs = bsdf_sample(b, vn, nn, rng.u(vslot + 4), rng.u(vslot + 5));
                float ct2 = dot3(nn, bs.dir);
                if (ct2 > 0.f) {
                    float tmax = bs.tmax;
                    float lpdf = light_pdf(l, P, bs.dir, AUR_RAY_BIAS, &tmax);
                    if (lpdf > 0.f && bs.pdf > 0.f) {
                        (*n_shadow)++;
                        if (!occluded<W>(nodes, sv.tris, P, bs.dir, AUR_RAY_BIAS, tmax, st)) {
                            float wgt = power_heuristic(bs.pdf, lpdf);
                            cb = div3(mul3(mul3v(mul3(bs.color, wgt), mk3(l.Le[0], l.Le[1], l.Le[2])), ct2), bs.pdf);
                            if (GRAD) dl_b = ct_dlogd_drough(b.dist, b.rough, dot3(nn, norm3(add3(bs.dir, vn))));
                        }
                    }
                }
            }
            const float scale = (float)nvis;
            cl = mul3v(mul3(cl, scale), T);
            cb = mul3v(mul3(cb, scale), T);
            L = add3(L, add3(cl, cb));
            if (GRAD) {
                if (alb) {
                    // d(f cos / pdf)/d(albedo) = f cos / pdf / albedo per channel
                    if (a_val[0] > 0.f) g.add(&a_grad[0], delta.x * cl.x / a_val[0]);
                    if (a_val[1] > 0.f) g.add(&a_grad[1], delta.y * cl.y / a_val[1]);
                    if (a_val[2] > 0.f) g.add(&a_grad[2], delta.z * cl.z / a_val[2]);
                }
                if (kel && spec_lobe) g.add(&g.spec[attr], dot3(delta, add3(cl, cb)) / m.spec[0]);
                if (rough_lobe) g.add(&g.rough[attr], (dot3(delta, cl) * dl_l + dot3(delta, cb) * dl_b) * ch.drough);
                R = sub3(R, add3(cl, cb));
            }
        }
        // continuation, integrator.cpp:208-257
        BsdfSample s = bsdf_sample(b, vn, nn, rng.u(vslot + 6), rng.u(vslot + 7));
        if (s.pdf <= 0.f) break;
        float ndl = dot3(nn, s.dir);
        T = mul3v(T, div3(mul3(s.color, ndl), s.pdf));
        if (bounces > sv.mindepth) cp *= lum3(s.color) * ndl / s.pdf;
        if (black3(T)) break;
        rtype = spec_lobe ? 2 : 1;
        if (bounces > sv.mindepth) {
            if (rng.u(vslot + 8) > cp) break;
            T = div3(T, cp);
        }
        if (GRAD) {
            if (alb) {
                if (a_val[0] > 0.f) g.add(&a_grad[0], delta.x * R.x / a_val[0]);
                if (a_val[1] > 0.f) g.add(&a_grad[1], delta.y * R.y / a_val[1]);
                if (a_val[2] > 0.f) g.add(&a_grad[2], delta.z * R.z / a_val[2]);
            }
            if (rough_lobe) {
                float dr = dot3(delta, R);
                if (kel) g.add(&g.spec[attr], dr / m.spec[0]);
                float c = dot3(nn, norm3(add3(s.dir, vn)));
                g.add(&g.rough[attr], dr * ct_dlogd_drough(b.dist, b.rough, c) * ch.drough);
            }
        }
        o = P;
        d = s.dir;
        ++tb;
        ++bounces;
        traverse<W, false>(nodes, sv.tris, o, d, AUR_RAY_BIAS, s.tmax, st, h);
        (*n_ext)++;
        if (h.prim < 0) break;
    }
    int k = 0;
    f3 Lc = L;
    while (k < AUR_FIREFLY_MAX_STEPS && (Lc.x > AUR_FIREFLY || Lc.y > AUR_FIREFLY || Lc.z > AUR_FIREFLY)) {
        Lc = mul3(Lc, 0.2f);
        ++k;
    }
    *L_out = L;
    *clamp_steps = k;
}

// per-wave sums of the ray counters, one atomic each
__device__ inline void adj_count(unsigned long long *stats, unsigned paths, unsigned ext, unsigned shadow) {
    for (int off = 32; off > 0; off >>= 1) {
        paths += __shfl_xor(paths, off);
        ext += __shfl_xor(ext, off);
        shadow += __shfl_xor(shadow, off);
    }
    if (__lane_id() == 0) {
        atomicAdd(&stats[0], (unsigned long long)paths);
        atomicAdd(&stats[1], (unsigned long long)ext);
        atomicAdd(&stats[2], (unsigned long long)shadow);
    }
}

__device__ inline void adj_path(const SceneView &sv, const AdjParams &ap, int i, int *pixel, PathRngT<false> *rng) {
    const int wh = sv.width * sv.height;
    const int pass = udiv(i, sv.div_wh);
    *pixel = tiled_pixel(sv, i - pass * wh);
    *rng = make_path_rng<false>(sv, nullptr, ap.seed, *pixel, (uint32_t)(ap.first_pass + pass));
}

// Second replay, carrying the remaining radiance, for paths with a nonzero
// incoming gradient.  Gradients of small scenes are summed per workgroup in
// dynamically sized LDS first.
// Sort keys of the gradient replay's work: longest paths first (the
// wavefront replay recorded each path's ray count), so the lanes of a wave
// replay paths of equal length and the longest ones start first.
__global__ void k_adj_keys(const float4 *__restrict__ rec, int n, unsigned *__restrict__ keys, int *__restrict__ idx) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int w = __float_as_int(rec[i].w);
    keys[i] = 255u - (unsigned)((w >> 17) & 0xff);
    idx[i] = i;
}

// Compiled for 6 waves per SIMD (at most 80 VGPRs): more paths in flight
// outweigh the spills that bound costs.
template <int W>
__global__ void __launch_bounds__(128, 6) k_adj_grad(SceneView sv, AdjParams ap, const float4 *rec, const int *order,
                                                  int *ovf, int ovf_per_thread) {
    __shared__ int s_stack[AUR_STACK_LDS * 128];
    extern __shared__ float s_grad[];
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    TravStack st;
    st.init(s_stack, 128, ovf, ovf_per_thread);
    const WNode<W> *nodes = static_cast<const WNode<W> *>(sv.nodes);
    const int n_alb = ap.g_albedo ? 3 * ap.n_tris : 0;
    GradSink g;
    if (ap.use_lds) {
        for (int i = threadIdx.x; i < n_alb + 2 * ap.n_attrs; i += blockDim.x) s_grad[i] = 0.f;
        __syncthreads();
        g.albedo = n_alb ? s_grad : nullptr;
        g.rough = s_grad + n_alb;
        g.spec = s_grad + n_alb + ap.n_attrs;
    } else {
        g.albedo = ap.g_albedo;
        g.rough = ap.g_rough;
        g.spec = ap.g_spec;
    }
    g.ptex = ap.g_ptex;
    unsigned long long n_ext = 0, n_shadow = 0;
    for (int j = gtid; j < ap.n_paths; j += gridDim.x * blockDim.x) {
        const int i = order[j];
        const float4 r = rec[i];
        const int kw = __float_as_int(r.w);
        if (kw & 0x10000) continue;  // camera miss: no vertex depends on the parameters
        int pixel;
        PathRngT<false> rng;
        adj_path(sv, ap, i, &pixel, &rng);
        const float4 dl = ap.dldi[pixel];
        float cf = ap.inv_passes;
        for (int j = 0; j < (kw & 0xffff); ++j) cf *= 0.2f;
        const f3 delta = mk3(dl.x * cf, dl.y * cf, dl.z * cf);
        if (delta.x == 0.f && delta.y == 0.f && delta.z == 0.f) continue;
        f3 L2;
        int k2;
        bool miss;
        replay<W, true>(sv, nodes, st, pixel, rng, &L2, &k2, &miss, delta, mk3(r.x, r.y, r.z), g, &n_ext, &n_shadow);
    }
    adj_count(ap.stats, 0u, (unsigned)n_ext, (unsigned)n_shadow);
    if (ap.use_lds) {
        __syncthreads();
        for (int i = threadIdx.x; i < n_alb; i += blockDim.x)
            if (s_grad[i] != 0.f) atomicAdd(&ap.g_albedo[i], s_grad[i]);
        for (int i = threadIdx.x; i < ap.n_attrs; i += blockDim.x) {
            if (s_grad[n_alb + i] != 0.f) atomicAdd(&ap.g_rough[i], s_grad[n_alb + i]);
            if (s_grad[n_alb + ap.n_attrs + i] != 0.f) atomicAdd(&ap.g_spec[i], s_grad[n_alb + ap.n_attrs + i]);
        }
    }
}

// The records a forward render kept for passes [first_pass, first_pass +
// n_passes), or null when any of them is missing or was recorded with other
// inputs or another image size.
static float4 *kept_records(Ctx *c, int first_pass, int n_passes) {
    const size_t wh = (size_t)c->scene.opt.width * c->scene.opt.height;
    if (c->rec_n <= 0 || !c->path_rec.p || c->rec_wh != wh || c->scene.opt.lightsamples != 1) return nullptr;
    if (first_pass < c->rec_first || first_pass + n_passes > c->rec_first + c->rec_n) return nullptr;
    for (int p = first_pass; p < first_pass + n_passes; ++p)
        if (c->rec_version[p - c->rec_first] != c->version) return nullptr;
    return c->path_rec.as<float4>() + (size_t)(first_pass - c->rec_first) * wh;
}

void adjoint(Ctx *c, int first_pass, int n_passes, const float4 *d_dldi, float *d_galb, float *d_grough, float *d_gspec,
             float *d_gptex, aur_stats *stats) {
    if (!c->built) throw StateError("adjoint before aur_build_accel");
    SceneView sv = scene_view(c);
    const int wh = c->scene.opt.width * c->scene.opt.height;
    const long long n = (long long)wh * n_passes;
    if (n > 0x7fffffffLL) throw CapacityError("too many paths for one adjoint call");
    DevBuf st;
    st.reserve(sizeof(unsigned long long) * 4);
    fill_zero(c, st.p, sizeof(unsigned long long) * 4);
    AdjParams ap;
    ap.dldi = d_dldi;
    ap.inv_passes = 1.f / (float)n_passes;
    ap.first_pass = first_pass;
    ap.seed = c->seed;
    ap.n_paths = (int)n;
    ap.g_albedo = d_galb;
    ap.g_rough = d_grough;
    ap.g_spec = d_gspec;
    ap.g_ptex = d_gptex;
    ap.n_tris = c->n_tris;
    ap.n_attrs = c->n_mats;
    ap.use_lds = ((d_galb ? 3 * c->n_tris : 0) + 2 * c->n_mats) <= AUR_ADJ_LDS_PARAMS;
    ap.stats = st.as<unsigned long long>();
    // traversal stack spill area, as for the wavefront kernels
    int need = (c->bvh.width - 1) * (c->bvh.depth + 1) + 2;
    size_t ovf = need > AUR_STACK_LDS ? (size_t)(need - AUR_STACK_LDS) : 1;
    // latency-bound replays: 24 waves per CU (6 per SIMD, the kernel's
    // launch bound: at 3 waves without it the replay measured 25 % slower
    // even though the bound spills a few registers)
    const int blocks = std::max(1, std::min((int)((n + 127) / 128), c->n_cu * 12));
    DevBuf sovf, own_rec;
    sovf.reserve(sizeof(int) * (size_t)blocks * 128 * ovf);
    const size_t lds = ap.use_lds ? sizeof(float) * (size_t)((d_galb ? 3 * c->n_tris : 0) + 2 * c->n_mats) : 0;
    aur_stats fst;
    std::memset(&fst, 0, sizeof fst);
    float4 *rec = kept_records(c, first_pass, n_passes);
    if (rec) {
        // a forward render of these passes with the current inputs kept the
        // records: the first replay is not needed
        fst.paths = n;
    } else {
        // first replay: the wavefront renderer traces the same paths (same
        // counter-based random numbers) and records each path's radiance and
        // clamp steps instead of adding it to a framebuffer
        own_rec.reserve(sizeof(float4) * (size_t)std::max<long long>(n, 1));
        rec = own_rec.as<float4>();
        PathSource src;
        src.mode = 0;
        src.first_pass = first_pass;
        src.ls = 1;
        src.pixel = src.sample = nullptr;
        src.slots = nullptr;
        src.adj_rec = rec;
        src.rec_base = 0;
        src.rec_count = n;
        src.rec_only = true;
        run_paths(c, src, (size_t)n, nullptr, &fst);
    }
    // replay order: longest paths first
    DevBuf keys[2], idx[2], stemp;
    for (int k = 0; k < 2; ++k) {
        keys[k].reserve(sizeof(unsigned) * (size_t)std::max<long long>(n, 1));
        idx[k].reserve(sizeof(int) * (size_t)std::max<long long>(n, 1));
    }
    hipLaunchKernelGGL(k_adj_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, rec, (int)n,
                       keys[0].as<unsigned>(), idx[0].as<int>());
    size_t st_bytes = 0;
    HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, st_bytes, keys[0].as<unsigned>(), keys[1].as<unsigned>(),
                                                 idx[0].as<int>(), idx[1].as<int>(), (int)n, 0, 8, c->stream));
    stemp.reserve(std::max<size_t>(st_bytes, 1));
    HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(stemp.p, st_bytes, keys[0].as<unsigned>(), keys[1].as<unsigned>(),
                                                 idx[0].as<int>(), idx[1].as<int>(), (int)n, 0, 8, c->stream));
    // second replay, carrying the remaining radiance
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    HIP_CHECK(hipEventRecord(e0, c->stream));
    if (c->bvh.width == 8)
        hipLaunchKernelGGL(k_adj_grad<8>, dim3(blocks), dim3(128), lds, c->stream, sv, ap, rec,
                           idx[1].as<int>(), sovf.as<int>(), (int)ovf);
    else
        hipLaunchKernelGGL(k_adj_grad<4>, dim3(blocks), dim3(128), lds, c->stream, sv, ap, rec,
                           idx[1].as<int>(), sovf.as<int>(), (int)ovf);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipEventRecord(e1, c->stream));
    HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    unsigned long long h[4];
    HIP_CHECK(hipMemcpy(h, st.p, sizeof h, hipMemcpyDeviceToHost));
    if (stats) {
        stats->paths = fst.paths;
        stats->ext_rays = fst.ext_rays + (int64_t)h[1];
        stats->hit_rays = 0;
        stats->shadow_rays = fst.shadow_rays + (int64_t)h[2];
        stats->ms = fst.ms + ms;
    }
}

}  // namespace aur
