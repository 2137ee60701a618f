// wavefront.hip -- the path-tracing hot path as a wavefront of kernels.
//
// Replaces the per-pixel recursion of `Integrator::integrateCameraSample`
// (src/frontend/integrator.cpp:152-300) run inside the TBB pixel loop
// (src/frontend/renderer.cpp:147-271).  One batch holds many paths (all
// pixels of several passes); per bounce:
//
//   extend  -- closest-hit traversal of the extension-ray queue
//   shade   -- one vertex of every live path: emission, getBrdf, next-event
//              estimation (light sample, and for specular lobes a BSDF sample
//              with the power heuristic), next direction, throughput, roulette;
//              writes shadow rays, the next extension queue and the finish list
//   shadow  -- any-hit traversal; unoccluded contributions are added to the
//              path's radiance (one accumulator per shadow-ray kind, so the
//              two rays of one path never race)
//   finish  -- firefly clamp (integrator.cpp:277-281) and per-path output
//
// Queues are compacted with one atomic per wave: a 64-lane ballot gives each
// lane its offset (mbcnt) and the wave's count (popcount).  Every kernel is
// a grid-stride loop over a count that lives in device memory, so a whole
// frame is a fixed sequence of launches with no host round trip.  Shadow and
// finish of bounce b run on a second stream beside extend of bounce b + 1
// (events order them before shade of bounce b + 1), so the tail of one
// traversal kernel overlaps the start of the next.
#include "wavefront_common.h"

#include <mutex>

namespace aur {

// Path id -> pixel, pass and light sample.  Mode 0 issues a batch of passes
// starting at first_pass: pass 0 holds one path per pixel, later passes ls
// (one block of w*h paths per light sample), each block in 8x8 tile order.
// `local` is the id's position in its block.
__device__ inline void path_decode(const SceneView &sv, const WfView &wv, int path, int *pixel, int *sample, int *sub,
                                   int *local) {
    if (wv.src.mode == 0) {
        const int wh = sv.width * sv.height;
        int id = path, p = wv.src.first_pass;
        if (p == 0 && id < wh) {
            *sample = 0; *sub = 0; *local = id;
        } else {
            if (p == 0) { id -= wh; p = 1; }
            const int per = wv.src.ls * wh;
            const int k = udiv(id, wv.div_per);
            const int r = id - k * per;
            *sample = p + k;
            *sub = udiv(r, sv.div_wh);
            *local = r - *sub * wh;
        }
        *pixel = tiled_pixel(sv, *local);
    } else {
        *pixel = wv.src.pixel[path];
        *sample = wv.src.sample[path];
        *sub = 0;
        *local = 0;
    }
}

template <bool QMC = true>
__device__ inline PathRngT<QMC> path_rng(const SceneView &sv, const WfView &wv, int path) {
    int pixel, sample, sub, local;
    path_decode(sv, wv, path, &pixel, &sample, &sub, &local);
    return make_path_rng<QMC>(sv, wv.src.slots ? wv.src.slots + (size_t)path * wv.nslots : nullptr, wv.seed, pixel,
                              (uint32_t)sample, sub, wv.src.ls);
}

// ---------------------------------------------------------------------------
// generate: camera.cpp:34-62
// ---------------------------------------------------------------------------
__global__ void k_generate(SceneView sv, WfView wv, int n_paths) {
    int stride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_paths; i += stride) {
        camera_ray_of(sv, path_rng(sv, wv, i), &wv.ray[0][2 * i], &wv.ray[0][2 * i + 1]);
        wv.T[0][i] = make_float4(1.f, 1.f, 1.f, 1.f);
        wv.La[0][i] = make_float4(0.f, 0.f, 0.f, 0.f);
        wv.Lb[0][i] = make_float4(0.f, 0.f, 0.f, 0.f);
        wv.meta[0][i] = make_int4(RT_CAMERA, i, 0, 0);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        wv.counts[C_Q0] = n_paths;
        atomicAdd(&wv.stats[S_PATHS], (unsigned long long)n_paths);
    }
}

// Hash-sampled renders store no camera rays: the first extension kernel and
// the first shade kernel recompute each path's camera ray from its id (the
// 96 bytes per path k_generate would write and they would read back).  This
// only sets the queue's size.
__global__ void k_start(WfView wv, int n_paths) {
    wv.counts[C_Q0] = n_paths;
    atomicAdd(&wv.stats[S_PATHS], (unsigned long long)n_paths);
}

#ifdef AUR_TRAV_STATS
// measurement builds only: per traversal kernel (0 extend, 1 shadow) nodes
// visited, leaves and triangles tested, summed over all lanes
__device__ unsigned long long g_trav_stats[2][4];
__device__ inline void trav_stats_add(int k, const TravStack &st) {
    unsigned long long v = st.n_visits, l = st.n_leaves, t = st.n_tris;
    for (int off = 32; off > 0; off >>= 1) {
        v += __shfl_xor(v, off);
        l += __shfl_xor(l, off);
        t += __shfl_xor(t, off);
    }
    if (__lane_id() == 0) {
        atomicAdd(&g_trav_stats[k][0], v);
        atomicAdd(&g_trav_stats[k][1], l);
        atomicAdd(&g_trav_stats[k][2], t);
    }
}
#endif

// ---------------------------------------------------------------------------
// extend: closest hit.  Lanes are refilled independently (traverse_stream):
// closest-hit traversals of incoherent rays differ in length several times
// over, and a wave that waited for its longest ray left most lanes idle.
// Both traversal kernels are held to 64 VGPRs (8 waves per SIMD) and 5 KB of
// LDS per wave, so a CU keeps its full 32 waves of dependent node fetches in
// flight.
// ---------------------------------------------------------------------------
template <int W, int B, bool CAM = false>
__global__ void __launch_bounds__(B, AUR_TRAV_WAVES(W)) k_extend(SceneView sv, WfView wv, int q) {
    __shared__ int s_stack[AUR_STACK_LDS * B];
    const int n = wv.counts[q];
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&wv.stats[S_EXT], (unsigned long long)n);
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const int stride = gridDim.x * blockDim.x;
    TravStack st;
    st.init(s_stack, B, wv.ovf, wv.ovf_per_thread);
    const WNode<W> *nodes = static_cast<const WNode<W> *>(sv.nodes);
    int hits = 0;
    traverse_stream<W, false>(
        nodes, sv.tris, st, gtid, n, stride, AUR_REFILL_LANES,
        [&](int i, f3 &o, f3 &d, float &tmin, float &tmax) {
            float4 a, b;
            if constexpr (CAM) camera_ray_of(sv, path_rng<false>(sv, wv, i), &a, &b);
            else { a = ld_s(&wv.ray[q][2 * i]); b = ld_s(&wv.ray[q][2 * i + 1]); }
            o = mk3(a.x, a.y, a.z); d = mk3(b.x, b.y, b.z); tmin = a.w; tmax = b.w;
        },
        [&](int i, const TravHit &h) {
            st_s(&wv.hit[i], make_float4(h.t, h.u, h.v, __int_as_float(h.prim)));
            hits += h.prim >= 0 ? 1 : 0;
        });
    // one add per wave into this workgroup's partial (no global same-address atomics)
    for (int off = 32; off > 0; off >>= 1) hits += __shfl_xor(hits, off);
    if (__lane_id() == 0 && hits) atomicAdd(&wv.partials[blockIdx.x], (unsigned long long)hits);
#ifdef AUR_TRAV_STATS
    trav_stats_add(0, st);
#endif
}

// ---------------------------------------------------------------------------
// shadow: any hit; unoccluded contributions go to La (light sample) or Lb
// (BSDF sample).  Any-hit traversals are short and end at the first hit; here
// the per-lane refill measured slower than whole-wave iteration.
// ---------------------------------------------------------------------------
template <int W, int B>
__global__ void __launch_bounds__(B, AUR_TRAV_WAVES(W)) k_shadow(SceneView sv, WfView wv, int nq) {
    __shared__ int s_stack[AUR_STACK_LDS * B];
    const int n = wv.counts[C_SHADOW];
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&wv.stats[S_SHADOW], (unsigned long long)n);
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const int stride = gridDim.x * blockDim.x;
    TravStack st;
    st.init(s_stack, B, wv.sovf, wv.ovf_per_thread);
    const WNode<W> *nodes = static_cast<const WNode<W> *>(sv.nodes);
    for (int i = gtid; i < n; i += stride) {
        float4 a = ld_s(&wv.sray[3 * i]), b = ld_s(&wv.sray[3 * i + 1]), c = ld_s(&wv.sray[3 * i + 2]);
        TravHit h;
        traverse<W, true>(nodes, sv.tris, mk3(a.x, a.y, a.z), mk3(b.x, b.y, b.z), AUR_RAY_BIAS, b.w, st, h);
        if (h.prim < 0) shadow_add(wv, nq, __float_as_int(c.w), c);
    }
#ifdef AUR_TRAV_STATS
    trav_stats_add(1, st);
#endif
}

// One finished path: its radiance La + Lb (clamped) goes to its output slot;
// ext_dup / hit_dup count the camera rays a whole-pass render traced more
// than once (later light samples of a pass share the pass's camera ray).
__device__ inline void finish_path(const SceneView &sv, const WfView &wv, const float4 a, const float4 b,
                                   const int4 m, unsigned &ext_dup, unsigned &hit_dup) {
    const f3 L = firefly_clamp(mk3(a.x + b.x, a.y + b.y, a.z + b.z), m.x);
    // whole passes land by (pass, light sample, pixel) for the resolve;
    // a camera miss is marked by alpha + 2 (the resolve takes it once);
    // injected paths land by id
    int slot = m.y;
    float w = a.w;
    if (wv.src.mode == 0) {
        int pixel, sample, sub, local;
        path_decode(sv, wv, m.y, &pixel, &sample, &sub, &local);
        slot = m.y - local + pixel;
        const bool cam_miss = (m.x & F_MISS) && path_ext_rays(m.x) == 1;
        if (cam_miss) w += 2.f;
        // the reference traces a pass's camera ray once for all its light
        // samples (integrator.cpp:160-164): later samples do not count it
        if (sub > 0) { ext_dup += 1; hit_dup += cam_miss ? 0 : 1; }
    }
    st_s(&wv.out[slot], make_float4(L.x, L.y, L.z, w));
    const long long ri = wv.src.rec_base + m.y;
    if (wv.src.adj_rec && ri >= 0 && ri < wv.src.rec_count) {
        // the adjoint's first replay: radiance before the clamp and the
        // number of 0.2 steps the clamp took (integrator.cpp:276-283)
        f3 Lc = mk3(a.x + b.x, a.y + b.y, a.z + b.z);
        int k = 0;
        if (!(m.x & F_NOCLAMP))
            while (k < AUR_FIREFLY_MAX_STEPS && (Lc.x > AUR_FIREFLY || Lc.y > AUR_FIREFLY || Lc.z > AUR_FIREFLY)) {
                Lc = mul3(Lc, 0.2f);
                ++k;
            }
        const bool cam_miss = (m.x & F_MISS) && path_ext_rays(m.x) == 1;
        // bits 0-15 clamp steps, bit 16 camera miss, bits 17-24 rays traced
        wv.src.adj_rec[ri] = make_float4(
            a.x + b.x, a.y + b.y, a.z + b.z,
            __int_as_float(k | (cam_miss ? 0x10000 : 0) | (min(path_ext_rays(m.x), 255) << 17)));
    }
}

// Takes the duplicate camera-ray counts out of the run's statistics (every
// lane of the wave calls this).
__device__ inline void finish_dups(const WfView &wv, unsigned ext_dup, unsigned hit_dup) {
    if (wv.src.mode == 0 && wv.src.ls > 1) {
        for (int o = 32; o > 0; o >>= 1) {
            ext_dup += __shfl_xor(ext_dup, o);
            hit_dup += __shfl_xor(hit_dup, o);
        }
        if ((threadIdx.x & 63) == 0 && ext_dup) {
            atomicAdd(&wv.stats[S_EXT], (unsigned long long)(0ull - ext_dup));
            atomicAdd(&wv.stats[S_HIT], (unsigned long long)(0ull - hit_dup));
        }
    }
}

// An entry of extension queue q with no vertex to shade -- its ray missed,
// or the vertex is past maxdepth (every entry of the last bounce: all
// entries of a bounce are the same vertex of their paths) -- finishes as
// shading would finish it (integrator.cpp:287-297, 262-271, then the clamp
// and output of integrator.cpp:276-283).
template <bool TEX>
__device__ inline void finish_unshaded(const SceneView &sv, const WfView &wv, int q, int i, int prim,
                                       unsigned &ext_dup, unsigned &hit_dup) {
    const int4 meta = ld_s(&wv.meta[q][i]);
    float4 La = ld_s(&wv.La[q][i]);
    const float4 Lb = ld_s(&wv.Lb[q][i]);
    int flags = meta.x;
    if (prim < 0) {
        miss_radiance<TEX>(
            sv, meta, [&](int4 m) { return path_rng<TEX>(sv, wv, m.y); },
            [&] { return ld_s(&wv.ray[q][2 * i + 1]); }, [&] { return ld_s(&wv.T[q][i]); }, La, flags);
        flags |= F_MISS;
    }
    finish_path(sv, wv, La, Lb, make_int4(flags, meta.y, meta.z, meta.w), ext_dup, hit_dup);
}

// ---------------------------------------------------------------------------
// shade: one path vertex per queue entry with a vertex to shade.  Each
// workgroup takes tiles of SHADE_TILE entries: it reads their hit records,
// finishes the entries with no vertex to shade (finish_unshaded), lists the
// others in LDS in queue order, and shades the list in rounds of a full
// workgroup, carrying fewer than a round's worth into its next tile.  Most
// bounce rays of an open scene miss, and a wave of queue entries almost
// always holds a few hits: shading the queue in place ran the whole vertex
// code on nearly every wave with a quarter of its lanes.
// ---------------------------------------------------------------------------
// 96 VGPRs (5 waves per SIMD) without the shader code; the shader, normal-map
// and environment-map code needs more registers and runs at 4.
#define SHADE_PER_THREAD 4
#define SHADE_TILE (SHADE_PER_THREAD * 256)
template <bool TEX, bool CAM = false>
__global__ void __launch_bounds__(256, TEX ? 4 : 5) k_shade(SceneView sv, WfView wv, int q, int last,
                                                                          int list_tiles) {
    __shared__ EmitScratch es;
    __shared__ int list[SHADE_TILE + 256];
    __shared__ int cnt[SHADE_PER_THREAD][4];
    const int n = wv.counts[q];
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
    const unsigned long long lt = (1ull << lane) - 1ull;
    // Listing pays off when each workgroup has tiles enough for the hit
    // records' read to hide behind other workgroups' shading (list_tiles:
    // four per workgroup by default); a shorter queue, and the camera
    // bounce, are shaded in place, a workgroup's worth of entries per round
    // (shade_compute finishes its misses itself).
    const bool listed = !CAM && (n + SHADE_TILE - 1) / SHADE_TILE >= list_tiles;
    const int tile_size = listed ? SHADE_TILE : 256;
    const int n_tiles = (n + tile_size - 1) / tile_size;
    unsigned ext_dup = 0, hit_dup = 0;
    int have = 0;  // entries waiting to be shaded (the same in every thread)
    for (int tile = blockIdx.x;; tile += gridDim.x) {
        const bool more = tile < n_tiles;
        const int base = tile * tile_size;
        if (more && !listed) {
            have = min(256, n - base);
        } else if (more) {
            int prim[SHADE_PER_THREAD];
#pragma unroll
            for (int k = 0; k < SHADE_PER_THREAD; ++k) {
                const int i = base + k * 256 + (int)threadIdx.x;
                prim[k] = i < n ? __float_as_int(ld_s(&wv.hit[i]).w) : 0;
            }
            unsigned long long m[SHADE_PER_THREAD];
#pragma unroll
            for (int k = 0; k < SHADE_PER_THREAD; ++k) {
                const int i = base + k * 256 + (int)threadIdx.x;
                const bool valid = i < n, fin = valid && (prim[k] < 0 || last);
                m[k] = __ballot(valid && !fin);
                if (fin) finish_unshaded<TEX>(sv, wv, q, i, prim[k], ext_dup, hit_dup);
            }
            if (lane == 0)
#pragma unroll
                for (int k = 0; k < SHADE_PER_THREAD; ++k) cnt[k][wave] = __popcll(m[k]);
            __syncthreads();
            // list positions in queue order: entry group k, then wave
            int pos = have, tot = 0;
#pragma unroll
            for (int k = 0; k < SHADE_PER_THREAD; ++k)
                for (int w = 0; w < 4; ++w) {
                    const int c = cnt[k][w];
                    if (w == wave && ((m[k] >> lane) & 1ull))
                        list[pos + tot + __popcll(m[k] & lt)] = base + k * 256 + (int)threadIdx.x;
                    tot += c;
                }
            have += __builtin_amdgcn_readfirstlane(tot);
            __syncthreads();
        }
        int head = 0;
        while (have - head >= 256 || ((!more || !listed) && have > head)) {
            const int j = head + (int)threadIdx.x;
            const bool valid = j < have;
            ShadeOut o;
            o.cont = o.finish = false;
            o.sh[0].valid = o.sh[1].valid = false;
            if (valid)
                shade_compute<TEX, CAM>(sv, wv, q, listed ? list[j] : base + j,
                                        [&](int4 m) { return path_rng<TEX>(sv, wv, m.y); }, 0, o);
            // (its barriers also end this round's reads of the list)
            shade_emit<true>(wv, q, valid, o, es,
                             [&](const ShadeOut &f) { finish_path(sv, wv, f.La, f.Lb, f.meta, ext_dup, hit_dup); });
            head += 256;
        }
        if (!more) break;
        if (!listed) {
            have = 0;
            continue;
        }
        // the rest (fewer than 256) moves to the front; head >= 256 when any
        // round ran, so the ranges do not overlap
        if (head > 0) {
            const int rest = have - head;
            if ((int)threadIdx.x < rest) list[threadIdx.x] = list[head + threadIdx.x];
            have = rest;
            __syncthreads();
        }
    }
    finish_dups(wv, ext_dup, hit_dup);
}

// ---------------------------------------------------------------------------
// finish: firefly clamp (integrator.cpp:277-281), per-path output
// ---------------------------------------------------------------------------
// The paths that finished with shadow rays of their last vertex pending
// (the others were written by k_shade).
__global__ void k_finish(SceneView sv, WfView wv) {
    const int n = wv.counts[C_FINISH];
    const int stride = gridDim.x * blockDim.x;
    unsigned ext_dup = 0, hit_dup = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        finish_path(sv, wv, wv.finLa[i], wv.finLb[i], wv.finMeta[i], ext_dup, hit_dup);
    finish_dups(wv, ext_dup, hit_dup);
}

__global__ void k_clear_counts(int *counts, int q_done) {
    // the consumed queue and the per-bounce shadow/finish lists restart at zero
    if (threadIdx.x == 0) {
        counts[q_done] = 0;
        counts[C_SHADOW] = 0;
        counts[C_FINISH] = 0;
    }
}

__global__ void k_reset_all(int *counts, unsigned long long *stats, int reset_stats) {
    int i = threadIdx.x;
    if (i < 16) counts[i] = 0;
    if (reset_stats && i < 8) stats[i] = 0;
}

// adds the per-workgroup hit counts into the run's statistics and clears them
__global__ void k_sum_partials(unsigned long long *partials, int n, unsigned long long *stats) {
    __shared__ unsigned long long s[256];
    unsigned long long v = 0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        v += partials[k];
        partials[k] = 0;
    }
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) stats[S_HIT] += s[0];
}

// per-pixel sum over the passes of a batch, in pass order (deterministic).  A
// pass with ls light samples contributes the sum of its paths times 1/ls in
// sample order, or its first path alone where the camera ray missed
// (integrator.cpp:176-292: result += Lo / numSamples; a miss returns the env).
__device__ inline float4 pass_value(const float4 *__restrict__ out, size_t off, int wh, int n_sub, int p) {
    const float4 v0 = out[off + p];
    if (v0.w >= 2.f) return make_float4(v0.x, v0.y, v0.z, v0.w - 2.f);
    if (n_sub == 1) return v0;
    const float inv = 1.f / (float)n_sub;
    float4 s = make_float4(0.f, 0.f, 0.f, v0.w);
    for (int j = 0; j < n_sub; ++j) {
        const float4 v = out[off + (size_t)j * wh + p];
        s.x += v.x * inv; s.y += v.y * inv; s.z += v.z * inv;
    }
    return s;
}

__global__ void k_resolve(const float4 *__restrict__ out, int wh, int first_pass, int n_passes, int ls,
                          float4 *__restrict__ accum, ushort4 *__restrict__ display, int display_base) {
    int stride = gridDim.x * blockDim.x;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < wh; p += stride) {
        float4 s = accum[p];
        ushort4 d = display ? display[p] : make_ushort4(0, 0, 0, 0);
        size_t off = 0;
        for (int k = 0; k < n_passes; ++k) {
            const int n_sub = (first_pass + k == 0) ? 1 : ls;
            const float4 v = pass_value(out, off, wh, n_sub, p);
            s.x += v.x; s.y += v.y; s.z += v.z; s.w += v.w;
            // the display's n-th appended pass since it was cleared has weight 1/n
            if (display) d = display_update(d, v, display_base + k + 1);
            off += (size_t)n_sub * wh;
        }
        accum[p] = s;
        if (display) display[p] = d;
    }
}

// ---------------------------------------------------------------------------
// per-kernel timing (events on the library stream around each launch)
// ---------------------------------------------------------------------------
const char *const kKernelNames[KT_COUNT] = {"k_generate", "k_extend", "k_shade", "k_shadow", "k_finish", "k_resolve"};

static hipEvent_t timer_event(Ctx *c) {
    KernelTimer &t = c->timer;
    if (!t.pool.empty()) {
        hipEvent_t e = t.pool.back();
        t.pool.pop_back();
        return e;
    }
    hipEvent_t e;
    HIP_CHECK(hipEventCreate(&e));
    return e;
}

void timer_begin(Ctx *c, int kernel, hipEvent_t *start, hipStream_t stream) {
    (void)kernel;
    *start = nullptr;
    if (!c->timer.on) return;
    *start = timer_event(c);
    HIP_CHECK(hipEventRecord(*start, stream));
}

void timer_end(Ctx *c, int kernel, hipEvent_t start, hipStream_t stream) {
    if (!c->timer.on || !start) return;
    hipEvent_t e = timer_event(c);
    HIP_CHECK(hipEventRecord(e, stream));
    c->timer.pending.push_back({kernel, {start, e}});
}

void timer_collect(Ctx *c) {
    KernelTimer &t = c->timer;
    for (auto &p : t.pending) {
        HIP_CHECK(hipEventSynchronize(p.second.second));
        float ms = 0.f;
        HIP_CHECK(hipEventElapsedTime(&ms, p.second.first, p.second.second));
        t.total_ms[p.first] += ms;
        t.launches[p.first] += 1;
        t.pool.push_back(p.second.first);
        t.pool.push_back(p.second.second);
    }
    t.pending.clear();
}

void timer_release(Ctx *c) {
    timer_collect(c);
    for (hipEvent_t e : c->timer.pool) (void)hipEventDestroy(e);
    c->timer.pool.clear();
}

#define TIMED_ON(c, kid, strm, ...)       \
    do {                                  \
        hipEvent_t ev_;                   \
        timer_begin(c, kid, &ev_, strm);  \
        __VA_ARGS__;                      \
        timer_end(c, kid, ev_, strm);     \
    } while (0)
#define TIMED(c, kid, ...) TIMED_ON(c, kid, (c)->stream, __VA_ARGS__)

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------
bool shade_needs_tex(const Ctx *c) {
    if (c->has_attrx || c->sampler != AUR_SAMPLER_HASH || !c->ptex.empty()) return true;
    for (const DLight &l : c->dlights)
        if (l.type == LIGHT_ENV && l.tex >= 0) return true;
    return false;
}

SceneView scene_view(Ctx *c) {
    SceneView sv;
    sv.nodes = c->bvh.nodes.p;
    sv.tris = c->bvh.tri_sorted.as<float4>();
    sv.shade = c->d_shade.as<float4>();
    sv.mats = c->d_mats.as<DMat>();
    sv.lights = c->d_lights.as<DLight>();
    sv.tables = c->d_tables.as<float>();
    sv.albedo = c->tri_albedo.empty() ? nullptr : c->d_albedo.as<float>();
    sv.uv = c->has_uv ? c->d_uv.as<float2>() : nullptr;
    sv.attrx = c->has_attrx ? c->d_attrx.as<DAttrX>() : nullptr;
    sv.shaders = c->d_shaders.as<DShader>();
    sv.texs = c->d_texs.as<DTex>();
    sv.texels = c->d_texels.as<ushort4>();
    sv.qmc = (c->sampler == AUR_SAMPLER_HALTON && c->d_halton_view.p) ? c->d_halton_view.as<HaltonView>() : nullptr;
    sv.ptex = c->ptex.empty() ? nullptr : c->d_ptex.as<float>();
    sv.ptex_w = c->ptex_w;
    sv.ptex_h = c->ptex_h;
    sv.n_lights = c->n_lights;
    sv.has_env = c->n_env > 0;
    sv.env[0] = c->env_le[0]; sv.env[1] = c->env_le[1]; sv.env[2] = c->env_le[2];
    sv.fov = c->scene.opt.fov;
    sv.width = c->scene.opt.width;
    sv.height = c->scene.opt.height;
    sv.div_w = udiv_make((uint32_t)std::max(sv.width, 1));
    sv.div_8w = udiv_make(8u * (uint32_t)std::max(sv.width, 1));
    sv.div_wh = udiv_make((uint32_t)std::max(sv.width * sv.height, 1));
    sv.maxdepth = c->scene.opt.maxdepth;
    sv.mindepth = c->scene.opt.mindepth;
    return sv;
}

static int persistent_blocks(Ctx *c, int block) {
    int per_cu = block <= 128 ? 8 : 4;
    return c->n_cu * per_cu;
}

// Persistent grid of the traversal kernels: trace_waves_per_cu waves per CU
// in workgroups of trace_block lanes.
int trace_blocks(const Ctx *c) {
    return c->n_cu * std::max(1, c->launch.trace_waves_per_cu * 64 / c->launch.trace_block);
}
size_t trace_threads(const Ctx *c) { return (size_t)trace_blocks(c) * c->launch.trace_block; }

static void ensure_wf(Ctx *c, Wavefront &w, size_t cap, int width, int depth) {
    int need = (width - 1) * (depth + 1) + 2;
    size_t ovf = need > AUR_STACK_LDS ? (size_t)(need - AUR_STACK_LDS) : 0;
    if (cap <= w.cap && ovf <= w.stack_ovf_per_thread && w.counts.p &&
        w.stack_ovf.bytes >= 2 * sizeof(int) * trace_threads(c) * std::max<size_t>(w.stack_ovf_per_thread, 1) &&
        w.partials.bytes >= sizeof(unsigned long long) * trace_blocks(c))
        return;
    cap = std::max(cap, w.cap);
    // a failed allocation leaves some buffers released: forget the capacity
    // so that the next call allocates them all again
    w.cap = 0;
    for (int k = 0; k < 2; ++k) {
        w.ray[k].reserve(sizeof(float4) * 2 * cap);
        w.T[k].reserve(sizeof(float4) * cap);
        w.La[k].reserve(sizeof(float4) * cap);
        w.Lb[k].reserve(sizeof(float4) * cap);
        w.meta[k].reserve(sizeof(int4) * cap);
    }
    w.hit.reserve(sizeof(float4) * cap);
    w.sray.reserve(sizeof(float4) * 3 * 2 * cap);
    w.finLa.reserve(sizeof(float4) * cap);
    w.finLb.reserve(sizeof(float4) * cap);
    w.finMeta.reserve(sizeof(int4) * cap);
    w.out.reserve(sizeof(float4) * cap);
    w.counts.reserve(sizeof(int) * 16);
    w.stats.reserve(sizeof(unsigned long long) * 8);
    w.stack_ovf_per_thread = std::max(ovf, w.stack_ovf_per_thread);
    // two regions: extension and shadow traversals may run at the same time
    w.stack_ovf.reserve(2 * sizeof(int) * trace_threads(c) * std::max<size_t>(w.stack_ovf_per_thread, 1));
    w.partials.reserve(sizeof(unsigned long long) * trace_blocks(c));
    w.cap = cap;
}

void ensure_wavefront_for(Ctx *c, size_t cap, int width, int depth) { ensure_wf(c, c->wf, cap, width, depth); }
void ensure_wavefront_on(Ctx *c, Wavefront &w, size_t cap, int width, int depth) { ensure_wf(c, w, cap, width, depth); }

static void ensure_wavefront(Ctx *c, size_t cap) { ensure_wavefront_for(c, cap, c->bvh.width, c->bvh.depth); }

static WfView wf_view_of(Ctx *c, Wavefront &w, const PathSource &src) {
    WfView v;
    for (int k = 0; k < 2; ++k) {
        v.ray[k] = w.ray[k].as<float4>();
        v.T[k] = w.T[k].as<float4>();
        v.La[k] = w.La[k].as<float4>();
        v.Lb[k] = w.Lb[k].as<float4>();
        v.meta[k] = w.meta[k].as<int4>();
    }
    v.hit = w.hit.as<float4>();
    v.sray = w.sray.as<float4>();
    v.finLa = w.finLa.as<float4>();
    v.finLb = w.finLb.as<float4>();
    v.finMeta = w.finMeta.as<int4>();
    v.out = w.out.as<float4>();
    v.counts = w.counts.as<int>();
    v.stats = w.stats.as<unsigned long long>();
    v.partials = w.partials.as<unsigned long long>();
    v.ovf = w.stack_ovf.as<int>();
    v.ovf_per_thread = (int)w.stack_ovf_per_thread;
    v.sovf = v.ovf + trace_threads(c) * std::max<size_t>(w.stack_ovf_per_thread, 1);
    v.src = src;
    v.div_per = udiv_make((uint32_t)std::max(src.ls, 1) * (uint32_t)std::max(c->scene.opt.width * c->scene.opt.height, 1));
    v.seed = c->seed;
    v.nslots = 2 + AUR_SLOTS_PER_VERTEX * c->scene.opt.maxdepth;
    return v;
}

WfView wf_view(Ctx *c, const PathSource &src) { return wf_view_of(c, c->wf, src); }
WfView wf_view_on(Ctx *c, Wavefront &w, const PathSource &src) { return wf_view_of(c, w, src); }

// The streams and events of one path pipeline.
struct Lane {
    hipStream_t main, side;
    hipEvent_t fork, shaded, done;
};

// One batch of n paths on lane L: generate, then maxdepth+1 rounds of
// extend/shade/shadow/finish.  Results land in the lane's out (by path id).
static void run_batch(Ctx *c, const Lane &L, const SceneView &sv, WfView wv, int n) {
    // hash-sampled paths recompute their camera rays where they are used
    const bool cam_inline = sv.qmc == nullptr;
    if (cam_inline) {
        TIMED_ON(c, KT_GENERATE, L.main, hipLaunchKernelGGL(k_start, dim3(1), dim3(1), 0, L.main, wv, n));
    } else {
        const int gen_blocks = std::min((n + 255) / 256, persistent_blocks(c, 256));
        TIMED_ON(c, KT_GENERATE, L.main,
                 hipLaunchKernelGGL(k_generate, dim3(gen_blocks), dim3(256), 0, L.main, sv, wv, n));
    }
    const int tb = trace_blocks(c), sb = persistent_blocks(c, 256);
    const bool tex = shade_needs_tex(c);
    // A bounce's shadow rays and finished paths go to the side stream, so
    // they run beside the next bounce's extension rays, which read and write
    // none of their buffers (the shadow traversal has its own stack overflow
    // region); the next shade waits for them, and the lane's stream for the
    // last of them.  With kernel timing on, everything stays on the
    // context's stream so that each kernel's time is its own.
    const bool overlap = !c->timer.on;
    hipStream_t s2 = overlap ? L.side : L.main;
    if (overlap) {
        HIP_CHECK(hipEventRecord(L.fork, L.main));
        HIP_CHECK(hipStreamWaitEvent(s2, L.fork, 0));
    }
    int q = 0;
    for (int it = 0; it <= sv.maxdepth; ++it) {
        const bool cam = cam_inline && it == 0;
#define AUR_EXT(w_, b_)                                                                                  \
    do {                                                                                                 \
        if (cam) hipLaunchKernelGGL((k_extend<w_, b_, true>), dim3(tb), dim3(b_), 0, L.main, sv, wv, q); \
        else hipLaunchKernelGGL((k_extend<w_, b_>), dim3(tb), dim3(b_), 0, L.main, sv, wv, q);          \
    } while (0)
#define AUR_SHD(w_, b_) hipLaunchKernelGGL((k_shadow<w_, b_>), dim3(tb), dim3(b_), 0, s2, sv, wv, 1 - q)
        TIMED_ON(c, KT_EXTEND, L.main, AUR_TRAV_DISPATCH(c->bvh.width, c->launch.trace_block, AUR_EXT));
        if (overlap && it > 0) HIP_CHECK(hipStreamWaitEvent(L.main, L.done, 0));
        const int last = it == sv.maxdepth;
        // the camera bounce is shaded in place: a frame's camera rays mostly
        // hit (at 1080p, listing them cost 1 ms of shading per frame)
        const int list_tiles = it == 0 ? 0x7fffffff : c->launch.shade_list_tiles;
#define AUR_SHADE(t_, c_) \
    hipLaunchKernelGGL((k_shade<t_, c_>), dim3(sb), dim3(256), 0, L.main, sv, wv, q, last, list_tiles)
        if (tex && cam) TIMED_ON(c, KT_SHADE, L.main, AUR_SHADE(true, true));
        else if (tex) TIMED_ON(c, KT_SHADE, L.main, AUR_SHADE(true, false));
        else if (cam) TIMED_ON(c, KT_SHADE, L.main, AUR_SHADE(false, true));
        else TIMED_ON(c, KT_SHADE, L.main, AUR_SHADE(false, false));
#undef AUR_SHADE
        if (overlap) {
            HIP_CHECK(hipEventRecord(L.shaded, L.main));
            HIP_CHECK(hipStreamWaitEvent(s2, L.shaded, 0));
        }
        TIMED_ON(c, KT_SHADOW, s2, AUR_TRAV_DISPATCH(c->bvh.width, c->launch.trace_block, AUR_SHD));
#undef AUR_EXT
#undef AUR_SHD
        TIMED_ON(c, KT_FINISH, s2, hipLaunchKernelGGL(k_finish, dim3(sb), dim3(256), 0, s2, sv, wv));
        hipLaunchKernelGGL(k_clear_counts, dim3(1), dim3(64), 0, s2, wv.counts, q);
        if (overlap) HIP_CHECK(hipEventRecord(L.done, s2));
        q = 1 - q;
    }
    if (overlap) HIP_CHECK(hipStreamWaitEvent(L.main, L.done, 0));
    HIP_CHECK(hipGetLastError());
}

void ensure_lane_streams(Ctx *c) {
    if (c->lane_stream[0]) return;
    for (hipStream_t &ls : c->lane_stream) HIP_CHECK(hipStreamCreateWithFlags(&ls, hipStreamNonBlocking));
    for (int l = 0; l < 2; ++l) {
        HIP_CHECK(hipEventCreateWithFlags(&c->lane_fork[l], hipEventDisableTiming));
        HIP_CHECK(hipEventCreateWithFlags(&c->lane_shaded[l], hipEventDisableTiming));
        HIP_CHECK(hipEventCreateWithFlags(&c->lane_done[l], hipEventDisableTiming));
        HIP_CHECK(hipEventCreateWithFlags(&c->ev_resolved[l], hipEventDisableTiming));
    }
    HIP_CHECK(hipEventCreateWithFlags(&c->ev_start, hipEventDisableTiming));
    HIP_CHECK(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
}

// One chunk of run_paths: whole passes (mode 0) or a range of paths.
struct Chunk {
    size_t done, n;
    int first_pass, passes;
};

void run_paths(Ctx *c, const PathSource &src, size_t n_paths, float4 *out_rgba, aur_stats *stats) {
    if (!c->built) throw StateError("render before aur_build_accel");
    const size_t wh = (size_t)c->scene.opt.width * c->scene.opt.height;
    const int ls = std::max(1, src.ls);
    // Chunks of up to pool_paths paths; with kernel timing off, renders of at
    // least kTwoLaneMin paths alternate their chunks between two lanes of half
    // the pool each, so one chunk's shading and latency-bound last bounces run
    // beside the other's traversal (two contexts rendering halves of the 1080p
    // frame side by side measured 2060-2160 Mpaths/s against 1870-1900 for
    // one; three or four, 1880-1980).  Below about a million paths per lane
    // the second lane costs more than it overlaps (box room 256x256: 4 passes
    // 330 against 500 Mpaths/s, 16 passes 970 against 997, 64 passes 1583
    // against 1312).  mode 0 traces whole passes per chunk (a pass of ls
    // paths per pixel fits).
    constexpr size_t kTwoLaneMin = size_t(1) << 21;
    const bool two = !c->timer.on && n_paths >= kTwoLaneMin;
    const size_t pool = c->launch.pool_paths;
    size_t want = std::min<size_t>(n_paths, two ? std::max<size_t>(pool / 2, 1) : pool);
    if (two) want = std::min(want, (n_paths + 1) / 2);
    const size_t cap = src.mode == 0 ? std::max<size_t>(wh * ls, want) : std::max<size_t>(want, 1);
    if (cap > kMaxPoolPaths) throw CapacityError("one pass needs more path slots than the queues can index");
    std::vector<Chunk> chunks;
    {
        size_t done = 0;
        int pass = src.first_pass;
        while (done < n_paths) {
            Chunk ch{done, std::min(cap, n_paths - done), pass, 0};
            if (src.mode == 0) {
                // as many whole passes as fit
                size_t m = 0;
                while (done + m < n_paths) {
                    const size_t np = wh * (size_t)(pass + ch.passes == 0 ? 1 : ls);
                    if (ch.passes > 0 && m + np > cap) break;
                    m += np;
                    ++ch.passes;
                }
                if (done + m > n_paths) throw std::invalid_argument("path count is not a whole number of passes");
                ch.n = m;
                pass += ch.passes;
            }
            chunks.push_back(ch);
            done += ch.n;
        }
    }
    const int n_lanes = two && chunks.size() > 1 ? 2 : 1;
    Wavefront *wfs[2] = {&c->wf, &c->wf2};
    for (int l = 0; l < n_lanes; ++l) ensure_wf(c, *wfs[l], cap, c->bvh.width, c->bvh.depth);
    // with kernel timing on, everything runs on the context's stream
    ensure_lane_streams(c);
    const bool lanes_on = n_lanes == 2;
    const Lane lanes[2] = {
        {c->stream, c->timer.on ? c->stream : c->lane_stream[0], c->lane_fork[0], c->lane_shaded[0], c->lane_done[0]},
        {c->lane_stream[1], c->lane_stream[2], c->lane_fork[1], c->lane_shaded[1], c->lane_done[1]},
    };
    SceneView sv = scene_view(c);
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    HIP_CHECK(hipEventRecord(e0, c->stream));
    for (int l = 0; l < n_lanes; ++l) {
        hipLaunchKernelGGL(k_reset_all, dim3(1), dim3(64), 0, c->stream, wfs[l]->counts.as<int>(),
                           wfs[l]->stats.as<unsigned long long>(), 1);
        fill_zero(c, wfs[l]->partials.p, sizeof(unsigned long long) * trace_blocks(c));
    }
    if (lanes_on) {
        HIP_CHECK(hipEventRecord(c->ev_start, c->stream));
        HIP_CHECK(hipStreamWaitEvent(lanes[1].main, c->ev_start, 0));
    }
    for (size_t k = 0; k < chunks.size(); ++k) {
        const Chunk &ch = chunks[k];
        const int l = (int)(k % (size_t)n_lanes);
        const Lane &L = lanes[l];
        PathSource s = src;
        s.ls = ls;
        if (src.mode == 0) {
            s.first_pass = ch.first_pass;
        } else {
            s.pixel = src.pixel + ch.done;
            s.sample = src.sample + ch.done;
            s.slots = src.slots ? src.slots + ch.done * (size_t)(2 + AUR_SLOTS_PER_VERTEX * c->scene.opt.maxdepth) : nullptr;
        }
        if (src.adj_rec) s.rec_base = src.rec_base + (long long)ch.done;
        WfView wv = wf_view_of(c, *wfs[l], s);
        hipLaunchKernelGGL(k_reset_all, dim3(1), dim3(64), 0, L.main, wv.counts, wv.stats, 0);
        run_batch(c, L, sv, wv, (int)ch.n);
        if (src.adj_rec && src.rec_only) {
            // records only
        } else if (src.mode == 0) {
            // passes go into the framebuffer in order: after the previous
            // chunk's resolve, which ran on the other lane
            if (n_lanes == 2 && k > 0) HIP_CHECK(hipStreamWaitEvent(L.main, c->ev_resolved[1 - l], 0));
            int blocks = std::min<int>((int)((wh + 255) / 256), persistent_blocks(c, 256));
            TIMED_ON(c, KT_RESOLVE, L.main,
                     hipLaunchKernelGGL(k_resolve, dim3(blocks), dim3(256), 0, L.main, wv.out, (int)wh,
                                        s.first_pass, ch.passes, ls, out_rgba, src.display,
                                        (int)(src.display_base + (s.first_pass - src.first_pass))));
            if (n_lanes == 2) HIP_CHECK(hipEventRecord(c->ev_resolved[l], L.main));
        } else {
            HIP_CHECK(hipMemcpyAsync(out_rgba + ch.done, wv.out, sizeof(float4) * ch.n, hipMemcpyDeviceToDevice, L.main));
        }
    }
    if (lanes_on) {
        // the context's stream (lane 0) waits for lane 1
        HIP_CHECK(hipEventRecord(c->ev_join, lanes[1].main));
        HIP_CHECK(hipStreamWaitEvent(c->stream, c->ev_join, 0));
    }
    for (int l = 0; l < n_lanes; ++l)
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, c->stream, wfs[l]->partials.as<unsigned long long>(),
                           trace_blocks(c), wfs[l]->stats.as<unsigned long long>());
    HIP_CHECK(hipEventRecord(e1, c->stream));
    HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    timer_collect(c);
    if (stats) {
        unsigned long long tot[8] = {0};
        for (int l = 0; l < n_lanes; ++l) {
            unsigned long long s8[8];
            HIP_CHECK(hipMemcpy(s8, wfs[l]->stats.p, sizeof s8, hipMemcpyDeviceToHost));
            for (int j = 0; j < 8; ++j) tot[j] += s8[j];
        }
        stats->paths = (int64_t)tot[S_PATHS];
        stats->ext_rays = (int64_t)tot[S_EXT];
        stats->hit_rays = (int64_t)tot[S_HIT];
        stats->shadow_rays = (int64_t)tot[S_SHADOW];
        stats->ms = ms;
    }
}

// ---------------------------------------------------------------------------
// aur_trace: batch closest-hit / any-hit queries for the parity tests
// ---------------------------------------------------------------------------
template <int W, int B>
__global__ void __launch_bounds__(B) k_trace(SceneView sv, const aur_ray *__restrict__ rays, int n, aur_hit *__restrict__ hits,
                                             int any_hit, int *ovf, int ovf_per_thread) {
    __shared__ int s_stack[AUR_STACK_LDS * B];
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    TravStack st;
    st.init(s_stack, B, ovf, ovf_per_thread);
    const WNode<W> *nodes = static_cast<const WNode<W> *>(sv.nodes);
    for (int i = gtid; i < n; i += gridDim.x * blockDim.x) {
        aur_ray r = rays[i];
        TravHit h;
        if (any_hit) traverse<W, true>(nodes, sv.tris, mk3(r.o[0], r.o[1], r.o[2]), mk3(r.d[0], r.d[1], r.d[2]), r.tmin, r.tmax, st, h);
        else traverse<W, false>(nodes, sv.tris, mk3(r.o[0], r.o[1], r.o[2]), mk3(r.d[0], r.d[1], r.d[2]), r.tmin, r.tmax, st, h);
        aur_hit o;
        o.t = h.t; o.u = h.u; o.v = h.v; o.prim = h.prim;
        hits[i] = o;
    }
}

// AUR_TRACE_EXTEND: the rays go through the renderer's own closest-hit
// kernel.  They are written to extension queue 0 in its layout, k_extend runs
// with the launch shape of a render (persistent grid, per-lane refill), and
// its hit records are read back.
__global__ void k_rays_to_queue(const aur_ray *__restrict__ rays, int n, WfView wv) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const aur_ray r = rays[i];
        wv.ray[0][2 * i] = make_float4(r.o[0], r.o[1], r.o[2], r.tmin);
        wv.ray[0][2 * i + 1] = make_float4(r.d[0], r.d[1], r.d[2], r.tmax);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) wv.counts[C_Q0] = n;
}

__global__ void k_queue_hits(WfView wv, int n, aur_hit *__restrict__ hits) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float4 h = wv.hit[i];
        aur_hit o;
        o.t = h.x; o.u = h.y; o.v = h.z; o.prim = __float_as_int(h.w);
        hits[i] = o;
    }
}

void trace_rays(Ctx *c, const aur_ray *d_rays, int n, aur_hit *d_hits, int mode) {
    if (!c->built) throw StateError("trace before aur_build_accel");
    if (mode == AUR_TRACE_EXTEND) {
        ensure_wavefront(c, std::max(n, 1));
        SceneView sv = scene_view(c);
        PathSource src;
        src.mode = 1;
        src.first_pass = 0;
        src.pixel = src.sample = nullptr;
        src.slots = nullptr;
        WfView wv = wf_view(c, src);
        const int blocks = std::max(1, std::min((n + 255) / 256, persistent_blocks(c, 256)));
        hipLaunchKernelGGL(k_reset_all, dim3(1), dim3(64), 0, c->stream, wv.counts, wv.stats, 1);
        hipLaunchKernelGGL(k_rays_to_queue, dim3(blocks), dim3(256), 0, c->stream, d_rays, n, wv);
        const int tb = trace_blocks(c);
#define AUR_EXT(w_, b_) hipLaunchKernelGGL((k_extend<w_, b_>), dim3(tb), dim3(b_), 0, c->stream, sv, wv, 0)
        AUR_TRAV_DISPATCH(c->bvh.width, c->launch.trace_block, AUR_EXT);
#undef AUR_EXT
        hipLaunchKernelGGL(k_queue_hits, dim3(blocks), dim3(256), 0, c->stream, wv, n, d_hits);
        // the extend kernel's per-workgroup hit partials start the next render at zero
        fill_zero(c, c->wf.partials.p, sizeof(unsigned long long) * trace_blocks(c));
        HIP_CHECK(hipGetLastError());
        return;
    }
    const int any_hit = mode == AUR_TRACE_ANY ? 1 : 0;
    ensure_wavefront(c, 1);
    SceneView sv = scene_view(c);
    const int B = c->launch.trace_block;
    int blocks = std::min((n + B - 1) / B, trace_blocks(c));
    if (blocks < 1) return;
#define AUR_TRC(w_, b_)                                                                                      \
    hipLaunchKernelGGL((k_trace<w_, b_>), dim3(blocks), dim3(b_), 0, c->stream, sv, d_rays, n, d_hits, any_hit, \
                       c->wf.stack_ovf.as<int>(), (int)c->wf.stack_ovf_per_thread)
    AUR_TRAV_DISPATCH(c->bvh.width, B, AUR_TRC);
#undef AUR_TRC
    HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// aur_trace_noalloc: the ray queries of aur_trace for callers that cannot
// allocate device memory (code under stream capture, serving loops).  The
// only scratch of these kernels is the traversal stack's overflow area; here
// it is the file-scope array below: AUR_NOALLOC_THREADS lanes (the fixed
// grid of these launches) times AUR_NOALLOC_OVF entries, 16 MiB per GPU, held
// until the process ends.  Every call on one GPU shares it: calls that overlap
// on that GPU (two streams, two host threads) race on it.  Nothing in it has
// to start at zero -- a lane writes each overflow entry before it reads it.
// ---------------------------------------------------------------------------
#define AUR_NOALLOC_THREADS 32768
#define AUR_NOALLOC_OVF 128
static __device__ __align__(256) int scratch[AUR_NOALLOC_THREADS * AUR_NOALLOC_OVF];

// the array's device address, one per GPU (the host-side name is not one)
static std::mutex g_scratch_mu;
static std::vector<int *> g_scratch_ptr;

void noalloc_init(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
    if (g_scratch_ptr[device]) return;
    // the caller made `device` current
    void *p = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw HipError("hipGetSymbolAddress returned a null address for the traversal scratch");
    g_scratch_ptr[device] = static_cast<int *>(p);
}

static int *noalloc_scratch(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= (int)g_scratch_ptr.size() || !g_scratch_ptr[device])
        throw StateError("the traversal scratch of this device was not set up (aur_create)");
    return g_scratch_ptr[device];
}

void trace_rays_noalloc(Ctx *c, const aur_ray *d_rays, int n, aur_hit *d_hits, int mode) {
    if (!c->built) throw StateError("trace before aur_build_accel");
    if (mode != AUR_TRACE_CLOSEST && mode != AUR_TRACE_ANY)
        throw std::invalid_argument("aur_trace_noalloc traces AUR_TRACE_CLOSEST or AUR_TRACE_ANY only");
    // stack entries one ray can need: W - 1 per level below the root, plus two
    const int need = (c->bvh.width - 1) * (c->bvh.depth + 1) + 2;
    if (need - AUR_STACK_LDS > AUR_NOALLOC_OVF)
        throw CapacityError("BVH too deep for the fixed traversal scratch (" + std::to_string(need - AUR_STACK_LDS) +
                            " overflow entries per ray, at most " + std::to_string(AUR_NOALLOC_OVF) + ")");
    int *ovf = noalloc_scratch(c->device);
    SceneView sv = scene_view(c);
    const int B = c->launch.trace_block;
    const int blocks = std::min((n + B - 1) / B, AUR_NOALLOC_THREADS / B);
    if (blocks < 1) return;
#define AUR_TRC(w_, b_)                                                                                      \
    hipLaunchKernelGGL((k_trace<w_, b_>), dim3(blocks), dim3(b_), 0, c->stream, sv, d_rays, n, d_hits,        \
                       mode == AUR_TRACE_ANY ? 1 : 0, ovf, AUR_NOALLOC_OVF)
    AUR_TRAV_DISPATCH(c->bvh.width, B, AUR_TRC);
#undef AUR_TRC
    HIP_CHECK(hipGetLastError());
}

}  // namespace aur

#ifdef AUR_TRAV_STATS
// measurement builds only: copies the traversal counters (extend, then
// shadow: visits, leaves, triangles, unused) and clears them
extern "C" int aur_debug_trav_stats(unsigned long long *out8) {
    unsigned long long z[8] = {0};
    if (hipMemcpyFromSymbol(out8, HIP_SYMBOL(aur::g_trav_stats), sizeof z) != hipSuccess) return -4;
    if (hipMemcpyToSymbol(HIP_SYMBOL(aur::g_trav_stats), z, sizeof z) != hipSuccess) return -4;
    return 0;
}
#endif
