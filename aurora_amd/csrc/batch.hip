// batch.hip -- many scenes in one persistent wavefront launch.
//
// The reference renders one scene per process: `Renderer::renderImage`
// (src/frontend/renderer.cpp:201-271) runs its passes with one TBB
// parallel_for each.  Here a list of jobs (each a built scene, a pass range
// and its own framebuffer) shares one pool of path slots.  The jobs' paths
// form one work sequence in launch order (shortest estimated job first by
// default): work item w of job j is pass (w - start_j) / (W*H) and pixel
// tiled_pixel((w - start_j) mod (W*H)), the order of a single-scene launch.
// Per iteration:
//
//   plan    -- one thread: as many new paths as the pool has room for (the
//              previous iteration's finished paths left it) come off the
//              work cursor
//   regen   -- the new paths' camera rays and state are appended to the
//              extension queue, coalesced, no atomics
//   extend, shade, shadow -- the single-scene kernels' bodies; each wave
//              loops once per distinct job among its lanes (almost always
//              one) with that job's scene record read wave-uniformly
//   finish  -- firefly clamp; the path's RGBA and its packed ray counts and
//              finishing iteration go to ring buffers at its work index
//
// so the GPU stays full until the last job drains, whatever the jobs' sizes,
// resolutions, sample counts and depths.  An iteration is a fixed sequence of
// launches whose counts live in device memory; two iterations (the extension
// queues swap) are captured once as a hipGraph and replayed.  Every few
// iterations the host reads the cursor history: work issued at least
// maxdepth + 1 iterations ago has finished, and that prefix of the ring is
// added into the jobs' framebuffers pixel by pixel in pass order -- the same
// float additions, in the same order, as a single-scene render of the same
// passes, so each job's framebuffer is bit-identical to it -- and summed into
// per-job ray counts and the job's completion time (the wall clock stamped at
// the start of the iteration after its last path finished).
#include "wavefront_common.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

namespace aur {

struct JobDev {
    SceneView sv;
    float4 *accum;       // the job's framebuffer: sum of per-pass rgba
    ushort4 *display;    // and its FP16 running mean (display_update)
    long long display_base;  // passes appended to display before this launch
    uint32_t seed;
    int first_pass;
    int ls;              // light samples: paths per pixel in passes after pass 0
    long long start;     // first work index of the job in its pipeline's work sequence
    long long n_paths;
    int root;            // the job's root node in the launch's node arena
};

enum { J_EXT = 0, J_HIT = 1, J_SHADOW = 2, J_PATHS = 3, J_STRIDE = 4 };

// A job's work is a sequence of blocks of w*h paths: pass 0 (when the job
// starts there) is one block, every later pass ls blocks (one per light
// sample).  Block -> pass offset from the job's first pass, and light sample.
__device__ inline void job_block(const JobDev &jd, int blk, int *pass_rel, int *sub) {
    int b = blk, p = 0;
    if (jd.first_pass == 0) {
        if (b == 0) { *pass_rel = 0; *sub = 0; return; }
        b -= 1;
        p = 1;
    }
    const int k = b / jd.ls;
    *pass_rel = p + k;
    *sub = b - k * jd.ls;
}
// control block: cursor, iterations done, the current iteration's plan, the
// cursor after each of the last BATCH_HIST iterations and the wall clock at
// the start of each of them
#define BATCH_HIST 64
enum { CTL_CURSOR = 0, CTL_ITER = 1, CTL_LIVE = 2, CTL_NEW = 3, CTL_BASE = 4, CTL_HIST = 8,
       CTL_TIME = CTL_HIST + BATCH_HIST, CTL_SIZE = CTL_TIME + BATCH_HIST };
// per finished path, next to its rgba: extension and hit ray counts (at most
// maxdepth + 1 <= 201 each: 8 bits), shadow rays (at most 2 maxdepth <= 400:
// 9 bits) and the iteration it finished in (mod BATCH_HIST: 6 bits)
__device__ inline unsigned pack_counts(int ext, int hit, int shadow, long long iter) {
    return (unsigned)min(ext, 255) | ((unsigned)min(hit, 255) << 8) | ((unsigned)min(shadow, 511) << 16) |
           ((unsigned)(iter % BATCH_HIST) << 25);
}

struct BatchView {
    const JobDev *jobs;      // in launch order
    const void *nodes;       // every job's BVH in one node arena (WNode<W>)
    const float4 *tris;      // and their triangle records in one array
    const long long *starts; // jobs[k].start, for the binary search
    int n_jobs;
    long long total;         // paths over all jobs
    long long *ctl;
    int pool;                // paths in flight at most
    float4 *ring;            // finished paths' rgba by work index
    unsigned *ring_cnt;      // finished paths' packed counts by work index
    long long ring_mask;
    unsigned long long *t_end;  // per job: wall clock after its last path finished
    long long *t0;           // wall clock at the start of the launch
    unsigned long long *jstats;  // per job: J_STRIDE counters
};

struct ResolveSeg {
    long long la, lb;   // work range of the job, relative to its start
    long long tbase;    // first thread of the segment
    int job, pad;
};

__global__ void k_binit(BatchView bv, WfView wv) {
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const int stride = gridDim.x * blockDim.x;
    for (int j = gtid; j < bv.n_jobs; j += stride) {
        bv.t_end[j] = 0;
        for (int k = 0; k < J_STRIDE; ++k) bv.jstats[(size_t)j * J_STRIDE + k] = 0;
    }
    for (int k = gtid; k < CTL_SIZE; k += stride) bv.ctl[k] = 0;
    if (gtid == 0) {
        for (int k = 0; k < 16; ++k) wv.counts[k] = 0;
        for (int k = 0; k < 8; ++k) wv.stats[k] = 0;
        *bv.t0 = wall_clock64();
    }
}

// The pool's free places (the paths that finished in the previous iteration)
// take new paths while work remains.  Also starts the iteration's shadow and
// finish lists empty (the previous finish list has been written out).
__global__ void k_plan(BatchView bv, WfView wv, int q) {
    if (threadIdx.x != 0) return;
    const long long live = wv.counts[q];
    const long long cur = bv.ctl[CTL_CURSOR];
    long long n_new = bv.total - cur;
    if (n_new > bv.pool - live) n_new = bv.pool - live;
    if (n_new < 0) n_new = 0;
    bv.ctl[CTL_LIVE] = live;
    bv.ctl[CTL_NEW] = n_new;
    bv.ctl[CTL_BASE] = cur;
    bv.ctl[CTL_CURSOR] = cur + n_new;
    const long long it = bv.ctl[CTL_ITER];
    bv.ctl[CTL_HIST + (int)(it % BATCH_HIST)] = cur + n_new;
    bv.ctl[CTL_TIME + (int)(it % BATCH_HIST)] = wall_clock64();
    bv.ctl[CTL_ITER] = it + 1;
    wv.counts[q] = (int)(live + n_new);
    wv.counts[C_FINISH] = 0;
    wv.counts[C_SHADOW] = 0;
    wv.stats[S_PATHS] += (unsigned long long)n_new;
}

__global__ void k_regen(BatchView bv, WfView wv, int q) {
    const int live = (int)bv.ctl[CTL_LIVE];
    const int n_new = (int)bv.ctl[CTL_NEW];
    const long long base = bv.ctl[CTL_BASE];
    const int stride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_new; i += stride) {
        const long long w = base + i;
        int a = 0, b = bv.n_jobs - 1;  // last job whose start <= w
        while (a < b) {
            int mid = (a + b + 1) >> 1;
            if (bv.starts[mid] <= w) a = mid;
            else b = mid - 1;
        }
        const JobDev &jd = bv.jobs[a];
        const long long local = w - jd.start;
        const int wh = jd.sv.width * jd.sv.height;
        const int blk = (int)(local / wh);
        const int l = (int)(local - (long long)blk * wh);
        const int pixel = tiled_pixel(jd.sv, l);
        int pass, sub;
        job_block(jd, blk, &pass, &sub);
        const PathRng rng = make_path_rng(jd.sv, nullptr, jd.seed, pixel, (uint32_t)(jd.first_pass + pass), sub, jd.ls);
        const f3 d = camera_dir(jd.sv, pixel, rng.u(0), rng.u(1));
        const int pos = live + i;
        wv.ray[q][2 * pos] = make_float4(0.f, 0.f, 0.f, 0.1f);
        wv.ray[q][2 * pos + 1] = make_float4(d.x, d.y, d.z, 10000000.f);
        wv.T[q][pos] = make_float4(1.f, 1.f, 1.f, 1.f);
        wv.La[q][pos] = make_float4(0.f, 0.f, 0.f, 0.f);
        wv.Lb[q][pos] = make_float4(0.f, 0.f, 0.f, 0.f);
        wv.meta[q][pos] = make_int4(RT_CAMERA, l, blk, a);
    }
}

// Stamps the start of the iteration that would follow the last one, so paths
// that finished in the last iteration get a completion time too.
__global__ void k_bstamp(BatchView bv) {
    if (threadIdx.x == 0) bv.ctl[CTL_TIME + (int)(bv.ctl[CTL_ITER] % BATCH_HIST)] = wall_clock64();
}

__global__ void k_bclear_queue(int *counts, int q) {
    if (threadIdx.x == 0) counts[q] = 0;
}

// The node arena of a batched launch: every job's nodes with inner child
// indices moved by the job's node offset and leaf triangle ranges by its
// triangle offset, so one pair of pointers (wave-uniform, as in the
// single-scene kernels) serves every job and a ray only names its root.
template <int W>
__global__ void k_arena_nodes(const WNode<W> *__restrict__ src, int n, int node_off, int tri_off,
                              WNode<W> *__restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    WNode<W> nd = src[i];
#pragma unroll
    for (int k = 0; k < W; ++k) {
        const int c = nd.child[k];
        if (c == 0x7fffffff) continue;  // empty slot
        if (c >= 0) {
            nd.child[k] = c + node_off;
        } else {
            const int enc = ~c;
            nd.child[k] = ~((((enc >> 4) + tri_off) << 4) | (enc & 15));
        }
    }
    nd.pad[0] += node_off;  // first inner child
    dst[node_off + i] = nd;
}

// Batched extension: the single-scene kernel's refilled lanes on the node
// arena; a ray starts at its job's root.
template <int W, int B>
__global__ void __launch_bounds__(B, AUR_TRAV_WAVES(W)) k_bextend(const JobDev *__restrict__ jobs,
                                                                   const WNode<W> *__restrict__ nodes,
                                                                   const float4 *__restrict__ tris, WfView wv, int q) {
    __shared__ int s_stack[AUR_STACK_LDS * B];
    const int n = wv.counts[q];
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const int stride = gridDim.x * blockDim.x;
    TravStack st;
    st.init(s_stack, B, wv.ovf, wv.ovf_per_thread);
    traverse_stream<W, false, true>(
        nodes, tris, st, gtid, n, stride, AUR_REFILL_LANES,
        [&](int i, f3 &o, f3 &d, float &tmin, float &tmax, int &root) {
            root = jobs[wv.meta[q][i].w].root;
            const float4 a = wv.ray[q][2 * i], b = wv.ray[q][2 * i + 1];
            o = mk3(a.x, a.y, a.z); d = mk3(b.x, b.y, b.z); tmin = a.w; tmax = b.w;
        },
        [&](int i, const TravHit &h) { wv.hit[i] = make_float4(h.t, h.u, h.v, __int_as_float(h.prim)); });
}

// The kernels below run each wave once per distinct job among its lanes
// (almost always one): the job index is made wave-uniform with
// readfirstlane, so the scene record is read with scalar loads and the
// traversal and shading code sees uniform scene pointers, as in the
// single-scene kernels.

// One finished path's ring entry: clamped radiance and its packed counts.
__device__ inline void bfinish_path(const BatchView &bv, long long iter, const float4 a, const float4 b,
                                    const int4 m) {
    const f3 L = firefly_clamp(mk3(a.x + b.x, a.y + b.y, a.z + b.z), m.x);
    const JobDev &jd = bv.jobs[m.w];
    const long long w = (jd.start + (long long)m.z * (jd.sv.width * jd.sv.height) + m.y) & bv.ring_mask;
    // a camera miss is marked by alpha + 2 (the resolve takes it once); a
    // pass's later light samples do not count the shared camera ray
    const bool cam_miss = (m.x & F_MISS) && path_ext_rays(m.x) == 1;
    int pass, sub;
    job_block(jd, m.z, &pass, &sub);
    const int dup = sub > 0 ? 1 : 0;
    bv.ring[w] = make_float4(L.x, L.y, L.z, a.w + (cam_miss ? 2.f : 0.f));
    bv.ring_cnt[w] = pack_counts(path_ext_rays(m.x) - dup, path_hit_rays(m.x) - (dup && !cam_miss ? 1 : 0),
                                 (int)b.w, iter);
}

// The paths that finished with shadow rays of their last vertex pending
// (the others were written by k_bshade).
__global__ void __launch_bounds__(256) k_bfinish(BatchView bv, WfView wv) {
    const int n = wv.counts[C_FINISH];
    const long long iter = bv.ctl[CTL_ITER] - 1;
    const int stride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        bfinish_path(bv, iter, wv.finLa[i], wv.finLb[i], wv.finMeta[i]);
}

template <bool TEX>
__global__ void __launch_bounds__(256, TEX ? 4 : 5) k_bshade(BatchView bv, WfView wv, int q) {
    __shared__ EmitScratch es;
    const JobDev *__restrict__ jobs = bv.jobs;
    const long long iter = bv.ctl[CTL_ITER] - 1;
    const int n = wv.counts[q];
    const int stride = gridDim.x * blockDim.x;
    for (int base = blockIdx.x * blockDim.x; base < n; base += stride) {
        const int i = base + threadIdx.x;
        ShadeOut o;
        o.cont = o.finish = false;
        o.sh[0].valid = o.sh[1].valid = false;
        if (i < n) {
            const int job = wv.meta[q][i].w;
            for (;;) {
                const int j = __builtin_amdgcn_readfirstlane(job);
                if (job == j) {
                    const JobDev &jd = jobs[j];
                    shade_compute<TEX>(jd.sv, wv, q, i, [&](int4 m) {
                        int pass, sub;
                        job_block(jd, m.z, &pass, &sub);
                        return make_path_rng<TEX>(jd.sv, nullptr, jd.seed,
                                                  tiled_pixel(jd.sv, m.y),
                                                  (uint32_t)(jd.first_pass + pass), sub, jd.ls);
                    }, j, o);
                    break;
                }
            }
        }
        shade_emit<true>(wv, q, i < n, o, es,
                         [&](const ShadeOut &f) { bfinish_path(bv, iter, f.La, f.Lb, f.meta); });
    }
}

// Batched shadow rays on the node arena, from their job's root; the
// contribution record is read only for unoccluded rays.
template <int W, int B>
__global__ void __launch_bounds__(B, AUR_TRAV_WAVES(W)) k_bshadow(const JobDev *__restrict__ jobs,
                                                                   const WNode<W> *__restrict__ nodes,
                                                                   const float4 *__restrict__ tris, WfView wv, int nq) {
    __shared__ int s_stack[AUR_STACK_LDS * B];
    const int n = wv.counts[C_SHADOW];
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const int stride = gridDim.x * blockDim.x;
    TravStack st;
    st.init(s_stack, B, wv.ovf, wv.ovf_per_thread);
    for (int i = gtid; i < n; i += stride) {
        const float4 a = wv.sray[3 * i], b = wv.sray[3 * i + 1];
        TravHit h;
        traverse<W, true>(nodes, tris, mk3(a.x, a.y, a.z), mk3(b.x, b.y, b.z), AUR_RAY_BIAS, b.w, st, h,
                          jobs[__float_as_int(a.w)].root);
        if (h.prim < 0) {
            const float4 c = wv.sray[3 * i + 2];
            shadow_add(wv, nq, __float_as_int(c.w), c);
        }
    }
}

// Per-job sums over chunks of finished work: ray counts, paths, and the latest
// finish time (the wall clock at the start of the iteration after the path's
// last one).  One workgroup per chunk, one atomic per counter per chunk.
struct StatChunk {
    long long begin, end;  // work indices
    int job, pad;
};

__global__ void __launch_bounds__(256) k_bstats(BatchView bv, const StatChunk *__restrict__ chunks) {
    __shared__ unsigned long long s_sum[4][256];
    __shared__ unsigned long long s_tmax[256];
    const StatChunk ch = chunks[blockIdx.x];
    unsigned long long e = 0, h = 0, sh = 0, np = 0, tmax = 0;
    for (long long w = ch.begin + threadIdx.x; w < ch.end; w += blockDim.x) {
        const unsigned c = bv.ring_cnt[w & bv.ring_mask];
        e += c & 255u;
        h += (c >> 8) & 255u;
        sh += (c >> 16) & 511u;
        np += 1;
        const int it_next = (int)(((c >> 25) + 1) % BATCH_HIST);
        const unsigned long long t = (unsigned long long)bv.ctl[CTL_TIME + it_next];
        tmax = t > tmax ? t : tmax;
    }
    s_sum[0][threadIdx.x] = e;
    s_sum[1][threadIdx.x] = h;
    s_sum[2][threadIdx.x] = sh;
    s_sum[3][threadIdx.x] = np;
    s_tmax[threadIdx.x] = tmax;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            for (int k = 0; k < 4; ++k) s_sum[k][threadIdx.x] += s_sum[k][threadIdx.x + off];
            const unsigned long long o = s_tmax[threadIdx.x + off];
            if (o > s_tmax[threadIdx.x]) s_tmax[threadIdx.x] = o;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        unsigned long long *js = bv.jstats + (size_t)ch.job * J_STRIDE;
        atomicAdd(js + J_EXT, s_sum[0][0]);
        atomicAdd(js + J_HIT, s_sum[1][0]);
        atomicAdd(js + J_SHADOW, s_sum[2][0]);
        atomicAdd(js + J_PATHS, s_sum[3][0]);
        atomicMax(&bv.t_end[ch.job], s_tmax[0]);
    }
}

// Adds the finished ring entries of the given work ranges to the jobs'
// framebuffers: one thread per (segment, pixel), passes in increasing order.
// Adds the job-local work range [la, lb) -- whole passes -- into the job's
// framebuffer, per pixel in pass order: a pass contributes the sum of its
// light samples times 1/ls in sample order, or its first path alone where the
// camera ray missed (integrator.cpp:176-292).
__global__ void k_bresolve(BatchView bv, const ResolveSeg *__restrict__ segs, int n_segs, long long n_threads) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n_threads; t += stride) {
        int a = 0, b = n_segs - 1;
        while (a < b) {
            int mid = (a + b + 1) >> 1;
            if (segs[mid].tbase <= t) a = mid;
            else b = mid - 1;
        }
        const ResolveSeg sg = segs[a];
        const JobDev &jd = bv.jobs[sg.job];
        const int wh = jd.sv.width * jd.sv.height;
        const long long l = t - sg.tbase;
        const int pix = tiled_pixel(jd.sv, (int)l);
        float4 *px = jd.accum + pix;
        float4 s = *px;
        ushort4 d = jd.display[pix];
        const int blk1 = (int)(sg.lb / wh);
        for (int blk = (int)(sg.la / wh); blk < blk1;) {
            int pass, sub;
            job_block(jd, blk, &pass, &sub);
            const int n_sub = (jd.first_pass + pass == 0) ? 1 : jd.ls;
            const float4 v0 = bv.ring[(jd.start + (long long)blk * wh + l) & bv.ring_mask];
            float4 v;
            if (v0.w >= 2.f) {
                v = make_float4(v0.x, v0.y, v0.z, v0.w - 2.f);
            } else if (n_sub == 1) {
                v = v0;
            } else {
                const float inv = 1.f / (float)n_sub;
                v = make_float4(0.f, 0.f, 0.f, v0.w);
                for (int j = 0; j < n_sub; ++j) {
                    const float4 u = bv.ring[(jd.start + (long long)(blk + j) * wh + l) & bv.ring_mask];
                    v.x += u.x * inv; v.y += u.y * inv; v.z += u.z * inv;
                }
            }
            s.x += v.x; s.y += v.y; s.z += v.z; s.w += v.w;
            d = display_update(d, v, (int)(jd.display_base + pass + 1));
            blk += n_sub;
        }
        *px = s;
        jd.display[pix] = d;
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
// The device view of one pipeline's state; the node arena is `arena`'s.
static BatchView batch_view(BatchState &b, const BatchState &arena, int n_jobs, long long total, long long ring_mask) {
    BatchView v;
    v.jobs = b.jobs.as<JobDev>();
    v.starts = b.starts.as<long long>();
    v.n_jobs = n_jobs;
    v.total = total;
    v.ctl = b.ctl.as<long long>();
    v.ring = b.ring.as<float4>();
    v.ring_mask = ring_mask;
    v.ring_cnt = b.ring_cnt.as<unsigned>();
    v.t_end = b.t_end.as<unsigned long long>();
    v.t0 = b.t0.as<long long>();
    v.jstats = b.jstats.as<unsigned long long>();
    v.nodes = arena.arena_nodes.p;
    v.tris = arena.arena_tris.as<float4>();
    return v;
}

// One iteration over extension queue q, on stream st.
static void batch_iteration(Ctx *c, hipStream_t st, const BatchView &bv, const WfView &wv, int q, int width, bool tex) {
    const int tb = trace_blocks(c);
    const int sb = c->n_cu * 4;
    hipLaunchKernelGGL(k_plan, dim3(1), dim3(64), 0, st, bv, wv, q);
    hipLaunchKernelGGL(k_regen, dim3(sb), dim3(256), 0, st, bv, wv, q);
#define AUR_BEXT(w_, b_)                                                                                    \
    hipLaunchKernelGGL((k_bextend<w_, b_>), dim3(tb), dim3(b_), 0, st, bv.jobs,                              \
                       static_cast<const WNode<w_> *>(bv.nodes), bv.tris, wv, q)
#define AUR_BSHD(w_, b_)                                                                                    \
    hipLaunchKernelGGL((k_bshadow<w_, b_>), dim3(tb), dim3(b_), 0, st, bv.jobs,                              \
                       static_cast<const WNode<w_> *>(bv.nodes), bv.tris, wv, 1 - q)
    AUR_TRAV_DISPATCH(width, c->launch.trace_block, AUR_BEXT);
    if (tex) hipLaunchKernelGGL(k_bshade<true>, dim3(sb), dim3(256), 0, st, bv, wv, q);
    else hipLaunchKernelGGL(k_bshade<false>, dim3(sb), dim3(256), 0, st, bv, wv, q);
    AUR_TRAV_DISPATCH(width, c->launch.trace_block, AUR_BSHD);
#undef AUR_BEXT
#undef AUR_BSHD
    hipLaunchKernelGGL(k_bfinish, dim3(sb), dim3(256), 0, st, bv, wv);
    hipLaunchKernelGGL(k_bclear_queue, dim3(1), dim3(64), 0, st, wv.counts, q);
}

static double job_cost(const Ctx *j, long long paths) {
    // relative cost estimate: paths x depth x tris^0.3
    return (double)paths * (j->scene.opt.maxdepth + 1) * std::pow((double)std::max(j->n_tris, 1), 0.3);
}

// Adds ring entries [from, to) (work indices) into the framebuffers and the
// per-job statistics.
#define STAT_CHUNK 65536
// The largest pass boundary of a job at or below job-local work index x.
static long long pass_floor(long long x, long long wh, int first_pass, int ls) {
    long long blocks = x / wh;
    if (first_pass == 0) {
        if (blocks == 0) return 0;
        return (1 + ((blocks - 1) / ls) * ls) * wh;
    }
    return (blocks / ls) * ls * wh;
}

// Resolves every job's whole passes finished below global work index `to`
// (its frontier job_done moves pass by pass) and sums their statistics.
static void resolve_range(Ctx *c, BatchState &b, hipStream_t st, const BatchView &bv, const std::vector<long long> &starts,
                          const std::vector<long long> &npaths, const std::vector<int> &wh,
                          const std::vector<int> &jfirst, const std::vector<int> &jls, std::vector<long long> &job_done,
                          long long to) {
    ResolveSeg *seg = static_cast<ResolveSeg *>(b.host_segs);
    StatChunk *chk = static_cast<StatChunk *>(b.host_chunks);
    int n = 0, nc = 0;
    long long threads = 0;
    for (int r = 0; r < (int)starts.size() && starts[r] < to; ++r) {
        const long long la = job_done[r];
        const long long lraw = std::min(to, starts[r] + npaths[r]) - starts[r];
        const long long lb = lraw >= npaths[r] ? npaths[r] : pass_floor(lraw, wh[r], jfirst[r], jls[r]);
        if (la >= lb) continue;
        job_done[r] = lb;
        seg[n].la = la;
        seg[n].lb = lb;
        seg[n].tbase = threads;
        seg[n].job = r;
        seg[n].pad = 0;
        threads += wh[r];  // one thread per pixel of the job
        ++n;
        for (long long k = la; k < lb; k += STAT_CHUNK) {
            if ((size_t)nc >= b.host_chunks_cap) throw CapacityError("statistics chunk table overflow");
            chk[nc].begin = starts[r] + k;
            chk[nc].end = starts[r] + std::min(lb, k + STAT_CHUNK);
            chk[nc].job = r;
            chk[nc].pad = 0;
            ++nc;
        }
    }
    if (n == 0) return;
    HIP_CHECK(hipMemcpyAsync(b.segs.p, b.host_segs, sizeof(ResolveSeg) * n, hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemcpyAsync(b.chunks.p, b.host_chunks, sizeof(StatChunk) * nc, hipMemcpyHostToDevice, st));
    const long long blocks = std::min<long long>((threads + 255) / 256, (long long)c->n_cu * 8);
    hipLaunchKernelGGL(k_bresolve, dim3((unsigned)blocks), dim3(256), 0, st, bv, b.segs.as<ResolveSeg>(), n, threads);
    hipLaunchKernelGGL(k_bstats, dim3(nc), dim3(256), 0, st, bv, b.chunks.as<StatChunk>());
    HIP_CHECK(hipGetLastError());
}

void run_jobs(Ctx *c, Ctx *const *jobs, const int *first_pass, const int *n_passes, int n_jobs, int order, int use_graph,
              int iters_per_poll, aur_job_result *results, aur_stats *total_stats) {
    if (n_jobs <= 0) throw std::invalid_argument("no jobs");
    int width = 0, depth = 0, max_path = 0;
    bool tex = false;
    std::vector<long long> paths(n_jobs);
    for (int k = 0; k < n_jobs; ++k) {
        Ctx *j = jobs[k];
        if (!j->built) throw StateError("job " + std::to_string(k) + ": scene not built (aur_build_accel)");
        if (j->device != c->device) throw std::invalid_argument("job " + std::to_string(k) + " is on another device");
        if (first_pass[k] < 0 || n_passes[k] < 0) throw std::invalid_argument("negative pass range");
        for (int i = 0; i < k; ++i)
            if (jobs[i] == j)  // two jobs would resolve into one framebuffer at once
                throw std::invalid_argument("job " + std::to_string(k) + " renders the same context as job " +
                                            std::to_string(i));
        if (j->scene.opt.lightsamples < 1)
            throw std::invalid_argument("job " + std::to_string(k) + ": lightsamples must be at least 1");
        if (width == 0) width = j->bvh.width;
        if (j->bvh.width != width) throw std::invalid_argument("all jobs of a batch need the same BVH width");
        depth = std::max(depth, j->bvh.depth);
        max_path = std::max(max_path, j->scene.opt.maxdepth + 1);
        tex = tex || shade_needs_tex(j);
        paths[k] = (long long)pass_paths((size_t)j->scene.opt.width * j->scene.opt.height, first_pass[k], n_passes[k],
                                         j->scene.opt.lightsamples);
    }
    // launch order
    std::vector<int> ord(n_jobs);
    std::iota(ord.begin(), ord.end(), 0);
    if (order == 0)
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return job_cost(jobs[a], paths[a]) < job_cost(jobs[b], paths[b]); });
    long long total_all = 0;
    for (int k = 0; k < n_jobs; ++k) total_all += paths[k];
    // Two pipelines ("lanes") run side by side when the work is large enough
    // (as single-scene renders do): each takes jobs in launch order, a job
    // going to the lane with less estimated work so far, and has its own job
    // table, control block, ring, path pool and stream; the node arena is
    // shared.  The host polls the lanes in turn and queues a lane's next
    // iterations right after handling its status.
    // (subsets of the 64-job workload: 8 jobs, 2 M paths 460 against 425
    // Mpaths/s; 21 jobs, 16 M paths 1100 against 855; all 64, 1.58 G paths
    // 1590 against 1260)
    const int n_lanes = (n_jobs >= 2 && total_all >= (1LL << 21)) ? 2 : 1;
    struct BLane {
        BatchState *b;
        Wavefront *wf;
        hipStream_t stream;
        std::vector<int> rows;  // launch positions, in this lane's order
        std::vector<JobDev> jd;
        std::vector<long long> starts, npaths, job_done;
        std::vector<int> wh, jfirst, jls;
        long long total = 0, max_pass = 0, resolved = 0;
        size_t pool = 0, ring = 0;
        WfView wv;
        BatchView bv;
        hipGraph_t graph = nullptr;
        hipGraphExec_t gexec = nullptr;
        bool done = false;
    };
    BLane lanes[2];
    lanes[0].b = &c->batch;
    lanes[0].wf = &c->wf;
    lanes[0].stream = c->stream;
    lanes[1].b = &c->batch2;
    lanes[1].wf = &c->wf2;
    lanes[1].stream = c->lane_stream[1];
    {
        double cost[2] = {0.0, 0.0};
        for (int r = 0; r < n_jobs; ++r) {
            const int l = (n_lanes == 2 && cost[1] < cost[0]) ? 1 : 0;
            cost[l] += job_cost(jobs[ord[r]], paths[ord[r]]);
            lanes[l].rows.push_back(r);
        }
    }
    // node arena offsets per launch position
    std::vector<long long> node_off(n_jobs), tri_off(n_jobs);
    long long node_total = 0, tri_total = 0;
    for (int r = 0; r < n_jobs; ++r) {
        const Ctx *j = jobs[ord[r]];
        node_off[r] = node_total;
        tri_off[r] = tri_total;
        node_total += j->bvh.n_nodes;
        tri_total += j->n_tris;
    }
    if (node_total > 0x7ffffffeLL || tri_total >= (1LL << 27))
        throw CapacityError("the batch's BVHs exceed the node arena's index range");
    // job records and framebuffers
    for (int l = 0; l < n_lanes; ++l) {
        BLane &L = lanes[l];
        for (int r : L.rows) {
            Ctx *j = jobs[ord[r]];
            HIP_CHECK(hipStreamSynchronize(j->stream));
            const int w = j->scene.opt.width, h = j->scene.opt.height;
            const int whr = w * h;
            bool resized = j->fb_w != w || j->fb_h != h;
            j->accum.reserve(sizeof(float4) * (size_t)whr);
            j->display.reserve(sizeof(ushort4) * (size_t)whr);
            if (first_pass[ord[r]] == 0 || resized) {
                fill_zero(c, j->accum.p, sizeof(float4) * (size_t)whr);
                fill_zero(c, j->display.p, sizeof(ushort4) * (size_t)whr);
                j->display_passes = 0;
            }
            j->fb_w = w;
            j->fb_h = h;
            JobDev d{};
            d.sv = scene_view(j);
            d.accum = j->accum.as<float4>();
            d.display = j->display.as<ushort4>();
            d.display_base = j->display_passes;
            d.seed = j->seed;
            d.first_pass = first_pass[ord[r]];
            d.ls = j->scene.opt.lightsamples;
            d.start = L.total;
            d.n_paths = paths[ord[r]];
            d.root = (int)node_off[r];
            L.jd.push_back(d);
            L.wh.push_back(whr);
            L.jfirst.push_back(d.first_pass);
            L.jls.push_back(d.ls);
            L.starts.push_back(L.total);
            L.npaths.push_back(paths[ord[r]]);
            L.job_done.push_back(0);
            L.total += paths[ord[r]];
            L.max_pass = std::max(L.max_pass, (long long)whr * d.ls);
        }
    }
    // iterations between polls, and the lag after which issued work has finished
    const int lag = max_path + 1;
    const int pairs_per_poll = std::max(1, std::min((iters_per_poll > 0 ? iters_per_poll : 8), BATCH_HIST - lag - 4) / 2);
    for (int l = 0; l < n_lanes; ++l) {
        BLane &L = lanes[l];
        BatchState &b = *L.b;
        const int nj = (int)L.rows.size();
        L.pool = std::max<size_t>(64, std::min<size_t>(c->launch.pool_paths, (size_t)std::max<long long>(L.total, 64)));
        if (L.pool > kMaxPoolPaths) throw CapacityError("path pool too large");
        // the ring holds every item between the resolved prefix and the cursor;
        // a job resolves whole passes, so its prefix may trail by one pass
        L.ring = 1;
        while (L.ring < L.pool * (size_t)(2 * pairs_per_poll + lag + 2) + (size_t)L.max_pass) L.ring <<= 1;
        ensure_wavefront_on(c, *L.wf, L.pool, width, depth);
        b.jobs.reserve(sizeof(JobDev) * nj);
        b.starts.reserve(sizeof(long long) * nj);
        b.t_end.reserve(sizeof(unsigned long long) * nj);
        b.jstats.reserve(sizeof(unsigned long long) * J_STRIDE * nj);
        b.ctl.reserve(sizeof(long long) * CTL_SIZE);
        b.t0.reserve(sizeof(long long));
        b.ring.reserve(sizeof(float4) * L.ring);
        b.ring_cnt.reserve(sizeof(unsigned) * L.ring);
        b.segs.reserve(sizeof(ResolveSeg) * nj);
        // a resolve covers less than the ring: at most ring / STAT_CHUNK chunks plus one per job
        const size_t max_chunks = L.ring / STAT_CHUNK + (size_t)nj + 1;
        b.chunks.reserve(sizeof(StatChunk) * max_chunks);
        if (b.host_chunks_cap < max_chunks) {
            if (b.host_chunks) (void)hipHostFree(b.host_chunks);
            b.host_chunks = nullptr;
            b.host_chunks_cap = 0;
            HIP_CHECK(hipHostMalloc(&b.host_chunks, sizeof(StatChunk) * max_chunks, hipHostMallocDefault));
            b.host_chunks_cap = max_chunks;
        }
        if (!b.host_status)
            HIP_CHECK(hipHostMalloc(&b.host_status, sizeof(long long) * (CTL_SIZE + 8), hipHostMallocDefault));
        if (b.host_segs_cap < (size_t)nj) {
            if (b.host_segs) (void)hipHostFree(b.host_segs);
            b.host_segs = nullptr;
            b.host_segs_cap = 0;
            HIP_CHECK(hipHostMalloc(&b.host_segs, sizeof(ResolveSeg) * (size_t)nj, hipHostMallocDefault));
            b.host_segs_cap = (size_t)nj;
        }
    }
    // the node arena (in lane 0's state): job r's nodes from node_off, its
    // triangles from tri_off
    {
        BatchState &b = c->batch;
        const size_t nsz = width == 4 ? sizeof(WNode<4>) : sizeof(WNode<8>);
        b.arena_nodes.reserve(nsz * (size_t)std::max<long long>(node_total, 1));
        b.arena_tris.reserve(sizeof(float4) * 3 * (size_t)std::max<long long>(tri_total, 1));
        for (int r = 0; r < n_jobs; ++r) {
            const Ctx *j = jobs[ord[r]];
            const int nn = j->bvh.n_nodes;
            if (nn > 0) {
                if (width == 4)
                    hipLaunchKernelGGL(k_arena_nodes<4>, dim3((nn + 255) / 256), dim3(256), 0, c->stream,
                                       j->bvh.nodes.as<WNode<4>>(), nn, (int)node_off[r], (int)tri_off[r],
                                       b.arena_nodes.as<WNode<4>>());
                else
                    hipLaunchKernelGGL(k_arena_nodes<8>, dim3((nn + 255) / 256), dim3(256), 0, c->stream,
                                       j->bvh.nodes.as<WNode<8>>(), nn, (int)node_off[r], (int)tri_off[r],
                                       b.arena_nodes.as<WNode<8>>());
                HIP_CHECK(hipGetLastError());
            }
            if (j->n_tris > 0)
                HIP_CHECK(hipMemcpyAsync(b.arena_tris.as<float4>() + 3 * (size_t)tri_off[r], j->bvh.tri_sorted.p,
                                         sizeof(float4) * 3 * (size_t)j->n_tris, hipMemcpyDeviceToDevice, c->stream));
        }
    }
    PathSource src;
    src.mode = 0;
    src.first_pass = 0;
    src.pixel = src.sample = nullptr;
    src.slots = nullptr;
    for (int l = 0; l < n_lanes; ++l) {
        BLane &L = lanes[l];
        const int nj = (int)L.rows.size();
        // the copies read host vectors that outlive the stream synchronisation below
        HIP_CHECK(hipMemcpyAsync(L.b->jobs.p, L.jd.data(), sizeof(JobDev) * nj, hipMemcpyHostToDevice, c->stream));
        HIP_CHECK(hipMemcpyAsync(L.b->starts.p, L.starts.data(), sizeof(long long) * nj, hipMemcpyHostToDevice, c->stream));
        L.wv = wf_view_on(c, *L.wf, src);
        L.bv = batch_view(*L.b, c->batch, nj, L.total, (long long)L.ring - 1);
        L.bv.pool = (int)L.pool;
    }

    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    HIP_CHECK(hipEventRecord(e0, c->stream));
    for (int l = 0; l < n_lanes; ++l)
        hipLaunchKernelGGL(k_binit, dim3(std::max(1, std::min(((int)lanes[l].rows.size() + 255) / 256, c->n_cu * 4))),
                           dim3(256), 0, c->stream, lanes[l].bv, lanes[l].wv);
    HIP_CHECK(hipGetLastError());
    if (n_lanes == 2) {
        // lane 1 starts after the set-up above
        HIP_CHECK(hipEventRecord(c->ev_start, c->stream));
        HIP_CHECK(hipStreamWaitEvent(lanes[1].stream, c->ev_start, 0));
    }
    auto release = [&]() {
        for (BLane &L : lanes) {
            if (L.gexec) (void)hipGraphExecDestroy(L.gexec);
            if (L.graph) (void)hipGraphDestroy(L.graph);
            L.gexec = nullptr;
            L.graph = nullptr;
        }
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        e0 = e1 = nullptr;
    };
    // A path holds its slot for at most maxdepth + 1 iterations, so every slot
    // completes a path at least that often and a lane drains within
    // (total / pool + 1) * (maxdepth + 1) iterations; twice that guards the loop.
    float ms = 0.f;
    auto issue = [&](BLane &L) {
        for (int k = 0; k < pairs_per_poll; ++k) {
            if (L.gexec) HIP_CHECK(hipGraphLaunch(L.gexec, L.stream));
            else {
                batch_iteration(c, L.stream, L.bv, L.wv, 0, width, tex);
                batch_iteration(c, L.stream, L.bv, L.wv, 1, width, tex);
            }
        }
        HIP_CHECK(hipGetLastError());
        long long *hs = static_cast<long long *>(L.b->host_status);
        HIP_CHECK(hipMemcpyAsync(hs, L.b->ctl.p, sizeof(long long) * CTL_TIME, hipMemcpyDeviceToHost, L.stream));
        HIP_CHECK(hipMemcpyAsync(hs + CTL_SIZE, L.wv.counts, sizeof(int) * 4, hipMemcpyDeviceToHost, L.stream));
    };
    try {
        for (int l = 0; l < n_lanes; ++l) {
            BLane &L = lanes[l];
            if (use_graph) {
                HIP_CHECK(hipStreamBeginCapture(L.stream, hipStreamCaptureModeThreadLocal));
                batch_iteration(c, L.stream, L.bv, L.wv, 0, width, tex);
                batch_iteration(c, L.stream, L.bv, L.wv, 1, width, tex);
                HIP_CHECK(hipStreamEndCapture(L.stream, &L.graph));
                HIP_CHECK(hipGraphInstantiate(&L.gexec, L.graph, nullptr, nullptr, 0));
            }
        }
        for (int l = 0; l < n_lanes; ++l) issue(lanes[l]);
        for (int remaining = n_lanes; remaining > 0;) {
            for (int l = 0; l < n_lanes; ++l) {
                BLane &L = lanes[l];
                if (L.done) continue;
                HIP_CHECK(hipStreamSynchronize(L.stream));
                const long long *hs = static_cast<const long long *>(L.b->host_status);
                const int *hcounts = reinterpret_cast<const int *>(hs + CTL_SIZE);
                const long long cur = hs[CTL_CURSOR], iters = hs[CTL_ITER];
                // after a pair of iterations the live paths sit in queue 0
                const bool done = cur >= L.total && hcounts[C_Q0] == 0;
                long long safe = 0;
                if (done) safe = L.total;
                else if (iters - 1 - lag >= 0) safe = hs[CTL_HIST + (int)((iters - 1 - lag) % BATCH_HIST)];
                if (done) hipLaunchKernelGGL(k_bstamp, dim3(1), dim3(64), 0, L.stream, L.bv);
                if (safe > L.resolved) {
                    resolve_range(c, *L.b, L.stream, L.bv, L.starts, L.npaths, L.wh, L.jfirst, L.jls, L.job_done, safe);
                    L.resolved = safe;
                }
                if (done) {
                    L.done = true;
                    --remaining;
                    continue;
                }
                const long long max_iters = 2 * (L.total / (long long)L.pool + 1) * (long long)max_path + 64;
                if (iters > max_iters) throw StateError("batched launch did not drain");
                issue(L);
            }
        }
        if (n_lanes == 2) {
            HIP_CHECK(hipEventRecord(c->ev_join, lanes[1].stream));
            HIP_CHECK(hipStreamWaitEvent(c->stream, c->ev_join, 0));
        }
        HIP_CHECK(hipEventRecord(e1, c->stream));
        HIP_CHECK(hipEventSynchronize(e1));
        HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    } catch (...) {
        for (int l = 0; l < n_lanes; ++l) (void)hipStreamSynchronize(lanes[l].stream);
        release();
        throw;
    }
    release();

    int rate_khz = 0;
    HIP_CHECK(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, c->device));
    aur_stats tot;
    std::memset(&tot, 0, sizeof tot);
    for (int l = 0; l < n_lanes; ++l) {
        BLane &L = lanes[l];
        const int nj = (int)L.rows.size();
        std::vector<unsigned long long> t_end(nj);
        std::vector<unsigned long long> js((size_t)nj * J_STRIDE);
        long long t0 = 0;
        HIP_CHECK(hipMemcpy(t_end.data(), L.b->t_end.p, sizeof(unsigned long long) * nj, hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(js.data(), L.b->jstats.p, sizeof(unsigned long long) * js.size(), hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(&t0, L.b->t0.p, sizeof t0, hipMemcpyDeviceToHost));
        for (int i = 0; i < nj; ++i) {
            const int r = L.rows[i];
            if ((long long)js[(size_t)i * J_STRIDE + J_PATHS] != L.npaths[i])
                throw StateError("batched launch: job " + std::to_string(ord[r]) + " finished " +
                                 std::to_string(js[(size_t)i * J_STRIDE + J_PATHS]) + " of " +
                                 std::to_string(L.npaths[i]) + " paths");
            aur_job_result &o = results[ord[r]];
            o.paths = paths[ord[r]];
            o.ext_rays = (int64_t)js[(size_t)i * J_STRIDE + J_EXT];
            o.hit_rays = (int64_t)js[(size_t)i * J_STRIDE + J_HIT];
            o.shadow_rays = (int64_t)js[(size_t)i * J_STRIDE + J_SHADOW];
            o.latency_ms = (paths[ord[r]] > 0 && rate_khz > 0) ? (double)((long long)t_end[i] - t0) / (double)rate_khz : 0.0;
            o.order = r;
            jobs[ord[r]]->display_passes += n_passes[ord[r]];
            tot.paths += o.paths;
            tot.ext_rays += o.ext_rays;
            tot.hit_rays += o.hit_rays;
            tot.shadow_rays += o.shadow_rays;
        }
    }
    tot.ms = ms;
    if (total_stats) *total_stats = tot;
}

}  // namespace aur
