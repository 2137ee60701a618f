// batch.hip -- many scenes in one persistent wavefront launch.
//
// The reference renders one scene per process: `Renderer::renderImage`
// (src/frontend/renderer.cpp:201-271) runs its passes with one TBB
// parallel_for each.  Here a list of jobs (each a built scene, a pass range
// and its own framebuffer) shares one pool of path slots:
//
//   regen   -- every free slot takes the next path from a global work cursor
//              that runs over the jobs in launch order (shortest estimated job
//              first by default); the job is found by binary search over the
//              jobs' first work indices, the pixel and pass from the offset in
//              the job (the same 8x8-tiled order as a single-scene launch)
//   extend, shade, shadow -- the single-scene kernels' bodies, with the scene
//              of each path read from its job record
//   finish  -- firefly clamp, atomic add into the job's framebuffer, and a
//              per-job countdown; the last path of a job stamps the wall clock
//
// so the GPU stays full until the last job drains, whatever the jobs' sizes,
// resolutions, sample counts and depths.  One iteration is a fixed sequence
// of launches whose counts live in device memory; two iterations (the two
// extension queues swap) are captured once as a hipGraph and replayed, and
// the host reads the cursor and the live count only every few iterations.
//
// A path's random numbers depend only on (job seed, pixel, sample), so each
// job's framebuffer equals a single-scene render of the same passes up to
// the order of the floating-point sums.
#include "wavefront_common.h"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace aur {

struct JobDev {
    SceneView sv;
    float4 *accum;       // the job's framebuffer: sum of per-pass rgba
    uint32_t seed;
    int first_pass;
    long long start;     // first work index of the job in launch order
    long long n_paths;
};

enum { J_EXT = 0, J_HIT = 1, J_SHADOW = 2, J_STRIDE = 4 };

struct BatchView {
    const JobDev *jobs;      // in launch order
    const long long *starts; // jobs[k].start, for the binary search
    int n_jobs;
    long long total;         // paths over all jobs
    unsigned long long *cursor;
    int *pjob, *ppix, *psmp; // per slot: job (launch order), pixel, sample
    long long *remaining;    // per job: paths not yet finished
    long long *t_end;        // per job: wall clock when its last path finished
    long long *t0;           // wall clock at the start of the launch
    unsigned long long *jstats;  // per job: J_STRIDE counters
};

__global__ void k_binit(BatchView bv, WfView wv, int pool) {
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const int stride = gridDim.x * blockDim.x;
    for (int i = gtid; i < pool; i += stride) wv.finq[i] = i;
    for (int j = gtid; j < bv.n_jobs; j += stride) {
        bv.remaining[j] = bv.jobs[j].n_paths;
        bv.t_end[j] = 0;
        for (int k = 0; k < J_STRIDE; ++k) bv.jstats[(size_t)j * J_STRIDE + k] = 0;
    }
    if (gtid == 0) {
        for (int k = 0; k < 16; ++k) wv.counts[k] = 0;
        for (int k = 0; k < 8; ++k) wv.stats[k] = 0;
        wv.counts[C_FINISH] = pool;  // every slot starts free
        *bv.cursor = 0;
        *bv.t0 = wall_clock64();
    }
}

// Free slots (the finish list of the previous iteration) take new paths.
__global__ void k_regen(BatchView bv, WfView wv, int q) {
    const int n = wv.counts[C_FINISH];
    const int stride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int slot = wv.finq[i];
        // one cursor atomic per wave
        const unsigned long long mask = __ballot(true);
        const int lane = __lane_id();
        const int leader = __ffsll((long long)mask) - 1;
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(bv.cursor, (unsigned long long)__popcll(mask));
        const unsigned lo = __shfl((unsigned)(base & 0xffffffffu), leader);
        const unsigned hi = __shfl((unsigned)(base >> 32), leader);
        base = ((unsigned long long)hi << 32) | lo;
        const long long w = (long long)(base + (unsigned long long)__popcll(mask & ((1ull << lane) - 1ull)));
        const bool valid = w < bv.total;
        f3 d = mk3(0.f, 0.f, 1.f);
        int job = 0;
        if (valid) {
            int a = 0, b = bv.n_jobs - 1;  // last job whose start <= w
            while (a < b) {
                int mid = (a + b + 1) >> 1;
                if (bv.starts[mid] <= w) a = mid;
                else b = mid - 1;
            }
            job = a;
            const JobDev &jd = bv.jobs[a];
            const long long local = w - jd.start;
            const int wh = jd.sv.width * jd.sv.height;
            const int pass = (int)(local / wh);
            const int pixel = tiled_pixel((int)(local - (long long)pass * wh), jd.sv.width, jd.sv.height);
            const int sample = jd.first_pass + pass;
            PathRng rng;
            rng.slots = nullptr;
            rng.key = path_key(jd.seed, (uint32_t)pixel, (uint32_t)sample);
            d = camera_dir(jd.sv, pixel, rng.u(0), rng.u(1));
            bv.pjob[slot] = a;
            bv.ppix[slot] = pixel;
            bv.psmp[slot] = sample;
            wv.T[slot] = make_float4(1.f, 1.f, 1.f, 1.f);
            wv.La[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
            wv.Lb[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
            wv.flags[slot] = RT_CAMERA;
        }
        wave_count(&wv.stats[S_PATHS], valid);
        const int qs = wave_reserve(&wv.counts[q], valid);
        if (valid) {
            wv.ray[q][2 * qs] = make_float4(0.f, 0.f, 0.f, 0.1f);
            wv.ray[q][2 * qs + 1] = make_float4(d.x, d.y, d.z, 10000000.f);
            wv.qpath[q][qs] = slot;
            wv.qjob[q][qs] = job;
        }
    }
}

// The finish list has been consumed by regen; shade and shadow start empty.
__global__ void k_bclear_lists(int *counts) {
    if (threadIdx.x == 0) {
        counts[C_FINISH] = 0;
        counts[C_SHADOW] = 0;
    }
}

__global__ void k_bclear_queue(int *counts, int q) {
    if (threadIdx.x == 0) counts[q] = 0;
}

// The kernels below run each wave once per distinct job among its lanes
// (almost always one): the job index is made wave-uniform with
// readfirstlane, so the scene record is read with scalar loads and the
// traversal and shading code sees uniform scene pointers, as in the
// single-scene kernels.
template <int W, int B>
__global__ void __launch_bounds__(B) k_bextend(BatchView bv, const JobDev *__restrict__ jobs, WfView wv, int q) {
    __shared__ int s_stack[AUR_STACK_LDS * B];
    const int n = wv.counts[q];
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&wv.stats[S_EXT], (unsigned long long)n);
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const int stride = gridDim.x * blockDim.x;
    TravStack st;
    st.lds = s_stack + threadIdx.x;
    st.stride = B;
    st.ovf = wv.ovf + (size_t)gtid * wv.ovf_per_thread;
    for (int i = gtid; i < n; i += stride) {
        const int job = wv.qjob[q][i];
        const float4 a = wv.ray[q][2 * i], b = wv.ray[q][2 * i + 1];
        TravHit h;
        for (;;) {
            const int j = __builtin_amdgcn_readfirstlane(job);
            if (job == j) {
                traverse<W, false>(static_cast<const WNode<W> *>(jobs[j].sv.nodes), jobs[j].sv.tris, mk3(a.x, a.y, a.z),
                                   mk3(b.x, b.y, b.z), a.w, b.w, st, h);
                wave_count(&bv.jstats[(size_t)j * J_STRIDE + J_EXT], true);
                wave_count(&bv.jstats[(size_t)j * J_STRIDE + J_HIT], h.prim >= 0);
                break;
            }
        }
        wv.hit[i] = make_float4(h.t, h.u, h.v, __int_as_float(h.prim));
        wave_count(&wv.stats[S_HIT], h.prim >= 0);
    }
}

__global__ void __launch_bounds__(256) k_bshade(BatchView bv, const JobDev *__restrict__ jobs, WfView wv, int q) {
    const int n = wv.counts[q];
    const int stride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int job = wv.qjob[q][i];
        for (;;) {
            const int j = __builtin_amdgcn_readfirstlane(job);
            if (job == j) {
                const uint32_t seed = jobs[j].seed;
                shade_one(jobs[j].sv, wv, q, i, [&](int p) {
                    PathRng r;
                    r.slots = nullptr;
                    r.key = path_key(seed, (uint32_t)bv.ppix[p], (uint32_t)bv.psmp[p]);
                    return r;
                }, j);
                break;
            }
        }
    }
}

template <int W, int B>
__global__ void __launch_bounds__(B) k_bshadow(BatchView bv, const JobDev *__restrict__ jobs, WfView wv) {
    __shared__ int s_stack[AUR_STACK_LDS * B];
    const int n = wv.counts[C_SHADOW];
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&wv.stats[S_SHADOW], (unsigned long long)n);
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const int stride = gridDim.x * blockDim.x;
    TravStack st;
    st.lds = s_stack + threadIdx.x;
    st.stride = B;
    st.ovf = wv.ovf + (size_t)gtid * wv.ovf_per_thread;
    for (int i = gtid; i < n; i += stride) {
        const float4 a = wv.sray[3 * i], b = wv.sray[3 * i + 1], c = wv.sray[3 * i + 2];
        const int tag = __float_as_int(c.w);
        const int path = tag & 0x7fffffff;
        const int job = bv.pjob[path];
        TravHit h;
        for (;;) {
            const int j = __builtin_amdgcn_readfirstlane(job);
            if (job == j) {
                traverse<W, true>(static_cast<const WNode<W> *>(jobs[j].sv.nodes), jobs[j].sv.tris, mk3(a.x, a.y, a.z),
                                  mk3(b.x, b.y, b.z), a.w, b.w, st, h);
                wave_count(&bv.jstats[(size_t)j * J_STRIDE + J_SHADOW], true);
                break;
            }
        }
        if (h.prim < 0) {
            float4 *acc = (tag < 0) ? wv.Lb : wv.La;
            float4 v = acc[path];
            v.x += c.x; v.y += c.y; v.z += c.z;
            acc[path] = v;
        }
    }
}

__device__ inline void job_done_count(const BatchView &bv, int job) {
    const unsigned long long mask = __ballot(true);
    const int lane = __lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    const int j0 = __shfl(job, leader);
    const unsigned long long same = __ballot(job == j0);
    if (lane == leader) {
        const long long c = (long long)__popcll(same);
        if (atomicAdd((unsigned long long *)&bv.remaining[j0], (unsigned long long)(-c)) == (unsigned long long)c)
            bv.t_end[j0] = wall_clock64();
    }
    if (job != j0) {
        if (atomicAdd((unsigned long long *)&bv.remaining[job], (unsigned long long)(-1ll)) == 1ull)
            bv.t_end[job] = wall_clock64();
    }
}

__global__ void k_bfinish(BatchView bv, WfView wv) {
    const int n = wv.counts[C_FINISH];
    const int stride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int slot = wv.finq[i];
        const float4 a = wv.La[slot], b = wv.Lb[slot];
        const f3 L = firefly_clamp(mk3(a.x + b.x, a.y + b.y, a.z + b.z), wv.flags[slot]);
        const int job = bv.pjob[slot];
        float *acc = reinterpret_cast<float *>(bv.jobs[job].accum + bv.ppix[slot]);
        atomicAdd(acc + 0, L.x);
        atomicAdd(acc + 1, L.y);
        atomicAdd(acc + 2, L.z);
        atomicAdd(acc + 3, a.w);
        job_done_count(bv, job);
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static BatchView batch_view(Ctx *c, int n_jobs, long long total) {
    BatchState &b = c->batch;
    BatchView v;
    v.jobs = b.jobs.as<JobDev>();
    v.starts = b.starts.as<long long>();
    v.n_jobs = n_jobs;
    v.total = total;
    v.cursor = b.cursor.as<unsigned long long>();
    v.pjob = b.pjob.as<int>();
    v.ppix = b.ppix.as<int>();
    v.psmp = b.psmp.as<int>();
    v.remaining = b.remaining.as<long long>();
    v.t_end = b.t_end.as<long long>();
    v.t0 = b.t0.as<long long>();
    v.jstats = b.jstats.as<unsigned long long>();
    return v;
}

// One iteration: regen, extend, shade, shadow, finish over queue q.
static void batch_iteration(Ctx *c, const BatchView &bv, const WfView &wv, int q, int width) {
    const int tb = trace_blocks(c);
    const int sb = c->n_cu * 4;
    hipLaunchKernelGGL(k_regen, dim3(sb), dim3(256), 0, c->stream, bv, wv, q);
    hipLaunchKernelGGL(k_bclear_lists, dim3(1), dim3(64), 0, c->stream, wv.counts);
#define AUR_BEXT(w_, b_) hipLaunchKernelGGL((k_bextend<w_, b_>), dim3(tb), dim3(b_), 0, c->stream, bv, bv.jobs, wv, q)
#define AUR_BSHD(w_, b_) hipLaunchKernelGGL((k_bshadow<w_, b_>), dim3(tb), dim3(b_), 0, c->stream, bv, bv.jobs, wv)
    AUR_TRAV_DISPATCH(width, c->launch.trace_block, AUR_BEXT);
    hipLaunchKernelGGL(k_bshade, dim3(sb), dim3(256), 0, c->stream, bv, bv.jobs, wv, q);
    AUR_TRAV_DISPATCH(width, c->launch.trace_block, AUR_BSHD);
#undef AUR_BEXT
#undef AUR_BSHD
    hipLaunchKernelGGL(k_bfinish, dim3(sb), dim3(256), 0, c->stream, bv, wv);
    hipLaunchKernelGGL(k_bclear_queue, dim3(1), dim3(64), 0, c->stream, wv.counts, q);
}

static double job_cost(const Ctx *j, long long paths) {
    // relative cost estimate: paths x depth x tris^0.3
    return (double)paths * (j->scene.opt.maxdepth + 1) * std::pow((double)std::max(j->n_tris, 1), 0.3);
}

void run_jobs(Ctx *c, Ctx *const *jobs, const int *first_pass, const int *n_passes, int n_jobs, int order, int use_graph,
              int iters_per_poll, aur_job_result *results, aur_stats *total_stats) {
    if (n_jobs <= 0) throw std::invalid_argument("no jobs");
    int width = 0, depth = 0, max_path = 0;
    std::vector<long long> paths(n_jobs);
    for (int k = 0; k < n_jobs; ++k) {
        Ctx *j = jobs[k];
        if (!j->built) throw StateError("job " + std::to_string(k) + ": scene not built (aur_build_accel)");
        if (j->device != c->device) throw std::invalid_argument("job " + std::to_string(k) + " is on another device");
        if (first_pass[k] < 0 || n_passes[k] < 0) throw std::invalid_argument("negative pass range");
        if (width == 0) width = j->bvh.width;
        if (j->bvh.width != width) throw std::invalid_argument("all jobs of a batch need the same BVH width");
        depth = std::max(depth, j->bvh.depth);
        max_path = std::max(max_path, j->scene.opt.maxdepth + 1);
        paths[k] = (long long)j->scene.opt.width * j->scene.opt.height * n_passes[k];
    }
    // launch order
    std::vector<int> ord(n_jobs);
    std::iota(ord.begin(), ord.end(), 0);
    if (order == 0)
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return job_cost(jobs[a], paths[a]) < job_cost(jobs[b], paths[b]); });
    // job records and framebuffers
    std::vector<JobDev> jd(n_jobs);
    std::vector<long long> starts(n_jobs);
    long long total = 0;
    for (int r = 0; r < n_jobs; ++r) {
        Ctx *j = jobs[ord[r]];
        HIP_CHECK(hipStreamSynchronize(j->stream));
        const int w = j->scene.opt.width, h = j->scene.opt.height;
        const size_t wh = (size_t)w * h;
        bool resized = j->fb_w != w || j->fb_h != h;
        j->accum.reserve(sizeof(float4) * wh);
        if (first_pass[ord[r]] == 0 || resized) fill_zero(c, j->accum.p, sizeof(float4) * wh);
        j->fb_w = w;
        j->fb_h = h;
        jd[r].sv = scene_view(j);
        jd[r].accum = j->accum.as<float4>();
        jd[r].seed = j->seed;
        jd[r].first_pass = first_pass[ord[r]];
        jd[r].start = total;
        jd[r].n_paths = paths[ord[r]];
        starts[r] = total;
        total += paths[ord[r]];
    }
    // buffers
    const size_t pool = std::max<size_t>(64, std::min<size_t>(c->launch.batch_pool_paths, (size_t)std::max<long long>(total, 64)));
    if (pool > (size_t)0x3fffffff) throw CapacityError("path pool too large");
    ensure_wavefront_for(c, pool, width, depth);
    BatchState &b = c->batch;
    b.jobs.reserve(sizeof(JobDev) * n_jobs);
    b.starts.reserve(sizeof(long long) * n_jobs);
    b.remaining.reserve(sizeof(long long) * n_jobs);
    b.t_end.reserve(sizeof(long long) * n_jobs);
    b.jstats.reserve(sizeof(unsigned long long) * J_STRIDE * n_jobs);
    b.cursor.reserve(sizeof(unsigned long long));
    b.t0.reserve(sizeof(long long));
    b.pjob.reserve(sizeof(int) * pool);
    b.ppix.reserve(sizeof(int) * pool);
    b.psmp.reserve(sizeof(int) * pool);
    b.qjob[0].reserve(sizeof(int) * pool);
    b.qjob[1].reserve(sizeof(int) * pool);
    if (!b.host_status) HIP_CHECK(hipHostMalloc(&b.host_status, 256, hipHostMallocDefault));
    HIP_CHECK(hipMemcpyAsync(b.jobs.p, jd.data(), sizeof(JobDev) * n_jobs, hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipMemcpyAsync(b.starts.p, starts.data(), sizeof(long long) * n_jobs, hipMemcpyHostToDevice, c->stream));

    PathSource src;
    src.mode = 0;
    src.first_pass = 0;
    src.pixel = src.sample = nullptr;
    src.slots = nullptr;
    WfView wv = wf_view(c, src);
    wv.qjob[0] = b.qjob[0].as<int>();
    wv.qjob[1] = b.qjob[1].as<int>();
    BatchView bv = batch_view(c, n_jobs, total);

    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    HIP_CHECK(hipEventRecord(e0, c->stream));
    hipLaunchKernelGGL(k_binit, dim3(std::min<size_t>((pool + 255) / 256, (size_t)c->n_cu * 4)), dim3(256), 0, c->stream,
                       bv, wv, (int)pool);
    HIP_CHECK(hipGetLastError());

    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    if (use_graph) {
        HIP_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        batch_iteration(c, bv, wv, 0, width);
        batch_iteration(c, bv, wv, 1, width);
        hipError_t ec = hipStreamEndCapture(c->stream, &graph);
        if (ec != hipSuccess) throw HipError(std::string("hipStreamEndCapture failed: ") + hipGetErrorString(ec));
        HIP_CHECK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
    }
    const int pairs_per_poll = std::max(1, (iters_per_poll > 0 ? iters_per_poll : 8) / 2);
    const long long *st = static_cast<const long long *>(b.host_status);
    // A path holds its slot for at most maxdepth + 1 iterations, so every slot
    // completes a path at least that often and the launch drains within
    // (total / pool + 1) * (maxdepth + 1) iterations; twice that guards the loop.
    const long long max_iters = 2 * (total / (long long)pool + 1) * (long long)max_path + 64;
    long long iters = 0;
    for (;;) {
        for (int k = 0; k < pairs_per_poll; ++k) {
            if (gexec) HIP_CHECK(hipGraphLaunch(gexec, c->stream));
            else {
                batch_iteration(c, bv, wv, 0, width);
                batch_iteration(c, bv, wv, 1, width);
            }
            iters += 2;
        }
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpyAsync(b.host_status, b.cursor.p, sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipMemcpyAsync(static_cast<char *>(b.host_status) + 8, wv.counts, sizeof(int) * 4, hipMemcpyDeviceToHost,
                                 c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
        const unsigned long long cur = *reinterpret_cast<const unsigned long long *>(st);
        const int *cnt = reinterpret_cast<const int *>(static_cast<const char *>(b.host_status) + 8);
        // after a pair of iterations the live paths sit in queue 0
        if ((long long)cur >= total && cnt[C_Q0] == 0) break;
        if (iters > max_iters) {
            if (gexec) (void)hipGraphExecDestroy(gexec);
            if (graph) (void)hipGraphDestroy(graph);
            throw StateError("batched launch did not drain");
        }
    }
    HIP_CHECK(hipEventRecord(e1, c->stream));
    HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);

    std::vector<long long> t_end(n_jobs), rem(n_jobs);
    std::vector<unsigned long long> js((size_t)n_jobs * J_STRIDE);
    long long t0 = 0;
    unsigned long long s[8];
    HIP_CHECK(hipMemcpy(t_end.data(), b.t_end.p, sizeof(long long) * n_jobs, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(rem.data(), b.remaining.p, sizeof(long long) * n_jobs, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(js.data(), b.jstats.p, sizeof(unsigned long long) * js.size(), hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(&t0, b.t0.p, sizeof t0, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(s, c->wf.stats.p, sizeof s, hipMemcpyDeviceToHost));
    int rate_khz = 0;
    HIP_CHECK(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, c->device));
    for (int r = 0; r < n_jobs; ++r) {
        if (rem[r] != 0) throw StateError("batched launch left unfinished paths");
        aur_job_result &o = results[ord[r]];
        o.paths = paths[ord[r]];
        o.ext_rays = (int64_t)js[(size_t)r * J_STRIDE + J_EXT];
        o.hit_rays = (int64_t)js[(size_t)r * J_STRIDE + J_HIT];
        o.shadow_rays = (int64_t)js[(size_t)r * J_STRIDE + J_SHADOW];
        o.latency_ms = (paths[ord[r]] > 0 && rate_khz > 0) ? (double)(t_end[r] - t0) / (double)rate_khz : 0.0;
        o.order = r;
    }
    if (total_stats) {
        total_stats->paths = (int64_t)s[S_PATHS];
        total_stats->ext_rays = (int64_t)s[S_EXT];
        total_stats->hit_rays = (int64_t)s[S_HIT];
        total_stats->shadow_rays = (int64_t)s[S_SHADOW];
        total_stats->ms = ms;
    }
}

}  // namespace aur
