// context.h -- the library context behind `aur_ctx` and its device buffers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <utility>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/aurora_hip.h"
#include "dev_common.h"
#include "halton.h"
#include "host_scene.h"

namespace aur {

struct HipError : std::runtime_error {
    explicit HipError(const std::string &m) : std::runtime_error(m) {}
};
struct StateError : std::runtime_error {
    explicit StateError(const std::string &m) : std::runtime_error(m) {}
};
struct CapacityError : std::runtime_error {
    explicit CapacityError(const std::string &m) : std::runtime_error(m) {}
};
struct IoError : std::runtime_error {
    explicit IoError(const std::string &m) : std::runtime_error(m) {}
};

#define HIP_CHECK(expr)                                                                          \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            throw ::aur::HipError(std::string(#expr) + " failed: " + hipGetErrorString(e_));     \
    } while (0)

// Owned device allocation.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    // Grows (never shrinks) to at least n bytes; contents are not preserved.
    void reserve(size_t n) {
        if (n <= bytes) return;
        release();
        size_t want = n < 256 ? 256 : n;
        HIP_CHECK(hipMalloc(&p, want));
        bytes = want;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
    void swap(DevBuf &o) {
        std::swap(p, o.p);
        std::swap(bytes, o.bytes);
    }
};

// Wide-BVH node: W boxes as SoA rows + W child codes (see bvh_build.hip).
// child >= 0: inner node index; child < 0: leaf, ~child = (first << 4) | (count - 1).
// Inner children occupy the first slots and consecutive indices:
// child[k] = pad[0] + k for each inner slot k; pad[1..] are zero.
template <int W> struct alignas(16) WNode {
    float lox[W], hix[W], loy[W], hiy[W], loz[W], hiz[W];
    int child[W];
    int pad[W];
};

struct Bvh {
    int width = 4, max_leaf = 4;
    int n_nodes = 0, depth = 0;
    DevBuf nodes;       // WNode<width>[n_nodes]
    DevBuf tri_sorted;  // float4[3 * n_tris]: (v0, prim), (e1, 0), (e2, 0)
};

// Wavefront working set, sized for `cap` paths in flight.  Path state is
// stored by queue position and moves with the queue (see wavefront_common.h).
struct Wavefront {
    size_t cap = 0;
    DevBuf ray[2];      // float4[2 * cap]: (o, tmin), (d, tmax) per queue entry
    DevBuf T[2];        // float4[cap]: throughput rgb, continue probability
    DevBuf La[2], Lb[2];  // float4[cap]: radiance partial sums (alpha in La.w)
    DevBuf meta[2];     // int4[cap]: flags and the path's identity
    DevBuf hit;         // float4[cap]
    DevBuf sray;        // float4[3 * 2cap]: shadow ray (o, job), (d, tmax), (contribution, target)
    DevBuf finLa, finLb, finMeta;  // finished paths' state, after the shadow pass
    DevBuf out;         // float4[cap]: per-path rgba
    DevBuf counts;      // int[16]: queue sizes
    DevBuf stats;       // unsigned long long[8]
    DevBuf partials;    // unsigned long long per traversal workgroup
    DevBuf stack_ovf;   // traversal stack overflow
    size_t stack_ovf_per_thread = 0;
};

// Optional per-kernel timing with HIP events recorded on the library stream
// around each launch.
enum { KT_GENERATE = 0, KT_EXTEND, KT_SHADE, KT_SHADOW, KT_FINISH, KT_RESOLVE, KT_COUNT };
struct KernelTimer {
    bool on = false;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<hipEvent_t> pool;
    double total_ms[KT_COUNT] = {0};
    long long launches[KT_COUNT] = {0};
};

// Largest path pool: queue entries are indexed with 32-bit ints, and the
// shadow queue (two rays per path) stores three float4 records per ray
constexpr size_t kMaxPoolPaths = 0x7fffffff / 6;

// Launch shape of the tracer kernels (set by aur_set_launch / the autotuner).
struct LaunchCfg {
    int trace_block = 64;           // traversal workgroup size: 64, 128 or 256
    int trace_waves_per_cu = 32;    // persistent traversal grid, waves per CU (the traversal kernels fit 32)
    // path slots in flight: the batched launch's pool, and the largest chunk
    // of whole passes a single-scene render traces at once (368 bytes of
    // wavefront state per slot, allocated for the chunks a render needs: up
    // to 47 GB at the default 128 M, which holds a 64-spp 1080p frame).
    // Every chunk ends in a latency-bound drain (the last bounces of its
    // longest paths), so larger chunks amortise it over more paths
    size_t pool_paths = 128u << 20;
    // a bounce's queue is shaded from per-workgroup hit lists when it holds
    // at least this many 1024-entry tiles (in place below); aur_create sets
    // four per shade workgroup
    int shade_list_tiles = 4096;
};

// Device state of a batched launch (batch.hip), kept between calls.
struct BatchState {
    DevBuf jobs, starts, t_end, jstats, ctl, t0, segs, chunks;
    DevBuf ring, ring_cnt;          // finished paths' rgba and packed counts by work index
    DevBuf arena_nodes, arena_tris; // every job's BVH nodes / triangle records, back to back
    void *host_status = nullptr;    // pinned: control block + queue counts, polled by the host
    void *host_segs = nullptr;      // pinned: resolve segment table
    size_t host_segs_cap = 0;
    void *host_chunks = nullptr;    // pinned: statistics chunk table
    size_t host_chunks_cap = 0;
    BatchState() = default;
    BatchState(const BatchState &) = delete;
    BatchState &operator=(const BatchState &) = delete;
    ~BatchState() {
        if (host_status) (void)hipHostFree(host_status);
        if (host_segs) (void)hipHostFree(host_segs);
        if (host_chunks) (void)hipHostFree(host_chunks);
    }
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    uint32_t seed = 0x5EED1234u;
    int n_cu = 256;

    HostScene scene;
    bool have_scene = false;
    bool built = false;
    std::vector<float> tri_albedo;  // optional per-triangle override
    // optional FP32 RGB albedo texture driving the diffuse colour of the
    // attributes in ptex_attrs (aur_set_albedo_texture)
    std::vector<float> ptex;
    int ptex_w = 0, ptex_h = 0;
    std::vector<int> ptex_attrs;

    // device scene
    DevBuf d_verts, d_tris, d_shade, d_mats, d_lights, d_tables, d_albedo, d_ptex;
    DevBuf d_uv, d_attrx, d_shaders, d_texs, d_texels;  // shading inputs (empty without shaders)
    int sampler = 0;                                      // AUR_SAMPLER_*
    DevBuf d_halton_dims, d_halton_tables, d_halton_view; // Halton source (built on first use)
    int halton_w = 0, halton_h = 0;
    bool has_uv = false, has_attrx = false;
    int n_verts = 0, n_tris = 0, n_mats = 0, n_lights = 0;
    float env_le[3] = {0, 0, 0};
    int n_env = 0;
    Bvh bvh;
    std::map<std::pair<int, int>, int> table_ids;  // (reflectance/20 index, distribution) -> table
    std::vector<DMat> dmats;
    std::vector<DLight> dlights;

    // framebuffer (sum of per-pass rgba)
    DevBuf accum;
    DevBuf display;  // ushort4 per pixel: OpenexrDisplay's FP16 running mean (display_update)
    // passes appended to `display` since it was last cleared: the running
    // mean's sample count (openexrDisplay.cpp:61-72 counts appended values)
    long long display_passes = 0;
    int fb_w = 0, fb_h = 0;

    // Inputs generation: bumped by every call that can change a path's
    // radiance (scene, seed, sampler, material parameters, albedo overrides).
    uint64_t version = 1;
    // Per-path records of passes [rec_first, rec_first + rec_n) kept from
    // forward renders (aur_keep_path_records), in the adjoint's record layout;
    // rec_version[p] is the inputs generation pass rec_first + p was recorded
    // at (0: not recorded), rec_wh the image size they were recorded for.
    DevBuf path_rec;
    int rec_first = 0, rec_n = 0;
    size_t rec_wh = 0;
    std::vector<uint64_t> rec_version;

    // run_paths traces its chunks on two path pipelines ("lanes") at the same
    // time: lane 0 on `wf`, lane 1 on `wf2`.  Each lane has a main stream and
    // a side stream (a bounce's shadow rays and finished paths are traced on
    // the side stream while the next bounce's extension rays are traced on
    // the main one).  Lane 0 runs on `stream` and lane_stream[0], lane 1 on
    // lane_stream[1] and [2]; the three are created right after `stream`, so
    // that the runtime puts the four streams on different hardware queues
    // even when many contexts exist (default priority).  `stream` forks into
    // lane 1 (ev_start) and joins it at the end (ev_join); a chunk's resolve
    // waits for the previous chunk's (ev_resolved), so passes are added to
    // the framebuffer in order.
    Wavefront wf;
    Wavefront wf2;
    hipStream_t lane_stream[3] = {nullptr, nullptr, nullptr};
    hipEvent_t lane_fork[2] = {nullptr, nullptr}, lane_shaded[2] = {nullptr, nullptr};
    hipEvent_t lane_done[2] = {nullptr, nullptr}, ev_resolved[2] = {nullptr, nullptr};
    hipEvent_t ev_start = nullptr, ev_join = nullptr;
    // the second pipeline of a batched launch (run_jobs) keeps its state here
    // and runs on lane_stream[1] with wf2
    BatchState batch2;
    KernelTimer timer;
    LaunchCfg launch;
    BatchState batch;
};

// kernel timing helpers (wavefront.hip)
void timer_begin(Ctx *c, int kernel, hipEvent_t *start, hipStream_t stream);
void timer_end(Ctx *c, int kernel, hipEvent_t start, hipStream_t stream);
void timer_collect(Ctx *c);
void timer_release(Ctx *c);
extern const char *const kKernelNames[KT_COUNT];

// bvh_build.hip
void build_bvh(Ctx *c, int width, int max_leaf);
// wavefront.hip
struct PathSource {
    int mode;              // 0: render passes, 1: explicit (pixel, sample) lists
    int first_pass;
    int ls = 1;            // mode 0: paths per pixel in passes after the first (lightsamples)
    const int *pixel;      // device, mode 1
    const int *sample;     // device, mode 1
    const float *slots;    // device, optional injected uniforms
    // mode 0 with ls = 1: per-path records for the adjoint:
    // rec[rec_base + path] = (unclamped radiance, clamp steps | 0x10000 for a
    // camera miss | rays traced << 17), path = pass * w * h + tiled position;
    // only indices in [0, rec_count) are written.  rec_only: records instead
    // of the framebuffer (the adjoint's first replay); otherwise both (a
    // forward render keeping its records).
    float4 *adj_rec = nullptr;
    long long rec_base = 0;
    long long rec_count = 0;
    bool rec_only = false;
    // mode 0: the FP16 running mean updated with every resolved pass, or null;
    // display_base passes were appended to it before this run
    ushort4 *display = nullptr;
    long long display_base = 0;
};
void run_paths(Ctx *c, const PathSource &src, size_t n_paths, float4 *out_rgba, aur_stats *stats);
// Paths of passes [first_pass, first_pass + n_passes) of a w*h image: one per
// pixel in pass 0, ls per pixel after it (renderer.cpp:233, 257).
inline size_t pass_paths(size_t wh, int first_pass, int n_passes, int ls) {
    if (n_passes <= 0) return 0;
    size_t n = 0;
    int p = first_pass;
    if (p == 0) { n += wh; ++p; --n_passes; }
    return n + wh * (size_t)ls * (size_t)n_passes;
}
void resolve_passes(Ctx *c, const float4 *out_rgba, int n_passes, float4 *accum);
// mode: AUR_TRACE_CLOSEST, AUR_TRACE_ANY or AUR_TRACE_EXTEND
void trace_rays(Ctx *c, const aur_ray *d_rays, int n, aur_hit *d_hits, int mode);
// the allocation-free variant (closest / any hit) on the file-scope traversal
// scratch; noalloc_init looks up that scratch's address on the current device
void noalloc_init(int device);
void trace_rays_noalloc(Ctx *c, const aur_ray *d_rays, int n, aur_hit *d_hits, int mode);
// sizes the path pool and the traversal stack overflow for BVHs of the given width and depth
void ensure_wavefront_for(Ctx *c, size_t cap, int width, int depth);
// creates the lane streams and events of run_paths (aur_create, after `stream`)
void ensure_lane_streams(Ctx *c);
// the same for a given wavefront (lane 1 of a batched launch uses wf2)
void ensure_wavefront_on(Ctx *c, Wavefront &w, size_t cap, int width, int depth);
// batch.hip: jobs[k] renders passes [first_pass[k], first_pass[k] + n_passes[k]) into its framebuffer
void run_jobs(Ctx *c, Ctx *const *jobs, const int *first_pass, const int *n_passes, int n_jobs, int order, int use_graph,
              int iters_per_poll, aur_job_result *results, aur_stats *total_stats);
// adjoint.hip
void adjoint(Ctx *c, int first_pass, int n_passes, const float4 *d_dldi, float *d_galb, float *d_grough, float *d_gspec,
             float *d_gptex, aur_stats *stats);
// parity.hip
void bsdf_queries(Ctx *c, const aur_bsdf_query *d_q, int n, aur_bsdf_result *d_r);
void light_queries(Ctx *c, const aur_light_query *d_q, int n, aur_light_result *d_r);
// the Monte-Carlo estimate of the table; half_store rounds each entry to FP16
void kelemen_table(Ctx *c, float reflectance, int dist, int n_samples, uint32_t seed, float *d_out,
                   bool half_store = false);
// kernels used for device-side clears (graph safe, no hipMemsetAsync)
void fill_zero(Ctx *c, void *p, size_t bytes);

}  // namespace aur
