// aurora_gpu.hpp -- C++ host side of the MI355X path tracer over its C ABI
// (include/aurora_hip.h).  Header only; link with libaurora_hip.so.
//
// The classes mirror the interfaces of the reference's render path so that
// a C++ caller written against them keeps its shape:
//   * Renderer: Renderer(file) and render(), which runs parseSceneDescription,
//     buildRenderEnvironment, renderImage, outputStats and postRenderCleanup
//     (src/frontend/renderer.h:32-56, renderer.cpp:32-60, 201-310);
//   * AccelerationStructure: intersect / intersectBinary
//     (src/geometry/accelerationStructure.h:17-27), over batches of rays;
//   * Brdf and Light queries (brdf.h:17-40, light.h:23-81) in batches.
// Errors: the reference logs and carries on (LOG_ERROR is a no-op assert in
// release builds); here every failing call throws aurora_gpu::Error with the
// library's code and message, so a caller sees the first failure.
#pragma once

#include <chrono>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/aurora_hip.h"

namespace aurora_gpu {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc, const char *what) {
    if (rc != AUR_OK) {
        const char *m = aur_last_error();
        throw Error(rc, std::string(what) + ": " + (m ? m : "unknown error"));
    }
}

// One library context on one HIP device (not shared between host threads).
class Context {
public:
    explicit Context(int device = 0) { check(aur_create(&ctx_, device), "aur_create"); }
    ~Context() { aur_destroy(ctx_); }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    aur_ctx *get() const { return ctx_; }

private:
    aur_ctx *ctx_ = nullptr;
};

// Renderer (renderer.h:32-56): the scene file in, the EXR named by its
// options out.
class Renderer {
public:
    // renderer.cpp:32-37; config: the config.json that resolves $BASE,
    // $RENDER, ... in the scene's paths (empty: the one next to the scene)
    explicit Renderer(std::string file, std::string config = "", int device = 0)
        : m_filename(std::move(file)), m_config(std::move(config)), m_ctx(device) {}

    // renderer.cpp:39-60: parse, build (timed), render, stats, cleanup
    void render() {
        parseSceneDescription();
        const auto t0 = std::chrono::steady_clock::now();
        buildRenderEnvironment();
        const long secs = (long)std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("Total time building render env: %ld h %ld min %ld sec.\n", secs / 3600, (secs / 60) % 60,
                    secs % 60);
        renderImage();
        outputStats();
        postRenderCleanup();
    }

    // The parts, callable on their own.
    void parseSceneDescription() {
        check(aur_scene_load(m_ctx.get(), m_filename.c_str(), m_config.empty() ? nullptr : m_config.c_str()),
              "aur_scene_load");
    }
    void buildRenderEnvironment(int bvh_width = 4, int max_leaf = 2) {
        aur_accel_opts o = {bvh_width, max_leaf};
        check(aur_build_accel(m_ctx.get(), &o), "aur_build_accel");
    }
    // renderImage + the final draw of postRenderCleanup: all passes, the
    // display written after pass 0, every `draw_delay_s` seconds (<= 0: the
    // reference's DRAW_DELAY) and at the end with renderTime / RaySpeed
    void renderImage(const char *path = nullptr, double draw_delay_s = 0.0) {
        check(aur_render_image(m_ctx.get(), path, draw_delay_s, &m_stats, &m_draws), "aur_render_image");
    }
    void outputStats() const {
        const double s = m_stats.ms * 1e-3;
        std::printf("Rendered %lld paths, %lld rays in %.3f s (%.0fK rays/sec.), %d image writes.\n",
                    (long long)m_stats.paths, (long long)(m_stats.ext_rays + m_stats.shadow_rays), s,
                    s > 0 ? 1e-3 * (double)(m_stats.ext_rays + m_stats.shadow_rays) / s : 0.0, m_draws);
    }
    void postRenderCleanup() {}

    const aur_stats &stats() const { return m_stats; }
    int draws() const { return m_draws; }
    aur_ctx *context() const { return m_ctx.get(); }

private:
    std::string m_filename, m_config;
    Context m_ctx;
    aur_stats m_stats = {};
    int m_draws = 0;
};

// AccelerationStructure::intersect / intersectBinary
// (accelerationStructure.h:17-27) over a built scene, in batches: per-ray
// calls would be dominated by launch latency.
class AccelerationStructure {
public:
    explicit AccelerationStructure(aur_ctx *ctx) : m_ctx(ctx) {}
    // closest hit: t, u, v and the global triangle id (-1 on a miss)
    std::vector<aur_hit> intersect(const std::vector<aur_ray> &rays) const { return trace(rays, AUR_TRACE_CLOSEST); }
    // any hit (shadow rays): prim >= 0 when the segment is blocked
    std::vector<bool> intersectBinary(const std::vector<aur_ray> &rays) const {
        std::vector<aur_hit> h = trace(rays, AUR_TRACE_ANY);
        std::vector<bool> out(h.size());
        for (size_t i = 0; i < h.size(); ++i) out[i] = h[i].prim >= 0;
        return out;
    }

private:
    std::vector<aur_hit> trace(const std::vector<aur_ray> &rays, int mode) const {
        std::vector<aur_hit> hits(rays.size());
        if (!rays.empty())
            check(aur_trace(m_ctx, rays.data(), (int)rays.size(), hits.data(), mode), "aur_trace");
        return hits;
    }
    aur_ctx *m_ctx;
};

// Brdf::getSample / evalSampleWorld / pdf (brdf.h:17-40) and
// Light::generateSample / eval / pdf / visible (light.h:23-81), in batches.
inline std::vector<aur_bsdf_result> brdfQueries(aur_ctx *ctx, const std::vector<aur_bsdf_query> &q) {
    std::vector<aur_bsdf_result> r(q.size());
    if (!q.empty()) check(aur_bsdf(ctx, q.data(), (int)q.size(), r.data()), "aur_bsdf");
    return r;
}
inline std::vector<aur_light_result> lightQueries(aur_ctx *ctx, const std::vector<aur_light_query> &q) {
    std::vector<aur_light_result> r(q.size());
    if (!q.empty()) check(aur_light_queries(ctx, q.data(), (int)q.size(), r.data()), "aur_light_queries");
    return r;
}

}  // namespace aurora_gpu
