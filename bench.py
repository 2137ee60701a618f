#!/usr/bin/env python3
"""Benchmark: forward path-tracing throughput on one MI355X.

Workload: the 1920x1080, 64 spp, max-depth-8 scene of eight displaced, Loop-
subdivided (7 levels) icospheres (2,621,450 triangles with the light quads),
Kelemen/GGX materials, four quad lights and a constant environment.  A step is
one batch of ``--passes-per-step`` full-frame passes (each pass = one path per
pixel); by default a step is the whole 64-spp frame (64 passes, 132.7 M
paths, traced as one chunk), rendered twice after one warm-up frame.  Step k
renders passes (k * P) mod 64 onward, so with fewer passes per step the steps
walk through the frame's passes.

Prints one JSON line.  ``value`` is Mpaths/s of the timed steps with the scene
and BVH resident in HBM (a bounce's shadow and finish kernels run on a second
stream beside the next bounce's extension kernel).  ``roofline`` describes the
kernel with the largest share of device time, timed with HIP events on its
stream in one more step whose kernels run one after another, so that each
kernel's time is its own.
``cpu_baseline`` is the CPU restatement of the reference algorithm (the
oracle, with its own SAH BVH and OpenMP tiles) timed on this host on a bounded
sample of the same frame.  ``adjoint`` is the inverse-rendering workload (the
Kelemen box room at 512x512, 32 spp forward and adjoint, depth 6): forward and
adjoint times, adjoint Mpaths/s, and the largest relative error of the adjoint
gradient against central differences on 64 seeded parameters.  ``batch`` is
the batched uneven workload (64 seeded scenes of 1e3..5e5 triangles at
128^2..1024^2, 4..256 spp, depth 3..12): aggregate Mpaths/s of the batched
persistent launch and of per-scene launches, and per-job latency p50/p99.
``box_room`` is the CPU-runnable 256x256, 16 spp, depth-5 box room: GPU and
host-CPU Mpaths/s on the same frame, with both images' RMSE against a
4096-spp oracle reference.  ``large_scene`` is the 3840x2160, depth-16,
18.5M-triangle differentiable scene: forward and texture-adjoint Mpaths/s and
a central-difference check of the 2048x2048 albedo texture's gradient.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "Mpaths/s @1080p depth 8 vs host CPU; adjoint Mpaths/s at grad rel-err <1e-2"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per processed item for the wavefront kernels (FP32 SoA):
#   k_extend: ray record read (32) + hit record written (16)
#   k_shadow: shadow record read (48)
#   k_shade:  ray (32) + hit (16) + path state (64) read, path state (64) +
#             next ray (32) written, per extension ray
ALGO_BYTES = {"k_extend": 48.0, "k_shadow": 48.0, "k_shade": 208.0}


def _progress(msg: str):
    print("bench: " + msg, file=sys.stderr, flush=True)


def _hip_sync():
    lib = ctypes.CDLL("libamdhip64.so")
    lib.hipDeviceSynchronize.restype = ctypes.c_int
    rc = lib.hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError("hipDeviceSynchronize failed: %d" % rc)


def _host_threads() -> int:
    """Every host core this process may run on (its CPU affinity), limited
    by OMP_NUM_THREADS where the host sets it to the job's CPU share."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    try:
        omp = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        omp = 0
    return max(1, min(n, omp) if omp > 0 else n)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _load_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed counter profile
    (profiles/traffic.json: FETCH_SIZE and WRITE_SIZE of a profiled run of
    this benchmark), and where that figure comes from.  The counters cannot be
    read inside this run (a counter pass is a separate profiled process), so
    the figure belongs to the profiled build and launch shape named there."""
    path = os.path.join(HERE, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None, None
    try:
        with open(path) as fh:
            t = json.load(fh)
        v = t.get(kernel, {}).get("hbm_bytes_per_launch")
        src = "profiles/traffic.json: " + t.get("source", "rocprofv3 FETCH_SIZE + WRITE_SIZE per dispatch")
        return (float(v), src) if v is not None else (None, None)
    except (OSError, ValueError, AttributeError):
        return None, None


def cpu_baseline(exp: dict, target_s: float, threads: int, spp: int):
    """Times the oracle on an evenly spaced subset of the frame's rows, taken
    across the frame's passes (row g of the stacked passes is pass g // H, row
    g % H), sized to about target_s seconds of CPU work."""
    from oracle import oracle as O
    O.build()
    t0 = time.time()
    scene = O.Scene(exp, bvh=True)
    build_s = time.time() - t0
    _progress("oracle BVH built in %.1f s" % build_s)
    W, H = exp["width"], exp["height"]

    def sample(n_rows):
        V = H * spp
        n_rows = max(1, min(V, n_rows))
        g = (np.arange(n_rows, dtype=np.int64) * V) // n_rows
        rows, passes = g % H, g // H
        pix = (rows[:, None] * W + np.arange(W)[None, :]).reshape(-1).astype(np.int32)
        smp = np.repeat(passes, W).astype(np.int32)
        return pix, smp, n_rows

    # calibrate on growing samples until one takes a measurable time, then
    # size the timed sample from that rate
    n_rows = 16
    while True:
        pix, smp, n_rows = sample(n_rows)
        t0 = time.time()
        scene.render_paths(7, pix, smp, threads=threads)
        dt = time.time() - t0
        if dt >= 0.05 * target_s or n_rows >= H * spp:
            break
        n_rows *= 4
    rate = pix.size / max(dt, 1e-6)
    pix, smp, n_rows = sample(int(rate * target_s / W))
    t0 = time.time()
    _, cnt = scene.render_paths(7, pix, smp, threads=threads)
    dt = time.time() - t0
    return {"value": round(pix.size / dt / 1e6, 4), "unit": "Mpaths/s", "cores": threads, "kind": "port",
            "sample": "%d evenly spaced rows of the %d stacked passes of the same %dx%d depth-%d frame "
                      "(%d paths, %.1f s); oracle BVH build %.1f s excluded; host %s, nproc %d"
                      % (n_rows, spp, W, H, exp["maxdepth"], pix.size, dt, build_s, _cpu_model(),
                         os.cpu_count() or 0),
            "mrays_per_s": round((cnt["ext_rays"] + cnt["shadow_rays"]) / dt / 1e6, 4)}


def box_room_leg(threads: int, reps: int = 32, ref_spp: int = 4096, stride: int = 4, n_seeds: int = 8):
    """The CPU-runnable box room (256x256, 16 spp, depth 5): GPU Mpaths/s over
    repeated full frames, the oracle's Mpaths/s on the host cores (median of
    three full frames after a warm-up), and the RMSE of 16-spp images (mean
    squared error over n_seeds seeds on each side) against a 4096-spp oracle
    reference (independent seed, float64 mean) on every 4th pixel of every
    4th row."""
    from aurora_amd import scenes, tuning
    from aurora_amd.renderer import Renderer
    from oracle import oracle as O
    O.build()
    tmp = tempfile.mkdtemp(prefix="aurora_c1_")
    path = scenes.write_config("box_room", tmp)
    r = Renderer(path, os.path.join(tmp, "config.json"))
    r.build()
    cfg = tuning.choose(r)
    if (cfg["bvh_width"], cfg["max_leaf"]) != (4, 2):
        r.build(width=cfg["bvh_width"], max_leaf=cfg["max_leaf"])
    tuning.apply_launch(r, cfg)
    info = r.info()
    W, H, spp = info["width"], info["height"], info["pixelsamples"]
    r.set_seed(11)
    r.render_passes(0, spp, fetch=False)
    _hip_sync()
    t0 = time.perf_counter()
    for k in range(reps):
        acc, st = r.render_passes(0, spp, fetch=False)
    _hip_sync()
    gpu_s = time.perf_counter() - t0
    gpu_rate = reps * st["paths"] / gpu_s / 1e6
    ys, xs = np.mgrid[0:H:stride, 0:W:stride]
    sub = (ys.ravel(), xs.ravel())
    # 16-spp images of n_seeds independent seeds on both sides: the squared
    # error is heavy-tailed (caustics), so a single image's RMSE is noisy
    gpu_imgs = []
    for k in range(n_seeds):
        r.set_seed(1000 + k)
        a, _ = r.render_passes(0, spp)
        gpu_imgs.append(a[..., :3].astype(np.float64)[sub] / spp)
    # the Halton QMC source (the reference's USE_HALTON sampler): full-frame
    # rate, and 16-spp estimates of the same pixels from disjoint pass ranges
    r.set_sampler("halton")
    r.render_passes(0, spp, fetch=False)
    _hip_sync()
    t0 = time.perf_counter()
    for k in range(reps):
        r.render_passes(0, spp, fetch=False)
    _hip_sync()
    qmc_rate = reps * W * H * spp / (time.perf_counter() - t0) / 1e6
    pix_sub = (ys * W + xs).ravel().astype(np.int32)
    qmc_imgs = []
    for k in range(n_seeds):
        P = np.repeat(pix_sub, spp)
        S = np.tile(np.arange(k * spp, (k + 1) * spp, dtype=np.int32), pix_sub.size)
        o, _ = r.render_paths(P, S)
        qmc_imgs.append(o[:, :3].astype(np.float64).reshape(pix_sub.size, spp, 3).mean(1))
    r.set_sampler("hash")
    exp = r.export()
    r.close()
    sc = O.Scene(exp, bvh=True)
    sc.render(12, 0, 1, threads=threads)  # warm-up
    times = []
    for _ in range(3):  # median of three full frames
        t0 = time.time()
        cacc, _ = sc.render(12, 0, spp, threads=threads)
        times.append(time.time() - t0)
    cpu_s = float(np.median(times))
    cpu_rate = W * H * spp / cpu_s / 1e6
    cpu_imgs = []
    for k in range(n_seeds):
        cacc, _ = sc.render(2000 + k, 0, spp, threads=threads)
        cpu_imgs.append(cacc[..., :3].astype(np.float64)[sub] / spp)
    pix = (ys * W + xs).ravel().astype(np.int32)
    ref = np.zeros((pix.size, 3))
    t0 = time.time()
    chunk = 256
    for c0 in range(0, ref_spp, chunk):
        P = np.repeat(pix, chunk)
        S = np.tile(np.arange(c0, c0 + chunk, dtype=np.int32), pix.size)
        o, _ = sc.render_paths(999, P, S, threads=threads)
        ref += o[:, :3].astype(np.float64).reshape(pix.size, chunk, 3).sum(1)
    ref /= ref_spp
    ref_s = time.time() - t0
    mse_g = np.array([np.mean((g - ref) ** 2) for g in gpu_imgs])
    mse_c = np.array([np.mean((c - ref) ** 2) for c in cpu_imgs])
    mse_q = np.array([np.mean((q - ref) ** 2) for q in qmc_imgs])
    rmse_g, rmse_c = float(np.sqrt(mse_g.mean())), float(np.sqrt(mse_c.mean()))
    rmse_q = float(np.sqrt(mse_q.mean()))
    # standard error of the difference of the two mean squared errors
    se = float(np.sqrt(mse_g.var(ddof=1) / n_seeds + mse_c.var(ddof=1) / n_seeds))
    return {"workload": "box_room_256_spp16_depth5", "paths_per_frame": W * H * spp,
            "gpu_Mpaths_s": round(gpu_rate, 3), "gpu_frames_timed": reps,
            "cpu_Mpaths_s": round(cpu_rate, 4), "cpu_cores": threads, "cpu_s": round(cpu_s, 2),
            "gpu_vs_cpu": round(gpu_rate / cpu_rate, 2),
            "rmse_gpu_vs_ref": float("%.4g" % rmse_g), "rmse_cpu_vs_ref": float("%.4g" % rmse_c),
            "mse_diff_in_se": float("%.3g" % ((mse_g.mean() - mse_c.mean()) / max(se, 1e-30))),
            "rmse_seeds": n_seeds,
            "gpu_halton_Mpaths_s": round(qmc_rate, 3), "rmse_gpu_halton_vs_ref": float("%.4g" % rmse_q),
            "halton_rmse_vs_hash": float("%.3g" % (rmse_q / max(rmse_g, 1e-30))),
            "ref": "%d-spp oracle on %d pixels (every %dth row and column), %.1f s" % (ref_spp, pix.size, stride, ref_s)}


def adjoint_leg(res: int, spp: int, depth: int, n_fd: int = 64, adam_steps: int = 0):
    """Inverse-rendering workload: forward and path-replay adjoint timing on the
    Kelemen box room, plus a central-difference check (h = 1e-3) of the
    gradients on 64 seeded parameters with common random numbers."""
    from aurora_amd import inverse
    tmp = tempfile.mkdtemp(prefix="aurora_adj_")
    p = inverse.make_problem(tmp, res=res, spp=spp, depth=depth)
    r = p.renderer
    rng = np.random.default_rng(3)
    albedo = p.truth["albedo"] * rng.uniform(0.8, 1.2, p.truth["albedo"].shape).astype(np.float32)
    rough, spec = p.init["rough"], p.init["spec"]
    r.set_seed(5)
    img = inverse.render_image(r, albedo, rough, spec, 0, spp)
    dl = (2.0 / img.size) * (img.astype(np.float64) - p.target)  # d mean((I - T)^2) / dI at this seed
    r.adjoint(0, spp, dl.astype(np.float32))
    _hip_sync()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        inverse.render_image(r, albedo, rough, spec, 0, spp)
    _hip_sync()
    fwd_s = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        ga, gr, gs, st = r.adjoint(0, spp, dl.astype(np.float32))
    _hip_sync()
    adj_s = (time.perf_counter() - t0) / reps
    # the optimisation loop's form: each forward render keeps its per-path
    # records and the adjoint of the same passes skips its first replay
    r.keep_path_records(0, spp)
    fwd_k, adj_k = 0.0, 0.0
    for _ in range(reps):
        t0 = time.perf_counter()
        inverse.render_image(r, albedo, rough, spec, 0, spp)
        _hip_sync()
        t1 = time.perf_counter()
        r.adjoint(0, spp, dl.astype(np.float32))
        _hip_sync()
        fwd_k += t1 - t0
        adj_k += time.perf_counter() - t1
    fwd_k, adj_k = fwd_k / reps, adj_k / reps
    r.keep_path_records(0, 0)

    def loss(a, sw):
        r.set_seed(5)
        im = inverse.render_image(r, a, rough, sw, 0, spp)
        return float(np.mean((im.astype(np.float64) - p.target) ** 2))

    r.set_seed(5)
    ga, gr, gs, _ = r.adjoint(0, spp, dl.astype(np.float32))
    # 64 seeded parameters: every material roughness, then spec weights and
    # albedo channels whose gradient is not negligible.  The image is linear
    # in albedo and spec weight, so central differences with common random
    # numbers are exact up to rounding; roughness moves sampled directions and
    # Kelemen lobe choices (and the reference's table lookup jumps at every
    # 1/31), so its differences (h = 1e-3, inside one table cell) are compared
    # in the mean over seeds.
    rough_k = [int(k) for k in np.flatnonzero(np.abs(gr) > 1e-3 * np.abs(gr).max())]
    cand = [("s", int(k)) for k in np.flatnonzero(np.abs(gs) > 1e-3 * np.abs(gs).max())]
    cand_a = [("a", int(k)) for k in np.flatnonzero(np.abs(ga.ravel()) > 1e-3 * np.abs(ga).max())]
    cand += [cand_a[j] for j in np.random.default_rng(64).permutation(len(cand_a))]
    pick = cand[:max(0, n_fd - len(rough_k))]
    h, worst = 1e-3, 0.0
    for kind, k in pick:
        if kind == "a":
            t, c = divmod(k, 3)
            ap, am = albedo.copy(), albedo.copy()
            ap[t, c] += h
            am[t, c] -= h
            fd = (loss(ap, spec) - loss(am, spec)) / (2 * h)
            g, gmax = ga[t, c], np.abs(ga).max()
        else:
            sp, sm = spec.copy(), spec.copy()
            sp[k] += h
            sm[k] -= h
            fd = (loss(albedo, sp) - loss(albedo, sm)) / (2 * h)
            g, gmax = gs[k], np.abs(gs).max()
        worst = max(worst, abs(g - fd) / max(abs(fd), 1e-3 * gmax))
    # roughness over 8 seeds: adjoint and central differences estimate the
    # same derivative of the linearised loss <dL/dI, I> (dL/dI fixed at the
    # seed-5 image, for both sides) with different noise
    n_seeds = 8
    fdr, adr = np.zeros((n_seeds, len(rough_k))), np.zeros((n_seeds, len(rough_k)))
    for si in range(n_seeds):
        r.set_seed(40 + si)
        _, gr_s, _, _ = r.adjoint(0, spp, dl.astype(np.float32))
        for j, k in enumerate(rough_k):
            adr[si, j] = gr_s[k]
            rp, rm = rough.copy(), rough.copy()
            rp[k] += h
            rm[k] -= h
            vals = []
            for rv in (rp, rm):
                r.set_seed(40 + si)
                im = inverse.render_image(r, albedo, rv, spec, 0, spp)
                vals.append(float(np.sum(dl * im.astype(np.float64))))
            fdr[si, j] = (vals[0] - vals[1]) / (2 * h)
    se_r = np.sqrt(fdr.var(0, ddof=1) / n_seeds + adr.var(0, ddof=1) / n_seeds) + 1e-30
    z_r = np.abs(fdr.mean(0) - adr.mean(0)) / se_r
    rel_r = np.abs(fdr.mean(0) - adr.mean(0)) / np.maximum(np.abs(fdr.mean(0)), 1e-30)
    n_paths = res * res * spp
    adam = None
    if adam_steps:
        t0 = time.perf_counter()
        hist, prm = inverse.optimize(p, steps=adam_steps, lr=0.02)
        adam_s = time.perf_counter() - t0
        seen = np.abs(prm["albedo"] - p.init["albedo"]).sum(1) > 0
        adam = {"steps": adam_steps, "s": round(adam_s, 2), "loss_first": float("%.5g" % hist[0]),
                "loss_last": float("%.5g" % hist[-1]),
                "albedo_mae_init": round(float(np.abs(p.init["albedo"] - p.truth["albedo"])[seen].mean()), 4),
                "albedo_mae_final": round(float(np.abs(prm["albedo"] - p.truth["albedo"])[seen].mean()), 4)}
    r.close()
    return {"adam": adam, "workload": "box_room_kelemen_%d_spp%d_depth%d" % (res, spp, depth), "params": int(ga.size + gr.size + gs.size),
            "forward_ms": round(fwd_s * 1e3, 3), "adjoint_ms": round(adj_s * 1e3, 3),
            "adjoint_Mpaths_s": round(n_paths / adj_s / 1e6, 3), "adjoint_vs_forward": round(adj_s / fwd_s, 3),
            "kept_records": {"forward_ms": round(fwd_k * 1e3, 3), "adjoint_ms": round(adj_k * 1e3, 3),
                             "adjoint_Mpaths_s": round(n_paths / adj_k / 1e6, 3),
                             "adjoint_vs_forward": round(adj_k / fwd_k, 3)},
            "fd_checked": int(len(pick)) + len(rough_k), "fd_h": h, "fd_max_rel_err": float("%.3g" % worst),
            "fd_albedo_spec_checked": int(len(pick)), "fd_roughness_checked": len(rough_k),
            "fd_roughness_seeds": n_seeds, "fd_roughness_max_z": float("%.3g" % (z_r.max() if len(z_r) else 0.0)),
            "fd_roughness_rel_err_of_means": [float("%.3g" % x) for x in rel_r]}


def large_leg(fwd_passes: int = 256, adj_passes: int = 64, fd_passes: int = 1, n_fd: int = 64):
    """The large differentiable scene (3840x2160, depth 16, 32 displaced level-7
    icospheres on 8 terrain patches of 1M triangles: 18.5M triangles; a
    2048x2048 RGB FP32 albedo texture on the terrain): scene build time,
    forward Mpaths/s over fwd_passes passes, path-replay adjoint Mpaths/s of
    the texture gradient over adj_passes passes, and central differences
    (common random numbers, h = 1e-3) on n_fd seeded texels at fd_passes
    passes."""
    from aurora_amd import scenes, tuning
    from aurora_amd.renderer import Renderer
    tmp = tempfile.mkdtemp(prefix="aurora_large_")
    t0 = time.time()
    spec, truth = scenes.large_scene()
    path = spec.write(tmp)
    gen_s = time.time() - t0
    _progress("large scene written in %.1f s" % gen_s)
    t0 = time.time()
    r = Renderer(path, os.path.join(tmp, "config.json"))
    load_s = time.time() - t0
    r.build(width=4, max_leaf=1)
    _hip_sync()
    build_s = time.time() - t0 - load_s
    info = r.info()
    # the tuned BVH structure and launch shape for this triangle count (Loop
    # subdivision happens in build, so the table is read after it); when they
    # differ from the first build's, the scene is built again and that build,
    # the one the legs below run on, is the one timed
    cfg = tuning.choose(r)
    if (cfg["bvh_width"], cfg["max_leaf"]) != (4, 1):
        t1 = time.time()
        r.build(width=cfg["bvh_width"], max_leaf=cfg["max_leaf"])
        _hip_sync()
        build_s = time.time() - t1
    _progress("large scene: %d triangles loaded in %.1f s, built in %.1f s (BVH%d, %d per leaf)"
              % (info["n_tris"], load_s, build_s, cfg["bvh_width"], cfg["max_leaf"]))
    tuning.apply_launch(r, cfg)
    attrs = [spec.attr_index(k) for k in sorted(spec.objects) if k.startswith("terrain")]
    W, H = info["width"], info["height"]
    # the inverse problem's state: texture at 0.5, loss against the true texture's image
    r.set_albedo_texture(truth, attrs)
    r.set_seed(21)
    target, _ = r.render_passes(0, fd_passes)
    tex = np.full_like(truth, 0.5)
    r.set_albedo_texture(tex, attrs)
    r.set_seed(21)
    img, _ = r.render_passes(0, fd_passes)
    dl = ((2.0 / (W * H * 3)) * (img[..., :3].astype(np.float64) - target[..., :3]) / fd_passes).astype(np.float32)
    # forward rate
    r.render_passes(0, 1, fetch=False)
    _hip_sync()
    t0 = time.perf_counter()
    _, fst = r.render_passes(0, fwd_passes, fetch=False)
    _hip_sync()
    fwd_s = time.perf_counter() - t0
    _progress("large scene: %d forward passes in %.2f s" % (fwd_passes, fwd_s))
    # adjoint rate (texture gradient only); the wavefront kernels' time inside
    # the call is its first replay (radiance per path), the rest its gradient replay
    r.adjoint(0, 1, dl, albedo=False, texture=True)
    _hip_sync()
    t0 = time.perf_counter()
    _, _, _, ast, gt_rate = r.adjoint(0, adj_passes, dl, albedo=False, texture=True)
    _hip_sync()
    adj_s = time.perf_counter() - t0
    # the split, from the same adjoint with kernel timing on (kernels one
    # after another)
    r.set_profiling(True)
    _, _, _, ast_p, _ = r.adjoint(0, adj_passes, dl, albedo=False, texture=True)
    first_replay_ms = sum(v["total_ms"] for v in r.kernel_stats().values())
    r.set_profiling(False)
    _progress("large scene: %d adjoint passes in %.2f s" % (adj_passes, adj_s))
    # the same with the forward render keeping its per-path records for the
    # adjoint's passes (16 B per path), which then skips its first replay
    r.keep_path_records(0, adj_passes)
    t0 = time.perf_counter()
    r.render_passes(0, fwd_passes, fetch=False)
    _hip_sync()
    fwd_k = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, _, _, kst, gt_k = r.adjoint(0, adj_passes, dl, albedo=False, texture=True)
    _hip_sync()
    adj_k = time.perf_counter() - t0
    r.keep_path_records(0, 0)
    kept_diff = float(np.abs(gt_k - gt_rate).max() / max(np.abs(gt_rate).max(), 1e-30))
    _progress("large scene: kept records, forward %.2f s, adjoint %.2f s" % (fwd_k, adj_k))
    # gradient check at fd_passes passes
    r.set_seed(21)
    _, _, _, _, gt = r.adjoint(0, fd_passes, dl * fd_passes, albedo=False, texture=True)
    flat = np.abs(gt).reshape(-1)
    cand = np.flatnonzero(flat > 1e-2 * flat.max())
    pick = np.random.default_rng(64).permutation(cand)[:n_fd]
    h, worst = 1e-3, 0.0

    def loss(t):
        r.set_albedo_texture(t, attrs)
        r.set_seed(21)
        im, _ = r.render_passes(0, fd_passes)
        return float(np.sum(dl.astype(np.float64) * im[..., :3].astype(np.float64)))

    base = tex.reshape(-1)
    for k in pick:
        tp, tm = base.copy(), base.copy()
        tp[k] += h
        tm[k] -= h
        fd = (loss(tp.reshape(tex.shape)) - loss(tm.reshape(tex.shape))) / (2 * h)
        g = float(gt.reshape(-1)[k])
        worst = max(worst, abs(g - fd) / max(abs(fd), 1e-3 * float(flat.max())))
    r.close()
    return {"workload": "large_scene_3840x2160_depth16", "triangles": int(info["n_tris"]),
            "texture": [int(truth.shape[1]), int(truth.shape[0])], "scene_write_s": round(gen_s, 1),
            "load_s": round(load_s, 2), "bvh_build_s": round(build_s, 2),
            "launch": {k: cfg[k] for k in ("bvh_width", "max_leaf", "trace_block", "trace_waves_per_cu", "pool_paths")},
            "forward_passes": fwd_passes, "forward_Mpaths_s": round(fst["paths"] / fwd_s / 1e6, 3),
            "forward_Mrays_s": round((fst["ext_rays"] + fst["shadow_rays"]) / fwd_s / 1e6, 3),
            "adjoint_passes": adj_passes, "adjoint_Mpaths_s": round(ast["paths"] / adj_s / 1e6, 3),
            "adjoint_vs_forward": round((adj_s / adj_passes) / (fwd_s / fwd_passes), 3),
            "adjoint_first_replay_ms": round(first_replay_ms, 1),
            "adjoint_gradient_replay_ms": round(ast_p["ms"] - first_replay_ms, 1),
            "kept_records": {"forward_Mpaths_s": round(fst["paths"] / fwd_k / 1e6, 3),
                             "adjoint_Mpaths_s": round(kst["paths"] / adj_k / 1e6, 3),
                             "adjoint_vs_forward": round((adj_k / adj_passes) / (fwd_k / fwd_passes), 3),
                             "record_bytes": int(16 * info["width"] * info["height"] * adj_passes),
                             "gradient_max_rel_diff": float("%.3g" % kept_diff)},
            "texels_with_gradient": int(np.count_nonzero(gt_rate.reshape(-1, 3).any(1))),
            "fd_checked": int(len(pick)), "fd_h": h, "fd_passes": fd_passes, "fd_max_rel_err": float("%.3g" % worst)}


def batch_leg(n_jobs: int = 64, seed: int = 3, use_graph: bool = True):
    """Batched uneven workload: the same jobs rendered once through the batched
    persistent launch and once as per-scene launches (one aur_render per job,
    in the same order), with per-job latency from the start of the run."""
    from aurora_amd import scenes, tuning
    from aurora_amd.renderer import Renderer, batch_submit
    tmp = tempfile.mkdtemp(prefix="aurora_batch_")
    paths = scenes.batch_jobs(tmp, n_jobs=n_jobs, seed=seed)
    cfg = tuning.choose_batch()
    t0 = time.time()
    rs = []
    for k, p in enumerate(paths):
        r = Renderer(p, os.path.join(os.path.dirname(p), "config.json"))
        r.build(width=cfg["bvh_width"], max_leaf=cfg["max_leaf"])
        tuning.apply_launch(r, cfg)
        r.set_seed(1000 + k)
        rs.append(r)
    build_s = time.time() - t0
    info = [r.info() for r in rs]
    spp = [i["pixelsamples"] for i in info]
    tris = [i["n_tris"] for i in info]
    _progress("batch: %d scenes built in %.1f s (%d..%d triangles)" % (n_jobs, build_s, min(tris), max(tris)))
    ex = Renderer()
    tuning.apply_launch(ex, cfg)
    pool = int(cfg["pool_paths"])
    # warm both launch paths with their full working sets, so that no buffer
    # grows inside a timed region: the same batched launch (each pipeline's
    # path pool and ring are sized by the jobs it is given), and per job the
    # passes of one whole chunk (up to `pool` paths)
    batch_submit(ex, [(r, 0, n) for r, n in zip(rs, spp)], use_graph=use_graph)
    for r, i, n in zip(rs, info, spp):
        r.render_passes(0, min(n, max(1, pool // (i["width"] * i["height"]))), fetch=False)
    _hip_sync()
    t0 = time.perf_counter()
    res, tot = batch_submit(ex, [(r, 0, n) for r, n in zip(rs, spp)], use_graph=use_graph)
    batched_s = time.perf_counter() - t0
    _progress("batch: batched launch %.2f s" % batched_s)
    order = sorted(range(n_jobs), key=lambda k: res[k]["order"])
    check = set(order[:2] + order[-2:])
    fb = {k: rs[k].framebuffer() for k in check}
    lat_single = [0.0] * n_jobs
    worst = 0.0
    t0 = time.perf_counter()
    for k in order:
        acc, _ = rs[k].render_passes(0, spp[k], fetch=k in check)
        lat_single[k] = (time.perf_counter() - t0) * 1e3
        if k in check:
            worst = max(worst, float(np.max(np.abs(acc - fb[k]) / (np.abs(acc) + 1e-3))))
    single_s = time.perf_counter() - t0
    _progress("batch: per-scene launches %.2f s" % single_s)
    # What batching can remove: a job's per-scene time is its paths' marginal
    # cost plus a fixed part (launch gaps, the drain of its last bounces while
    # the queues no longer fill the GPU).  Rendering each job at spp and at
    # 2 spp passes separates the two: the marginal time of the second spp
    # passes is what the job costs with the fixed part amortised, so
    # sum(per-scene) / sum(marginal) bounds the speedup of any launch scheme
    # that does not make the paths themselves cheaper.
    t_once, t_twice = 0.0, 0.0
    for k in order:
        rs[k].render_passes(0, 2 * spp[k], fetch=False)  # grows the context's buffers untimed
        _hip_sync()
        t1 = time.perf_counter()
        rs[k].render_passes(0, spp[k], fetch=False)
        _hip_sync()
        t2 = time.perf_counter()
        rs[k].render_passes(0, 2 * spp[k], fetch=False)
        _hip_sync()
        t3 = time.perf_counter()
        t_once += t2 - t1
        t_twice += t3 - t2
        rs[k].close()
    marginal_s = max(t_twice - t_once, 1e-9)
    _progress("batch: per-scene %.2f s, marginal %.2f s" % (t_once, marginal_s))
    # the reference's launch granularity: one launch per scene and pass
    # (Renderer::renderImage runs one parallel_for per pass), on the first
    # third of the jobs in launch order, rebuilt since the others were closed
    sub = order[: n_jobs // 3]
    t_pp = 0.0
    paths_pp = 0
    rs_sub = []
    for k in sub:
        r = Renderer(paths[k], os.path.join(os.path.dirname(paths[k]), "config.json"))
        r.build(width=cfg["bvh_width"], max_leaf=cfg["max_leaf"])
        tuning.apply_launch(r, cfg)
        r.set_seed(1000 + k)
        r.render_passes(0, 1, fetch=False)
        _hip_sync()
        t0 = time.perf_counter()
        for p in range(spp[k]):
            r.render_passes(p, 1, fetch=False)
        t_pp += time.perf_counter() - t0
        paths_pp += spp[k] * info[k]["width"] * info[k]["height"]
        rs_sub.append((r, spp[k]))
    _progress("batch: per-pass launches of %d jobs %.2f s" % (len(sub), t_pp))
    # the same jobs through one batched launch
    ex2 = Renderer()
    tuning.apply_launch(ex2, cfg)
    batch_submit(ex2, [(r, 0, n) for r, n in rs_sub], use_graph=use_graph)
    _hip_sync()
    t0 = time.perf_counter()
    _, tot_sub = batch_submit(ex2, [(r, 0, n) for r, n in rs_sub], use_graph=use_graph)
    t_bsub = time.perf_counter() - t0
    for r, _ in rs_sub:
        r.close()
    ex2.close()
    ex.close()
    lat_b = np.array([o["latency_ms"] for o in res])
    lat_s = np.array(lat_single)
    total = int(tot["paths"])
    rays = int(tot["ext_rays"] + tot["shadow_rays"])
    return {"workload": "%d seeded jobs (seed %d): %d..%d triangles, 128^2..1024^2, spp 4..256, depth 3..12"
                        % (n_jobs, seed, min(tris), max(tris)),
            "paths": total, "rays": rays, "scene_build_s": round(build_s, 2),
            "pool_paths": cfg["pool_paths"], "bvh_width": cfg["bvh_width"], "max_leaf": cfg["max_leaf"],
            "trace_block": cfg["trace_block"], "trace_waves_per_cu": cfg["trace_waves_per_cu"], "graph_replay": use_graph,
            "batched_Mpaths_s": round(total / batched_s / 1e6, 3), "per_scene_Mpaths_s": round(total / single_s / 1e6, 3),
            "batched_Mrays_s": round(rays / batched_s / 1e6, 3),
            "speedup": round(single_s / batched_s, 3),
            "batched_latency_ms_p50": round(float(np.percentile(lat_b, 50)), 2),
            "batched_latency_ms_p99": round(float(np.percentile(lat_b, 99)), 2),
            "per_scene_latency_ms_p50": round(float(np.percentile(lat_s, 50)), 2),
            "per_scene_latency_ms_p99": round(float(np.percentile(lat_s, 99)), 2),
            "image_max_rel_diff_4_jobs": float("%.3g" % worst),
            "per_pass_launch_jobs": len(sub), "per_pass_launch_Mpaths_s": round(paths_pp / max(t_pp, 1e-9) / 1e6, 3),
            "batched_same_jobs_Mpaths_s": round(tot_sub["paths"] / max(t_bsub, 1e-9) / 1e6, 3),
            "speedup_vs_per_pass": round(t_pp / max(t_bsub, 1e-9), 3), "per_pass_launch_paths": int(paths_pp),
            "bound": {"per_scene_s": round(t_once, 3), "marginal_s": round(marginal_s, 3),
                      "max_speedup": round(t_once / marginal_s, 3),
                      "batched_vs_marginal": round(batched_s / marginal_s, 3)}}


def copy_peak_gbps(mib: int = 2048, reps: int = 5) -> float:
    """Achievable HBM bandwidth on this GPU: a device-to-device copy of `mib`
    MiB (read + write bytes per second), best of `reps`, timed with HIP
    events on the null stream."""
    hip = ctypes.CDLL("libamdhip64.so")
    vp = ctypes.c_void_p
    for f, args in (("hipMalloc", [ctypes.POINTER(vp), ctypes.c_size_t]), ("hipFree", [vp]),
                    ("hipMemcpy", [vp, vp, ctypes.c_size_t, ctypes.c_int]),
                    ("hipEventCreate", [ctypes.POINTER(vp)]), ("hipEventRecord", [vp, vp]),
                    ("hipEventSynchronize", [vp]), ("hipEventDestroy", [vp]),
                    ("hipEventElapsedTime", [ctypes.POINTER(ctypes.c_float), vp, vp])):
        getattr(hip, f).restype = ctypes.c_int
        getattr(hip, f).argtypes = args

    def ok(rc):
        if rc != 0:
            raise RuntimeError("HIP error %d in the copy measurement" % rc)

    n = mib << 20
    a, b, e0, e1 = vp(), vp(), vp(), vp()
    ok(hip.hipMalloc(ctypes.byref(a), n))
    ok(hip.hipMalloc(ctypes.byref(b), n))
    ok(hip.hipEventCreate(ctypes.byref(e0)))
    ok(hip.hipEventCreate(ctypes.byref(e1)))
    best = float("inf")
    try:
        ok(hip.hipMemcpy(b, a, n, 3))  # hipMemcpyDeviceToDevice, warm
        for _ in range(reps):
            ok(hip.hipEventRecord(e0, None))
            ok(hip.hipMemcpy(b, a, n, 3))
            ok(hip.hipEventRecord(e1, None))
            ok(hip.hipEventSynchronize(e1))
            ms = ctypes.c_float()
            ok(hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1))
            best = min(best, ms.value)
    finally:
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
        hip.hipFree(a)
        hip.hipFree(b)
    return 2.0 * n / (best * 1e-3) / 1e9


def copy_calibration(mib: int, reps: int):
    """A plain device copy of known size, for scaling the FETCH_SIZE / WRITE_SIZE
    counters to bytes."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 4, "copy_calibration_bytes_written": n * 4,
                      "reps": reps}), flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--passes-per-step", type=int, default=64)
    ap.add_argument("--bvh-width", type=int, default=0, help="0: autotuned choice")
    ap.add_argument("--max-leaf", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--levels", type=int, default=7)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-large", action="store_true", help="skip the 18.5M-triangle differentiable scene")
    ap.add_argument("--batch-plain-launches", action="store_true",
                    help="batched launch without graph replay (for kernel tracing)")
    ap.add_argument("--adam-steps", type=int, default=200,
                    help="Adam steps of the inverse-rendering run inside the adjoint leg (0: skip)")
    ap.add_argument("--serial-timed", action="store_true",
                    help="time the steps with kernel timing on (kernels one after another)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (counter calibration)")
    args = ap.parse_args(argv)
    if args.calibrate_copy_mib:
        return copy_calibration(args.calibrate_copy_mib, 3)

    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported "
              "(multi-GPU sharding is design-only)" % args.gpus, file=sys.stderr)
        return 2

    from aurora_amd import scenes, tuning
    from aurora_amd.renderer import Renderer

    tmp = tempfile.mkdtemp(prefix="aurora_bench_")
    scene_path = scenes.write_config("displaced_spheres", tmp, levels=args.levels)
    r = Renderer(scene_path, os.path.join(tmp, "config.json"))
    t0 = time.time()
    r.build()  # Loop subdivision happens here: the tuning table is keyed by the final triangle count
    build_s = time.time() - t0
    cfg = tuning.choose(r, width=args.bvh_width or None, max_leaf=args.max_leaf or None)
    if (cfg["bvh_width"], cfg["max_leaf"]) != (4, 2):
        t0 = time.time()
        r.build(width=cfg["bvh_width"], max_leaf=cfg["max_leaf"])
        build_s = time.time() - t0
    tuning.apply_launch(r, cfg)
    _progress("scene and BVH built in %.2f s" % build_s)
    info = r.info()
    wh = info["width"] * info["height"]
    P = args.passes_per_step

    spp = max(info["pixelsamples"], P)

    def first(k):
        return (k * P) % spp

    if args.serial_timed:
        r.set_profiling(True)  # the warm-up too, so that a trace holds one kind of step
    for k in range(args.warmup):
        r.render_passes(first(k), P, fetch=False)
    # the timed steps run as a render does (a bounce's shadow and finish
    # kernels beside the next bounce's extension kernel); with kernel timing
    # on the kernels run one after another, so the per-kernel times of the
    # roofline come from one more step timed that way (--serial-timed: the
    # timed steps themselves, e.g. under a profiler's kernel trace)
    if args.serial_timed:
        r.set_profiling(True)  # resets the totals after the warm-up
    _hip_sync()
    t0 = time.perf_counter()
    tot = {"paths": 0, "ext_rays": 0, "hit_rays": 0, "shadow_rays": 0}
    for k in range(args.steps):
        _, st = r.render_passes(first(k), P, fetch=False)
        for key in tot:
            tot[key] += st[key]
    _hip_sync()
    elapsed = time.perf_counter() - t0
    _progress("timed %d steps in %.3f s" % (args.steps, elapsed))
    kt = dict(tot)
    if not args.serial_timed:
        r.set_profiling(True)
        _, kt = r.render_passes(first(args.steps), P, fetch=False)
        _hip_sync()
    ks = r.kernel_stats()
    r.set_profiling(False)

    paths = tot["paths"]
    value = paths / elapsed / 1e6
    dom = max((k for k in ks if ks[k]["launches"] > 0), key=lambda k: ks[k]["total_ms"])
    launches = ks[dom]["launches"]
    avg_ms = ks[dom]["total_ms"] / max(launches, 1)
    items = {"k_extend": kt["ext_rays"], "k_shadow": kt["shadow_rays"], "k_shade": kt["ext_rays"]}.get(dom, 0)
    per_launch_bytes = ALGO_BYTES.get(dom, 0.0) * items / max(launches, 1)
    achieved = per_launch_bytes / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    traffic, traffic_src = _load_traffic(dom)
    copy_peak = copy_peak_gbps()
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic, "traffic_source": traffic_src,
                "kernel": dom, "avg_launch_ms": round(avg_ms, 4), "launches": launches,
                "algorithmic_bytes_per_launch": round(per_launch_bytes, 1),
                "copy_peak_measured": round(copy_peak, 1),
                "frac_of_copy_peak": round(achieved / copy_peak, 5)}
    algo_total = 24 * paths + 168 * tot["ext_rays"] + 112 * tot["hit_rays"] + 96 * tot["shadow_rays"]

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mpaths/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(elapsed * 1e3 / max(args.steps, 1), 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "displaced_spheres_1080p_depth8", "width": info["width"], "height": info["height"],
                   "spp_timed": args.steps * P, "passes_per_step": P, "maxdepth": info["maxdepth"],
                   "triangles": info["n_tris"], "bvh_width": info["bvh_width"], "max_leaf": info["bvh_max_leaf"],
                   "trace_block": cfg["trace_block"], "trace_waves_per_cu": cfg["trace_waves_per_cu"],
                   "bvh_build_s": round(build_s, 3)},
        "mrays_per_s": round((tot["ext_rays"] + tot["shadow_rays"]) / elapsed / 1e6, 3),
        "counters": tot,
        "path_algorithmic_GBps": round(algo_total / elapsed / 1e9, 2),
        "kernel_ms": {k: round(v["total_ms"], 3) for k, v in ks.items()},
        "kernel_timing": ("the timed steps, kernels one after another" if args.serial_timed else
                          "one more %d-pass step after the timed ones, kernels one after another" % P),
        "roofline": roofline,
        "cpu_baseline": None,
    }
    exp = None if args.no_cpu else r.export()
    r.close()
    if not args.no_cpu:
        out["box_room"] = box_room_leg(_host_threads())
        _progress("box room leg done")
    if not args.no_batch:
        out["batch"] = batch_leg(use_graph=not args.batch_plain_launches)
        _progress("batch leg done")
    if not args.no_adjoint:
        out["adjoint"] = adjoint_leg(512, 32, 6, adam_steps=args.adam_steps)
        _progress("adjoint leg done")
    if not args.no_large:
        out["large_scene"] = large_leg()
        _progress("large scene leg done")
    if exp is not None:
        threads = _host_threads()
        out["cpu_baseline"] = cpu_baseline(exp, args.cpu_seconds, threads, max(1, info["pixelsamples"]))
        _progress("cpu baseline done")
        out["gpu_vs_cpu"] = round(value / max(out["cpu_baseline"]["value"], 1e-9), 2)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
