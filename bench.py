#!/usr/bin/env python3
"""Benchmark: forward path-tracing throughput on one MI355X.

Workload: the 1920x1080, 64 spp, max-depth-8 scene of eight displaced, Loop-
subdivided (7 levels) icospheres (2,621,450 triangles with the light quads),
Kelemen/GGX materials, four quad lights and a constant environment.  A step is
one batch of ``--passes-per-step`` full-frame passes (each pass = one path per
pixel); the default 16 steps x 4 passes render the full 64-spp frame.

Prints one JSON line.  ``value`` is Mpaths/s of the timed steps with the scene
and BVH resident in HBM.  ``roofline`` describes the kernel with the largest
share of device time, timed with HIP events on the library stream.
``cpu_baseline`` is the CPU restatement of the reference algorithm (the
oracle, with its own SAH BVH and OpenMP tiles) timed on this host on a bounded
sample of the same frame.
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
#   k_shade:  ray (32) + hit (16) + path state read/write (2 x 36) per extension ray
ALGO_BYTES = {"k_extend": 48.0, "k_shadow": 48.0, "k_shade": 120.0}


def _hip_sync():
    lib = ctypes.CDLL("libamdhip64.so")
    lib.hipDeviceSynchronize.restype = ctypes.c_int
    rc = lib.hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError("hipDeviceSynchronize failed: %d" % rc)


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
    path = os.path.join(HERE, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as fh:
            t = json.load(fh)
        v = t.get(kernel, {}).get("hbm_bytes_per_launch")
        return float(v) if v is not None else None
    except (OSError, ValueError):
        return None


def cpu_baseline(exp: dict, target_s: float, threads: int):
    """Times the oracle on every k-th row of one pass, sized to ~target_s."""
    from oracle import oracle as O
    O.build()
    t0 = time.time()
    scene = O.Scene(exp, bvh=True)
    build_s = time.time() - t0
    W, H = exp["width"], exp["height"]

    def sample_rows(step):
        rows = np.arange(0, H, step)
        pix = (rows[:, None] * W + np.arange(W)[None, :]).reshape(-1).astype(np.int32)
        return pix

    # calibrate on a small sample, then size the timed sample
    pix = sample_rows(64)
    t0 = time.time()
    scene.render_paths(7, pix, np.zeros_like(pix), threads=threads)
    dt = max(time.time() - t0, 1e-4)
    rate = pix.size / dt
    want = int(rate * target_s)
    step = max(1, int(round(W * H / max(want, 1))))
    pix = sample_rows(step)
    smp = np.zeros_like(pix)
    t0 = time.time()
    _, cnt = scene.render_paths(7, pix, smp, threads=threads)
    dt = time.time() - t0
    return {"value": round(pix.size / dt / 1e6, 4), "unit": "Mpaths/s", "cores": threads, "kind": "port",
            "sample": "one pass (1 spp) over every %d-th row of the same 1920x1080 frame: %d paths, %.1f s; "
                      "BVH build %.1f s excluded; host %s, nproc %d" % (step, pix.size, dt, build_s, _cpu_model(),
                                                                       os.cpu_count() or 0),
            "mrays_per_s": round((cnt["ext_rays"] + cnt["shadow_rays"]) / dt / 1e6, 4)}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--passes-per-step", type=int, default=4)
    ap.add_argument("--bvh-width", type=int, default=0, help="0: autotuned choice")
    ap.add_argument("--max-leaf", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--levels", type=int, default=7)
    args = ap.parse_args(argv)

    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported "
              "(multi-GPU sharding is design-only)" % args.gpus, file=sys.stderr)
        return 2

    from aurora_amd import scenes, tuning
    from aurora_amd.renderer import Renderer

    tmp = tempfile.mkdtemp(prefix="aurora_bench_")
    scene_path = scenes.write_config("displaced_spheres", tmp, levels=args.levels)
    r = Renderer(scene_path, os.path.join(tmp, "config.json"))
    cfg = tuning.choose(r, width=args.bvh_width or None, max_leaf=args.max_leaf or None)
    t0 = time.time()
    r.build(width=cfg["bvh_width"], max_leaf=cfg["max_leaf"])
    build_s = time.time() - t0
    info = r.info()
    wh = info["width"] * info["height"]
    P = args.passes_per_step

    for k in range(args.warmup):
        r.render_passes(k * P, P, fetch=False)
    r.set_profiling(True)
    _hip_sync()
    t0 = time.perf_counter()
    tot = {"paths": 0, "ext_rays": 0, "hit_rays": 0, "shadow_rays": 0}
    for k in range(args.steps):
        _, st = r.render_passes(k * P, P, fetch=False)
        for key in tot:
            tot[key] += st[key]
    _hip_sync()
    elapsed = time.perf_counter() - t0
    ks = r.kernel_stats()
    r.set_profiling(False)

    paths = tot["paths"]
    value = paths / elapsed / 1e6
    dom = max((k for k in ks if ks[k]["launches"] > 0), key=lambda k: ks[k]["total_ms"])
    launches = ks[dom]["launches"]
    avg_ms = ks[dom]["total_ms"] / max(launches, 1)
    items = {"k_extend": tot["ext_rays"], "k_shadow": tot["shadow_rays"], "k_shade": tot["ext_rays"]}.get(dom, 0)
    per_launch_bytes = ALGO_BYTES.get(dom, 0.0) * items / max(launches, 1)
    achieved = per_launch_bytes / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": _load_traffic(dom), "kernel": dom,
                "avg_launch_ms": round(avg_ms, 4), "launches": launches,
                "algorithmic_bytes_per_launch": round(per_launch_bytes, 1)}
    algo_total = 24 * paths + 168 * tot["ext_rays"] + 112 * tot["hit_rays"] + 96 * tot["shadow_rays"]

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mpaths/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(elapsed * 1e3 / max(args.steps, 1), 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "displaced_spheres_1080p_depth8", "width": info["width"], "height": info["height"],
                   "spp_timed": args.steps * P, "passes_per_step": P, "maxdepth": info["maxdepth"],
                   "triangles": info["n_tris"], "bvh_width": info["bvh_width"], "max_leaf": info["bvh_max_leaf"],
                   "bvh_build_s": round(build_s, 3)},
        "mrays_per_s": round((tot["ext_rays"] + tot["shadow_rays"]) / elapsed / 1e6, 3),
        "counters": tot,
        "path_algorithmic_GBps": round(algo_total / elapsed / 1e9, 2),
        "kernel_ms": {k: round(v["total_ms"], 3) for k, v in ks.items()},
        "roofline": roofline,
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(r.export(), args.cpu_seconds, threads)
        out["gpu_vs_cpu"] = round(value / max(out["cpu_baseline"]["value"], 1e-9), 2)
    print(json.dumps(out), flush=True)
    r.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
