"""Launch-parameter autotuner for the wavefront tracer.

``autotune`` renders a few short passes of a scene for each candidate and
keeps the fastest, in three stages: the BVH structure (branching factor,
triangles per leaf), then the traversal launch shape (workgroup size,
persistent waves per CU) on the best structure, then the chunk size (paths
traced at once, ``pool_paths``).  ``autotune_batch`` times the
batched persistent launch over a list of jobs for each path-pool size (the
queue depth) and traversal grid.  Results are cached in ``tuning.json`` next
to this file, per triangle-count bucket for single scenes and under "batch"
for the batched launch; ``choose`` / ``choose_batch`` return the cached
choice (or the defaults), so benchmarks and users pick it up without
re-tuning.

Usage on a GPU box:
    python -m aurora_amd.tuning --workload displaced_spheres --out tuning.json
    python -m aurora_amd.tuning --batch --out tuning.json
"""
from __future__ import annotations

import argparse
import json
import math
import os
import time

from . import _lib as L

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning.json")
DEFAULT = {"bvh_width": 4, "max_leaf": 2, "trace_block": 64, "trace_waves_per_cu": 32, "pool_paths": 128 << 20}
BATCH_DEFAULT = {"bvh_width": 4, "max_leaf": 2, "trace_block": 64, "trace_waves_per_cu": 32, "pool_paths": 32 << 20}
WIDTHS = (4, 8)
LEAVES = (1, 2, 4, 8)
BLOCKS = (64, 128, 256)
WAVES = (8, 16, 24, 32)
POOLS = (2 << 20, 4 << 20, 8 << 20, 16 << 20, 32 << 20)
SINGLE_POOLS = (32 << 20, 64 << 20, 128 << 20, 256 << 20, 320 << 20)


def _bucket(n_tris: int) -> str:
    return "tris_2^%d" % max(0, int(math.log2(max(n_tris, 1))))


def _load(path: str = TABLE) -> dict:
    if not os.path.exists(path):
        return {}
    with open(path) as fh:
        return json.load(fh)


def _store(key: str, entry: dict, path: str) -> None:
    table = _load(path)
    table[key] = entry
    with open(path, "w") as fh:
        json.dump(table, fh, indent=2, sort_keys=True)


def choose(renderer, width=None, max_leaf=None, table: str = TABLE) -> dict:
    """Returns {"bvh_width", "max_leaf", "trace_block", "trace_waves_per_cu",
    "pool_paths"} for the renderer's loaded scene."""
    cfg = dict(DEFAULT)
    try:
        n = renderer.info()["n_tris"]
    except Exception:
        n = 0
    entry = _load(table).get(_bucket(n)) if n else None
    if entry:
        cfg.update({k: entry[k] for k in DEFAULT if k in entry})
    if width:
        cfg["bvh_width"] = width
    if max_leaf:
        cfg["max_leaf"] = max_leaf
    return cfg


def choose_batch(table: str = TABLE) -> dict:
    cfg = dict(BATCH_DEFAULT)
    entry = _load(table).get("batch")
    if entry:
        cfg.update({k: entry[k] for k in BATCH_DEFAULT if k in entry})
    return cfg


def apply_launch(renderer, cfg: dict) -> None:
    renderer.set_launch(trace_block=cfg.get("trace_block", 0), trace_waves_per_cu=cfg.get("trace_waves_per_cu", 0),
                        pool_paths=cfg.get("pool_paths", 0))


def _time_render(renderer, passes: int, repeats: int) -> float:
    renderer.render_passes(0, passes, fetch=False)  # warm (allocations, caches)
    best = float("inf")
    paths = 0
    for _ in range(repeats):
        t0 = time.perf_counter()
        _, st = renderer.render_passes(0, passes, fetch=False)
        best = min(best, time.perf_counter() - t0)
        paths = st["paths"]
    return paths / best / 1e6


# Short renders of the structure and launch stages trace at least this many
# paths: on a small image two passes measure launch overhead, not the
# candidates (box room 256x256: the 2-pass ranking picked a launch shape 3 %
# slower than the default at its 16-pass frame, and the previous entry was
# 12 % slower)
MIN_TIMED_PATHS = 1 << 20


def autotune(renderer, passes: int = 2, repeats: int = 3, table: str = TABLE) -> dict:
    """Times the structure candidates, then the launch shapes; returns the table entry."""
    info0 = renderer.info()
    wh0 = info0["width"] * info0["height"]
    passes = max(passes, min(info0["pixelsamples"], -(-MIN_TIMED_PATHS // wh0)))
    structure = []
    apply_launch(renderer, DEFAULT)
    for w in WIDTHS:
        for leaf in LEAVES:
            renderer.build(width=w, max_leaf=leaf)
            structure.append({"bvh_width": w, "max_leaf": leaf, "mpaths_per_s": _time_render(renderer, passes, repeats)})
    structure.sort(key=lambda r: -r["mpaths_per_s"])
    best = structure[0]
    renderer.build(width=best["bvh_width"], max_leaf=best["max_leaf"])
    launch = []
    for b in BLOCKS:
        for wv in WAVES:
            renderer.set_launch(trace_block=b, trace_waves_per_cu=wv)
            launch.append({"trace_block": b, "trace_waves_per_cu": wv,
                           "mpaths_per_s": _time_render(renderer, passes, repeats)})
    launch.sort(key=lambda r: -r["mpaths_per_s"])
    renderer.set_launch(trace_block=launch[0]["trace_block"], trace_waves_per_cu=launch[0]["trace_waves_per_cu"])
    # chunk size: passes traced at once, up to pool_paths paths (each chunk
    # ends in a latency-bound drain); timed on up to 64 passes
    info = renderer.info()
    wh = info["width"] * info["height"]
    n_pass = max(2, min(64, (SINGLE_POOLS[-1] + wh - 1) // wh))
    pools = []
    for pool in SINGLE_POOLS:
        if pool > SINGLE_POOLS[2] and n_pass * wh <= SINGLE_POOLS[2]:
            continue  # the timed passes already fit one 128 M chunk
        renderer.set_launch(pool_paths=pool)
        try:
            pools.append({"pool_paths": pool, "mpaths_per_s": _time_render(renderer, n_pass, 2)})
        except L.AuroraError:
            break  # 368 bytes of path state per slot: a pool the device memory cannot hold ends the sweep
    pools.sort(key=lambda r: -r["mpaths_per_s"])
    entry = {"bvh_width": best["bvh_width"], "max_leaf": best["max_leaf"], "trace_block": launch[0]["trace_block"],
             "trace_waves_per_cu": launch[0]["trace_waves_per_cu"], "pool_paths": pools[0]["pool_paths"],
             "mpaths_per_s": launch[0]["mpaths_per_s"], "structure_candidates": structure, "launch_candidates": launch,
             "pool_candidates": pools}
    apply_launch(renderer, entry)
    _store(_bucket(renderer.info()["n_tris"]), entry, table)
    return entry


def autotune_batch(exec_renderer, jobs, repeats: int = 2, table: str = TABLE) -> dict:
    """jobs: [(renderer, first_pass, n_passes)] with loaded scenes.  Times each
    triangles-per-leaf choice (all jobs rebuilt with BVH width 4), then each
    (path-pool size, traversal waves per CU, workgroup size); returns the
    table entry."""
    from .renderer import batch_submit

    def timed():
        batch_submit(exec_renderer, jobs)  # warm
        best = float("inf")
        for _ in range(repeats):
            t0 = time.perf_counter()
            _, tot = batch_submit(exec_renderer, jobs)
            best = min(best, time.perf_counter() - t0)
        return tot["paths"] / best / 1e6

    leaves = []
    apply_launch(exec_renderer, BATCH_DEFAULT)
    for leaf in (1, 2, 4):
        for r, _, _ in jobs:
            r.build(width=4, max_leaf=leaf)
        leaves.append({"max_leaf": leaf, "mpaths_per_s": timed()})
    leaves.sort(key=lambda r: -r["mpaths_per_s"])
    for r, _, _ in jobs:
        r.build(width=4, max_leaf=leaves[0]["max_leaf"])
    cands = []
    for pool in POOLS:
        for wv in (16, 24, 32):
            for blk in (64, 128):
                exec_renderer.set_launch(trace_block=blk, trace_waves_per_cu=wv, pool_paths=pool)
                cands.append({"pool_paths": pool, "trace_waves_per_cu": wv, "trace_block": blk, "mpaths_per_s": timed()})
    cands.sort(key=lambda r: -r["mpaths_per_s"])
    entry = {"bvh_width": 4, "max_leaf": leaves[0]["max_leaf"], "trace_block": cands[0]["trace_block"],
             "trace_waves_per_cu": cands[0]["trace_waves_per_cu"], "pool_paths": cands[0]["pool_paths"],
             "mpaths_per_s": cands[0]["mpaths_per_s"], "leaf_candidates": leaves, "candidates": cands}
    apply_launch(exec_renderer, entry)
    _store("batch", entry, table)
    return entry


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="displaced_spheres")
    ap.add_argument("--batch", action="store_true", help="tune the batched launch on the 64-job workload")
    ap.add_argument("--out", default=TABLE, help="table to update")
    args = ap.parse_args(argv)
    import tempfile

    from . import scenes
    from .renderer import Renderer
    tmp = tempfile.mkdtemp(prefix="aurora_tune_")
    if args.batch:
        paths = scenes.batch_jobs(tmp, n_jobs=64, seed=3)
        rs = []
        for k, p in enumerate(paths):
            r = Renderer(p, os.path.join(os.path.dirname(p), "config.json"))
            r.build()
            r.set_seed(1000 + k)
            rs.append(r)
        ex = Renderer()
        # up to four passes of every job: the same mix of scenes at ~1/15 of the work
        jobs = [(r, 0, min(4, r.info()["pixelsamples"])) for r in rs]
        print(json.dumps(autotune_batch(ex, jobs, table=args.out)), flush=True)
        return 0
    path = scenes.write_config(args.workload, tmp)
    r = Renderer(path, os.path.join(tmp, "config.json"))
    r.build()
    print(json.dumps(autotune(r, table=args.out)), flush=True)
    return 0


if __name__ == "__main__":
    main()
