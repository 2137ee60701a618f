"""Launch-parameter autotuner for the wavefront tracer.

``autotune`` renders a few short batches of a scene for each candidate
(BVH branching factor, triangles per leaf, paths per batch) and keeps the
fastest; results are cached per triangle-count bucket in ``tuning.json`` next
to this file.  ``choose`` returns the cached choice for a loaded scene (or the
defaults when no entry exists), so benchmarks and users pick it up without
re-tuning.

Usage on a GPU box:  python -m aurora_amd.tuning --workload displaced_spheres
"""
from __future__ import annotations

import argparse
import json
import math
import os
import time

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning.json")
DEFAULT = {"bvh_width": 4, "max_leaf": 4}
WIDTHS = (4, 8)
LEAVES = (1, 2, 4, 8)


def _bucket(n_tris: int) -> str:
    return "tris_2^%d" % max(0, int(math.log2(max(n_tris, 1))))


def _load() -> dict:
    if not os.path.exists(TABLE):
        return {}
    with open(TABLE) as fh:
        return json.load(fh)


def choose(renderer, width=None, max_leaf=None) -> dict:
    """Returns {"bvh_width", "max_leaf"} for the renderer's loaded scene."""
    cfg = dict(DEFAULT)
    try:
        n = renderer.info()["n_tris"]
    except Exception:
        n = 0
    entry = _load().get(_bucket(n)) if n else None
    if entry:
        cfg.update({k: entry[k] for k in ("bvh_width", "max_leaf") if k in entry})
    if width:
        cfg["bvh_width"] = width
    if max_leaf:
        cfg["max_leaf"] = max_leaf
    return cfg


def autotune(renderer, passes: int = 2, repeats: int = 3) -> dict:
    """Times every (width, leaf) candidate on the loaded scene; returns the table entry."""
    results = []
    for w in WIDTHS:
        for leaf in LEAVES:
            renderer.build(width=w, max_leaf=leaf)
            renderer.render_passes(0, passes, fetch=False)  # warm
            best = float("inf")
            for _ in range(repeats):
                t0 = time.perf_counter()
                _, st = renderer.render_passes(0, passes, fetch=False)
                best = min(best, time.perf_counter() - t0)
            paths = st["paths"]
            results.append({"bvh_width": w, "max_leaf": leaf, "mpaths_per_s": paths / best / 1e6})
    results.sort(key=lambda r: -r["mpaths_per_s"])
    n = renderer.info()["n_tris"]
    entry = dict(results[0])
    entry["candidates"] = results
    table = _load()
    table[_bucket(n)] = entry
    with open(TABLE, "w") as fh:
        json.dump(table, fh, indent=2, sort_keys=True)
    return entry


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="displaced_spheres")
    args = ap.parse_args(argv)
    import tempfile

    from . import scenes
    from .renderer import Renderer
    tmp = tempfile.mkdtemp(prefix="aurora_tune_")
    path = scenes.write_config(args.workload, tmp)
    r = Renderer(path, os.path.join(tmp, "config.json"))
    r.build()
    print(json.dumps(autotune(r)))


if __name__ == "__main__":
    main()
