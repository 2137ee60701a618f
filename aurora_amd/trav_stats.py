"""Traversal work per ray: nodes visited, leaves and triangles tested.

Builds a measurement copy of the library with the traversal counters compiled
in (-DAUR_TRAV_STATS, output in aurora_amd/_build/stats/, never loaded by the
package itself), renders a workload with it and reports the counts per
extension ray and per shadow ray for several BVH shapes.

    python -m aurora_amd.trav_stats --build        # compile the measurement build (CPU)
    python -m aurora_amd.trav_stats [--json PATH]  # run it (GPU)
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_build", "stats")
LIB = os.path.join(OUT, "libaurora_hip.so")
SHAPES = ((4, 1), (4, 2), (4, 4), (8, 2), (8, 4))


def build() -> str:
    from . import build as B
    B.COMMON = B.COMMON + ["-DAUR_TRAV_STATS"]
    B.BUILD = os.path.join(OUT, "obj")
    B.LIB = LIB
    return B.build()


def measure(workloads=(("displaced_spheres", 8), ("box_room", 16))) -> dict:
    from . import _lib as L
    lib = L.load_library(LIB)
    lib.aur_debug_trav_stats.argtypes = [C.POINTER(C.c_ulonglong)]
    lib.aur_debug_trav_stats.restype = C.c_int
    from . import scenes, tuning
    from .renderer import Renderer
    res = {}
    for name, passes in workloads:
        d = tempfile.mkdtemp(prefix="aurora_trav_")
        r = Renderer(scenes.write_config(name, d), os.path.join(d, "config.json"))
        for w, leaf in SHAPES:
            r.build(width=w, max_leaf=leaf)
            tuning.apply_launch(r, tuning.choose(r, width=w, max_leaf=leaf))
            r.set_seed(1)
            buf = (C.c_ulonglong * 8)()
            L.check(lib.aur_debug_trav_stats(buf))
            _, st = r.render_passes(0, passes, fetch=False)
            L.check(lib.aur_debug_trav_stats(buf))
            v = list(buf)
            ne, ns = st["ext_rays"], max(st["shadow_rays"], 1)
            key = "%s_w%d_leaf%d" % (name, w, leaf)
            res[key] = {"passes": passes, "ext_rays": ne, "shadow_rays": st["shadow_rays"],
                        "ext_visits": round(v[0] / ne, 3), "ext_leaves": round(v[1] / ne, 3),
                        "ext_tris": round(v[2] / ne, 3), "shadow_visits": round(v[4] / ns, 3),
                        "shadow_leaves": round(v[5] / ns, 3), "shadow_tris": round(v[6] / ns, 3),
                        "Mpaths_s_instrumented": round(st["paths"] / st["ms"] / 1e3, 1)}
            print(key, json.dumps(res[key]), flush=True)
        r.close()
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--build", action="store_true")
    ap.add_argument("--json", default="")
    a = ap.parse_args()
    if a.build:
        build()
        return 0
    res = measure()
    if a.json:
        with open(a.json, "w") as fh:
            json.dump(res, fh, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
