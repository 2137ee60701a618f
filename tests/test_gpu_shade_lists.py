"""GPU: the two ways a bounce is shaded give the same render bit for bit.

A long bounce queue after the camera bounce is shaded from per-workgroup hit
lists: each workgroup reads a tile of hit records, finishes the entries with
no vertex to shade (misses, and every entry after the last vertex) and
shades the listed hits in full waves; a shorter queue, and the camera
bounce, are shaded in place.  The listing threshold
(aur_launch_opts.shade_list_tiles, in 1024-entry tiles) is forced to one
tile (every bounce after the first listed) and to never, and framebuffer, FP16 display and
ray counts must agree on: an open scene under a constant environment
(displaced spheres), the textured stage (environment map, mirror rays that
miss, the shader code), the box room with the Halton sampler and with
light samples (later samples of a pass share its camera ray), and the
adjoint's first replay (per-path records written when paths finish; its
gradients agree to the float-atomic summation order of the second replay).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NEVER = 1 << 30


def _render(r, listed, seed, passes):
    r.set_launch(shade_list_tiles=1 if listed else NEVER)
    r.set_seed(seed)
    img, st = r.render_passes(0, passes)
    return img, r.display(), st


def _check(r, seed, passes, what):
    before = r.get_launch()
    try:
        print("shade lists:", what, flush=True)
        a, da, sa = _render(r, False, seed, passes)
        b, db, sb = _render(r, True, seed, passes)
    finally:
        r.set_launch(**before)
    assert np.array_equal(a, b), what
    assert np.array_equal(da, db), what
    for k in ("paths", "ext_rays", "hit_rays", "shadow_rays"):
        assert sa[k] == sb[k], (what, k)
    assert sa["ext_rays"] > sa["hit_rays"], what  # the scene has misses to set aside
    return sa


def _renderer(spec, d):
    from aurora_amd.renderer import Renderer
    path = spec.write(d)
    r = Renderer(path, os.path.join(d, "config.json"))
    r.build()
    return r


def test_open_scene_listed_equals_in_place(tmp_path):
    from aurora_amd import scenes
    r = _renderer(scenes.displaced_spheres(width=192, height=108, spp=8, levels=3), str(tmp_path))
    try:
        st = _check(r, 3, 8, "displaced spheres 192x108, 8 passes")
        # most bounce rays leave the scene
        assert st["ext_rays"] - st["hit_rays"] > 0.3 * (st["ext_rays"] - st["paths"])
    finally:
        r.close()


def test_textured_stage_listed_equals_in_place(tmp_path):
    from aurora_amd import scenes
    r = _renderer(scenes.textured_stage(res=96, spp=6, maxdepth=4, mindepth=4), str(tmp_path))
    try:
        _check(r, 9, 6, "textured stage 96x96, 6 passes")
    finally:
        r.close()


def test_halton_and_light_samples_listed_equal_in_place(tmp_path):
    from aurora_amd import scenes
    spec = scenes.box_room(res=96, spp=6, maxdepth=4, mindepth=1)
    spec.options["lightsamples"] = 3
    r = _renderer(spec, str(tmp_path))
    try:
        _check(r, 4, 6, "box room 96x96, 3 light samples, roulette after bounce 1")
        r.set_sampler("halton")
        _check(r, 4, 6, "box room 96x96, Halton sampler")
        r.set_sampler("hash")
    finally:
        r.close()


def test_adjoint_records_listed_equal_in_place(tmp_path):
    from aurora_amd import scenes
    r = _renderer(scenes.box_room(res=64, spp=2, maxdepth=4, mindepth=4, kelemen=True), str(tmp_path))
    before = r.get_launch()
    try:
        info = r.info()
        g = np.random.default_rng(2).standard_normal((info["height"], info["width"], 4)).astype(np.float32)
        out = []
        for listed in (False, True):
            r.set_launch(shade_list_tiles=1 if listed else NEVER)
            r.set_seed(6)
            out.append(r.adjoint(0, 2, g, albedo=False))
        (_, gr, gs, sa), (_, hr, hs, sb) = out
        # the gradient replay sums per-parameter gradients with float atomics,
        # whose order (and last bits) no two replays share
        for x, y in ((gr, hr), (gs, hs)):
            assert np.allclose(x, y, rtol=1e-5, atol=1e-6 * np.abs(x).max())
        for k in ("paths", "ext_rays", "hit_rays", "shadow_rays"):
            assert sa[k] == sb[k], k
        assert np.any(gr != 0) and np.any(gs != 0)
    finally:
        r.set_launch(**before)
        r.close()
