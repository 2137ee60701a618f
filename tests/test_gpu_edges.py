"""Edge cases of the render path against the oracle: image sizes that do not
fill the 8x8 tiles paths are issued in (and a single pixel), a scene with no
triangles, the shallowest and a deep path depth, the largest leaf size, and
zero passes."""
import os

import numpy as np
import pytest

from conftest import random_rays

pytestmark = pytest.mark.gpu


def _room(tmp, w, h, spp=4, maxdepth=5, mindepth=5):
    from aurora_amd import scenes
    from aurora_amd.renderer import Renderer
    spec = scenes.box_room(res=max(w, h), spp=spp, maxdepth=maxdepth, mindepth=mindepth)
    spec.options["resolution"] = [int(w), int(h)]
    path = spec.write(str(tmp))
    r = Renderer(path, os.path.join(str(tmp), "config.json"))
    r.build()
    return r


@pytest.mark.parametrize("w,h", [(37, 23), (1, 1), (1, 13), (9, 1)])
def test_ragged_resolutions_match_oracle(tmp_path, oracle, w, h):
    """Widths and heights that leave partial 8x8 tiles: the full render equals
    the oracle's on the same stream (every pixel written once per pass)."""
    print("ragged resolution %dx%d" % (w, h), flush=True)
    r = _room(tmp_path, w, h)
    e = r.export()
    assert (e["width"], e["height"]) == (w, h)
    r.set_seed(12)
    acc, gst = r.render_passes(0, 3)
    oacc, ost = oracle.Scene(e).render(12, 0, 3)
    assert acc.shape == (h, w, 4)
    diff = np.abs(acc - oacc)
    assert np.mean(diff > 1e-3 * (1 + np.abs(oacc))) <= max(1e-3, 1.0 / (w * h))
    assert gst["paths"] == 3 * w * h
    for k in ("paths", "ext_rays", "hit_rays", "shadow_rays"):
        assert abs(gst[k] - ost[k]) <= max(2, 1e-4 * ost[k]), k
    r.close()


def test_scene_without_triangles(tmp_path):
    """Only an environment light: no triangles, an empty BVH, every camera
    ray misses and sees color * 2^exposure with alpha 1 (integrator.cpp:
    287-297), no extension ray hits and no shadow ray is traced."""
    from aurora_amd import scenes
    from aurora_amd.renderer import Renderer
    spec = scenes.SceneSpec(options=scenes._options(24, 16, 2, 4, 4),
                            camera={"fov": 0.5, "transforms": []})
    spec.objects["sky"] = {"type": "envlight", "name": "sky", "color": [0.3, 0.2, 0.1], "exposure": 1.0,
                           "envmap": "", "transforms": []}
    path = spec.write(str(tmp_path))
    r = Renderer(path, os.path.join(str(tmp_path), "config.json"))
    r.build()
    i = r.info()
    assert i["n_tris"] == 0
    acc, st = r.render_passes(0, 2)
    want = np.float32([0.3, 0.2, 0.1]) * np.float32(2.0)
    assert np.array_equal(acc[..., :3], np.broadcast_to(2 * want, acc[..., :3].shape))
    assert np.all(acc[..., 3] == 2.0)
    assert st["paths"] == 2 * 24 * 16 and st["ext_rays"] == st["paths"]
    assert st["hit_rays"] == 0 and st["shadow_rays"] == 0
    hits = r.trace(random_rays(np.random.default_rng(1), 1000, [-1, -1, -1], [1, 1, 1]))
    assert np.all(hits["prim"] < 0)
    r.close()


@pytest.mark.parametrize("depth", [1, 40])
def test_shallow_and_deep_paths_match_oracle(tmp_path, oracle, depth):
    """maxdepth 1 (camera ray, one shaded vertex, one traced extension ray)
    and 40 (the injected-slot contract at 362 uniforms per path): per-path
    radiance and ray counts equal the oracle's."""
    r = _room(tmp_path, 32, 32, maxdepth=depth, mindepth=depth)
    e = r.export()
    s = oracle.Scene(e)
    rng = np.random.default_rng(40 + depth)
    n = 4000
    slots = rng.uniform(0, 1, size=(n, 2 + 9 * depth)).astype(np.float32)
    pix = rng.integers(0, 32 * 32, n).astype(np.int32)
    smp = rng.integers(0, 16, n).astype(np.int32)
    g, gs = r.render_paths(pix, smp, slots=slots)
    o, os_ = s.render_paths(0, pix, smp, slots=slots)
    rel = np.abs(g - o) / (np.abs(o) + 1e-3)
    assert np.mean(rel > 1e-4) < 3e-3, np.mean(rel > 1e-4)
    for k in ("paths", "ext_rays", "hit_rays", "shadow_rays"):
        assert abs(gs[k] - os_[k]) <= max(2, 1e-3 * os_[k]), k
    if depth == 1:
        assert gs["ext_rays"] <= 2 * n
    r.close()


@pytest.mark.parametrize("width", [4, 8])
def test_largest_leaves_trace_bit_exact(tmp_path, oracle, width):
    """max_leaf 16 (the largest leaf the node encoding holds): hits equal the
    brute-force oracle bit for bit on the box room."""
    r = _room(tmp_path, 16, 16)
    r.build(width=width, max_leaf=16)
    s = oracle.Scene(r.export())
    rays = random_rays(np.random.default_rng(5), 20000, [-0.99, -0.99, -4.49], [0.99, 0.99, -2.51])
    g, o = r.trace(rays), s.trace(rays)
    assert np.array_equal(g["prim"], o["prim"])
    hit = o["prim"] >= 0
    for k in ("t", "u", "v"):
        assert np.array_equal(g[k][hit], o[k][hit]), k
    ga, oa = r.trace(rays, any_hit=True), s.trace(rays, any_hit=True)
    assert np.array_equal(ga["prim"] >= 0, oa["prim"] >= 0)
    r.close()


def test_zero_passes(tmp_path):
    """A render of no passes clears the framebuffer and traces nothing; the
    display written for a scene of zero pixel samples is black."""
    from aurora_amd import exr
    r = _room(tmp_path, 16, 8)
    r.render_passes(0, 2, fetch=False)
    acc, st = r.render_passes(0, 0)
    assert np.all(acc == 0) and st["paths"] == 0 and st["ext_rays"] == 0
    r.close()
    d = tmp_path / "zero"
    d.mkdir()
    r = _room(d, 16, 8, spp=0)
    out = str(tmp_path / "zero.exr")
    got = r.render_image(out)
    assert got["draws"] == 1 and got["paths"] == 0
    assert np.all(exr.read_exr(out) == 0)
    r.close()
