"""GPU: the large differentiable scene at full size (3840x2160, depth 16,
32 displaced level-7 icospheres on 8 terrain patches: 18,506,000 triangles,
2048x2048 FP32 albedo texture on the terrain).  Built on the GPU, one pass
rendered: finite non-negative image, ray counts consistent, and sampled
paths equal to the oracle's BVH render of the exported scene with the same
texture; the texture gradient of one pass is finite and non-zero exactly on
texels the terrain shows."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def large(tmp_path_factory):
    from aurora_amd import scenes, tuning
    from aurora_amd.renderer import Renderer
    d = str(tmp_path_factory.mktemp("large"))
    print("large scene: writing", flush=True)
    spec, truth = scenes.large_scene()
    path = spec.write(d)
    r = Renderer(path, os.path.join(d, "config.json"))
    print("large scene: building", flush=True)
    r.build()
    cfg = tuning.choose(r)
    r.build(width=cfg["bvh_width"], max_leaf=cfg["max_leaf"])
    tuning.apply_launch(r, cfg)
    attrs = [spec.attr_index(k) for k in sorted(spec.objects) if k.startswith("terrain")]
    r.set_albedo_texture(truth, attrs)
    yield r, truth, attrs
    r.close()


def test_large_scene_full_size_pass(large, oracle):
    r, truth, attrs = large
    info = r.info()
    assert info["n_tris"] == 18506000 and (info["width"], info["height"], info["maxdepth"]) == (3840, 2160, 16)
    r.set_seed(3)
    print("large scene: one pass", flush=True)
    img, st = r.render_passes(0, 1)
    assert np.isfinite(img).all() and img[..., :3].min() >= 0 and img[..., :3].mean() > 0.01
    assert st["paths"] == 3840 * 2160
    assert st["paths"] <= st["ext_rays"] <= 17 * st["paths"]
    assert st["hit_rays"] <= st["ext_rays"] and st["shadow_rays"] <= 2 * st["hit_rays"]
    print("large scene: oracle BVH", flush=True)
    e = r.export()
    s = oracle.Scene(e, bvh=True, ptex=truth, ptex_attrs=attrs)
    rng = np.random.default_rng(5)
    pix = rng.integers(0, 3840 * 2160, 3000).astype(np.int32)
    smp = rng.integers(0, 256, 3000).astype(np.int32)
    g, _ = r.render_paths(pix, smp)
    o, _ = s.render_paths(3, pix, smp, threads=16)
    rel = np.abs(g - o) / (np.abs(o) + 1e-3)
    assert np.mean(rel > 1e-4) < 5e-3, np.mean(rel > 1e-4)


def test_large_scene_texture_gradient(large):
    r, truth, attrs = large
    w = np.ones((2160, 3840, 3), np.float32)
    r.set_seed(4)
    print("large scene: adjoint pass", flush=True)
    _, _, _, st, gt = r.adjoint(0, 1, w, albedo=False, texture=True)
    assert gt.shape == truth.shape and np.isfinite(gt).all()
    # a brighter texel brightens the image: the texels the terrain shows have
    # positive gradients, the others none (the replay's remaining radiance
    # L - sum of the contributions so far can round a hair below zero at a
    # path's last vertex, hence the relative floor)
    assert gt.min() >= -1e-4 * gt.max(), (gt.min(), gt.max())
    assert np.count_nonzero(gt > 1e-4 * gt.max()) > 0.2 * gt.size
    assert st["paths"] == 3840 * 2160
