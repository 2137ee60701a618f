"""Light samples on the GPU (lightsamples > 1: passes after the first trace
`lightsamples` paths per pixel through the pass's one camera sample,
renderer.cpp:233, 257, integrator.cpp:152-300) against the oracle: full frames
from the same stream, ray counts (the camera ray counted once per pixel and
pass), pass ranges that start after pass 0, batches of passes split over
several wavefront runs, and camera misses onto an environment map."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _renderer(tmp, spec, L):
    from aurora_amd.renderer import Renderer
    spec.options["lightsamples"] = L
    path = spec.write(tmp)
    r = Renderer(path, os.path.join(tmp, "config.json"))
    r.build()
    return r


def _match(acc, oacc, tol_frac=1e-3):
    diff = np.abs(acc - oacc)
    assert np.mean(diff > 1e-3 * (1 + np.abs(oacc))) < tol_frac
    np.testing.assert_allclose(acc.mean(axis=(0, 1)), oacc.mean(axis=(0, 1)), rtol=2e-4)


def test_light_samples_box_room(tmp_path, oracle):
    from aurora_amd import scenes
    r = _renderer(str(tmp_path), scenes.box_room(res=64, spp=4, maxdepth=4, mindepth=4), 4)
    e = r.export()
    assert e["lightsamples"] == 4
    s = oracle.Scene(e)
    r.set_seed(21)
    acc, gst = r.render_passes(0, 3)
    oacc, ost = s.render(21, 0, 3)
    _match(acc, oacc)
    wh = 64 * 64
    assert gst["paths"] == ost["paths"] == wh * (1 + 2 * 4)
    for k in ("ext_rays", "hit_rays", "shadow_rays"):
        assert abs(gst[k] - ost[k]) <= max(2, 1e-3 * ost[k]), (k, gst[k], ost[k])
    # passes [2, 4) on top of [0, 2) equal passes [0, 4)
    a4, _ = r.render_passes(0, 4)
    r.render_passes(0, 2, fetch=False)
    b4, _ = r.render_passes(2, 2)
    assert np.array_equal(a4, b4)
    r.close()


def test_light_samples_split_over_runs(tmp_path, oracle):
    """16 paths per pixel at 256^2: a pass is 1 M paths, so 12 passes run as
    several wavefront batches of whole passes."""
    from aurora_amd import scenes
    r = _renderer(str(tmp_path), scenes.box_room(res=256, spp=12, maxdepth=2, mindepth=2), 16)
    s = oracle.Scene(r.export())
    r.set_seed(3)
    acc, gst = r.render_passes(0, 12)
    oacc, ost = s.render(3, 0, 12)
    _match(acc, oacc)
    assert gst["paths"] == ost["paths"] == 256 * 256 * (1 + 11 * 16)
    r.close()


def test_light_samples_with_environment_and_shaders(tmp_path, oracle):
    from aurora_amd import scenes
    r = _renderer(str(tmp_path), scenes.textured_stage(res=48, spp=3, maxdepth=3, mindepth=3), 3)
    s = oracle.Scene(r.export())
    r.set_seed(8)
    acc, gst = r.render_passes(0, 3)
    oacc, ost = s.render(8, 0, 3)
    _match(acc, oacc, 4e-3)
    assert np.all(acc[..., 3] == 3.0)
    r.close()


def test_light_samples_halton(tmp_path, oracle):
    from aurora_amd import scenes
    r = _renderer(str(tmp_path), scenes.box_room(res=48, spp=3, maxdepth=3, mindepth=3), 2)
    e = r.export()
    e["sampler"] = "halton"
    s = oracle.Scene(e)
    r.set_sampler("halton")
    acc, _ = r.render_passes(0, 3)
    oacc, _ = s.render(0, 0, 3)
    _match(acc, oacc)
    r.close()


def test_light_samples_in_batched_launch(tmp_path):
    """Jobs with 1, 2 and 3 light samples (one starting after pass 0) in one
    persistent launch with a small pool: framebuffers bit-identical to
    per-scene renders, ray counts equal."""
    from aurora_amd import scenes
    from aurora_amd.renderer import Renderer, batch_submit
    rs, refs, jobs = [], [], []
    for k, (L, first, n) in enumerate(((1, 0, 3), (2, 0, 3), (3, 1, 2))):
        d = str(tmp_path / ("job%d" % k))
        r = _renderer(d, scenes.box_room(res=48 + 16 * k, spp=4, maxdepth=3, mindepth=3, seed=k), L)
        r.set_seed(50 + k)
        if first > 0:
            r.render_passes(0, first, fetch=False)
            base = r.framebuffer()
        acc, st = r.render_passes(first, n)
        refs.append((acc, st, base if first > 0 else None))
        rs.append(r)
        jobs.append((r, first, n))
    ex = Renderer()
    ex.set_launch(pool_paths=1 << 14)
    for r, (acc, st, base) in zip(rs, refs):
        if base is not None:  # restore the framebuffer of the passes before the batch
            r.render_passes(0, jobs[rs.index(r)][1], fetch=False)
    res, tot = batch_submit(ex, jobs)
    for r, (acc, st, base), o in zip(rs, refs, res):
        assert np.array_equal(r.framebuffer(), acc)
        for key in ("paths", "ext_rays", "hit_rays", "shadow_rays"):
            assert o[key] == st[key], key
    for r in rs:
        r.close()
    ex.close()
