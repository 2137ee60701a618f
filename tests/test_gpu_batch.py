"""Batched persistent launch: each job's framebuffer and ray counts against a
single-scene render of the same passes (same seed, so the same paths; the
batch adds each pixel's passes in pass order, so the framebuffers are
bit-identical), graph replay against plain launches, both launch orders, a
small path pool that recycles every slot many times, and pass ranges split
over two launches."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def jobs(tmp_path_factory):
    from aurora_amd import scenes
    from aurora_amd.renderer import Renderer
    d = str(tmp_path_factory.mktemp("batch"))
    paths = scenes.batch_jobs(d, n_jobs=6, seed=11, max_spp=3)
    rs = []
    for k, p in enumerate(paths):
        r = Renderer(p, os.path.join(os.path.dirname(p), "config.json"))
        r.build()
        r.set_seed(100 + k)
        rs.append(r)
    ex = Renderer()
    ex.set_launch(pool_paths=1 << 20)
    refs = []
    for r in rs:
        acc, st = r.render_passes(0, r.info()["pixelsamples"])
        refs.append((acc, st))
    yield ex, rs, refs
    for r in rs:
        r.close()
    ex.close()


def _check(rs, refs, res, tot):
    for k, (r, (acc, st), o) in enumerate(zip(rs, refs, res)):
        for key in ("paths", "ext_rays", "hit_rays", "shadow_rays"):
            assert o[key] == st[key], (k, key, o[key], st[key])
        fb = r.framebuffer()
        assert np.array_equal(fb, acc), (k, float(np.abs(fb - acc).max()))
        assert o["latency_ms"] > 0.0
    assert tot["paths"] == sum(st["paths"] for _, st in refs)
    assert tot["ext_rays"] == sum(st["ext_rays"] for _, st in refs)
    assert tot["shadow_rays"] == sum(st["shadow_rays"] for _, st in refs)
    assert max(o["latency_ms"] for o in res) <= tot["ms"] * 1.05 + 1.0


def test_batch_matches_single_scene_renders(jobs):
    from aurora_amd.renderer import batch_submit
    ex, rs, refs = jobs
    print("batch: 6 jobs, graph replay, shortest first", flush=True)
    res, tot = batch_submit(ex, [(r, 0, r.info()["pixelsamples"]) for r in rs])
    _check(rs, refs, res, tot)
    # shortest estimated job first: latencies follow the launch order
    order = sorted(range(len(rs)), key=lambda k: res[k]["order"])
    assert sorted(res[k]["order"] for k in range(len(rs))) == list(range(len(rs)))
    lat = [res[k]["latency_ms"] for k in order]
    assert lat[0] <= lat[-1]


def test_batch_plain_launches_and_given_order(jobs):
    from aurora_amd.renderer import batch_submit
    ex, rs, refs = jobs
    print("batch: plain launches, given order, poll every 2 bounces", flush=True)
    res, tot = batch_submit(ex, [(r, 0, r.info()["pixelsamples"]) for r in rs], order=1, use_graph=False,
                            iters_per_poll=2)
    _check(rs, refs, res, tot)
    assert [o["order"] for o in res] == list(range(len(rs)))


def test_batch_graph_replays_repeat_exactly(jobs):
    """Two launches of the same jobs with a 4096-path pool replay the captured
    graph many times each; counts and images must repeat exactly."""
    from aurora_amd.renderer import batch_submit
    ex, rs, refs = jobs
    print("batch: two graph launches with a small pool", flush=True)
    ex.set_launch(pool_paths=4096)
    try:
        a, ta = batch_submit(ex, [(r, 0, r.info()["pixelsamples"]) for r in rs[:3]], iters_per_poll=16)
        fa = [r.framebuffer() for r in rs[:3]]
        b, tb = batch_submit(ex, [(r, 0, r.info()["pixelsamples"]) for r in rs[:3]], iters_per_poll=16)
    finally:
        ex.set_launch(pool_paths=1 << 20)
    for k in range(3):
        for key in ("paths", "ext_rays", "hit_rays", "shadow_rays"):
            assert a[k][key] == b[k][key] == refs[k][1][key]
        assert np.array_equal(rs[k].framebuffer(), fa[k])
        assert np.array_equal(fa[k], refs[k][0])


def test_batch_pass_ranges_accumulate(jobs):
    from aurora_amd.renderer import batch_submit
    ex, rs, refs = jobs
    print("batch: passes [0, 2) then [2, spp)", flush=True)
    sel = rs[:4]
    batch_submit(ex, [(r, 0, 2) for r in sel])
    res, _ = batch_submit(ex, [(r, 2, r.info()["pixelsamples"] - 2) for r in sel])
    for k, r in enumerate(sel):
        assert np.array_equal(r.framebuffer(), refs[k][0])
        i = r.info()
        assert res[k]["paths"] == (i["pixelsamples"] - 2) * i["width"] * i["height"]


def test_batch_rejects_mixed_bvh_widths(tmp_path):
    from aurora_amd import scenes, _lib as L
    from aurora_amd.renderer import Renderer, batch_submit
    print("batch: mixed BVH widths", flush=True)
    paths = scenes.batch_jobs(str(tmp_path), n_jobs=2, seed=5, max_spp=1)
    a = Renderer(paths[0], os.path.join(os.path.dirname(paths[0]), "config.json"))
    b = Renderer(paths[1], os.path.join(os.path.dirname(paths[1]), "config.json"))
    a.build(width=4)
    b.build(width=8)
    with pytest.raises(L.AuroraError) as ei:
        batch_submit(a, [(a, 0, 1), (b, 0, 1)])
    assert ei.value.code == -1
    a.close()
    b.close()
