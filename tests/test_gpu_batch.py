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


def test_batch_counts_shadow_rays_of_deep_paths(tmp_path, oracle):
    """A closed room (walls extended behind the camera), depth 100, of matte
    albedo e/2: each vertex multiplies the throughput by 2 albedo cos with cos
    uniform, whose log has mean 0, so a path ends mostly on the light and many
    trace more than 63 shadow rays (counted path by path with the oracle).
    The batched launch counts shadow rays per job where they are traced, so
    its counts still equal a single-scene render's."""
    from aurora_amd import scenes
    from aurora_amd.renderer import Renderer, batch_submit
    spec = scenes.box_room(res=32, spp=2, maxdepth=100, mindepth=100)
    x0, x1, y0, y1, z0, z1 = -1.0, 1.0, 0.0, 2.0, -1.0, 3.6
    fv, ff = [], []
    for q in (([x0, y0, z1], [x1, y0, z1], [x1, y0, z0], [x0, y0, z0]),
              ([x0, y1, z0], [x1, y1, z0], [x1, y1, z1], [x0, y1, z1]),
              ([x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0]),
              ([x1, y0, z1], [x0, y0, z1], [x0, y1, z1], [x1, y1, z1])):
        v, f = scenes.quad(*q)
        ff += [[a + len(fv) for a in t] for t in f]
        fv += v
    spec.meshes["walls_white.obj"] = (fv, ff, None)
    spec.meshes["wall_left.obj"] = (*scenes.quad([x0, y0, z1], [x0, y0, z0], [x0, y1, z0], [x0, y1, z1]), None)
    spec.meshes["wall_right.obj"] = (*scenes.quad([x1, y0, z0], [x1, y0, z1], [x1, y1, z1], [x1, y1, z0]), None)
    for name in list(spec.materials):
        spec.materials[name] = {"type": "matteMaterial", "color": [1.36, 1.36, 1.36]}
    path = spec.write(str(tmp_path / "deep"))
    r = Renderer(path, os.path.join(os.path.dirname(path), "config.json"))
    r.build()
    r.set_seed(5)
    acc, st = r.render_passes(0, 2)
    o = oracle.Scene(r.export())
    deep = 0
    for p in range(0, 32 * 32, 4):
        _, c = o.render_paths(5, np.array([p], np.int32), np.array([1], np.int32))
        deep += c["shadow_rays"] > 63
    assert deep >= 3, deep
    ex = Renderer()
    res, tot = batch_submit(ex, [(r, 0, 2)])
    for key in ("paths", "ext_rays", "hit_rays", "shadow_rays"):
        assert res[0][key] == st[key], (key, res[0][key], st[key])
    assert np.array_equal(r.framebuffer(), acc)
    ex.close()
    r.close()


def test_batched_workload_seed3_full(tmp_path, oracle):
    """The batched workload as benchmarked: the 64 seed-3 jobs (996..327,728
    triangles, 128^2..1024^2, 4..256 spp, depth 3..12) through one batched
    launch with the tuned launch shape.  Every job's framebuffer and ray counts
    equal its own single-scene render bit for bit, and sampled paths of three
    jobs (the smallest and largest image, the deepest paths) equal the
    oracle's BVH render."""
    from aurora_amd import scenes, tuning
    from aurora_amd.renderer import Renderer, batch_submit
    paths = scenes.batch_jobs(str(tmp_path), n_jobs=64, seed=3)
    cfg = tuning.choose_batch()
    rs, info = [], []
    for k, p in enumerate(paths):
        r = Renderer(p, os.path.join(os.path.dirname(p), "config.json"))
        r.build(width=cfg["bvh_width"], max_leaf=cfg["max_leaf"])
        r.set_seed(1000 + k)
        rs.append(r)
        info.append(r.info())
    res_set = {(i["width"], i["height"]) for i in info}
    depths = {i["maxdepth"] for i in info}
    assert len(res_set) == 4 and min(depths) <= 4 and max(depths) >= 11
    ex = Renderer()
    tuning.apply_launch(ex, cfg)
    print("batch: 64 jobs", flush=True)
    res, tot = batch_submit(ex, [(r, 0, i["pixelsamples"]) for r, i in zip(rs, info)])
    fbs = [r.framebuffer() for r in rs]
    assert tot["paths"] == sum(i["width"] * i["height"] * i["pixelsamples"] for i in info)
    print("batch: per-scene renders", flush=True)
    for k, (r, i) in enumerate(zip(rs, info)):
        tuning.apply_launch(r, cfg)
        acc, st = r.render_passes(0, i["pixelsamples"])
        assert np.array_equal(fbs[k], acc), k
        for key in ("paths", "ext_rays", "hit_rays", "shadow_rays"):
            assert res[k][key] == st[key], (k, key)
    pick = {int(np.argmin([i["width"] for i in info])), int(np.argmax([i["width"] for i in info])),
            int(np.argmax([i["maxdepth"] for i in info]))}
    rng = np.random.default_rng(3)
    for k in sorted(pick):
        i = info[k]
        s = oracle.Scene(rs[k].export(), bvh=True)
        pix = rng.integers(0, i["width"] * i["height"], 2000).astype(np.int32)
        smp = rng.integers(0, i["pixelsamples"], 2000).astype(np.int32)
        g, _ = rs[k].render_paths(pix, smp)
        o, _ = s.render_paths(1000 + k, pix, smp, threads=16)
        rel = np.abs(g - o) / (np.abs(o) + 1e-3)
        assert np.mean(rel > 1e-4) < 5e-3, (k, np.mean(rel > 1e-4))
    for r in rs:
        r.close()
    ex.close()
