"""Host logic that runs without a GPU: scene input, flattening, error codes,
and the exported C ABI."""
import json
import math
import os
import re

import numpy as np
import pytest

from conftest import ROOT, host_renderer


def test_library_exports_every_declared_symbol(lib):
    hdr = open(os.path.join(ROOT, "include", "aurora_hip.h")).read()
    names = sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*\*?\s*(aur_\w+)\s*\(", hdr, re.M)))
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    from aurora_amd import _lib as L
    assert set(names) == set(L.SIGNATURES), "binding table and header differ"


def test_box_room_flattening_matches_world_geometry(box_room_dir, box_room_export):
    """Camera transform is translate(0, 1, 3.5): camera space = world - (0, 1, 3.5).
    Objects come in sorted key order, then the light quad (renderer.cpp:118-134)."""
    from aurora_amd import scenes
    spec = scenes.box_room()
    order = sorted(k for k, o in spec.objects.items() if o["type"] == "geometry")
    assert order == ["box_short", "box_tall", "wall_left", "wall_right", "walls_white"]
    world = []
    for name in order:
        v, f, _ = spec.meshes[name + ".obj"]
        v = np.asarray(v, np.float64)
        for tri in f:
            world.extend(v[list(tri)])
    world = np.array(world)
    e = box_room_export
    flat = e["verts"][e["tris"][:len(world) // 3].ravel()]
    np.testing.assert_allclose(flat, world - np.array([0.0, 1.0, 3.5]), atol=2e-6)
    # light quad: translate(0, 1.98, 0) rotateX(-90) of (+-0.25, +-0.25, 0); tris (p1,p2,p4), (p3,p2,p4)
    a = math.radians(-90.0)
    R = np.array([[1, 0, 0], [0, math.cos(a), -math.sin(a)], [0, math.sin(a), math.cos(a)]])
    corners = np.array([[-0.25, -0.25, 0], [-0.25, 0.25, 0], [0.25, 0.25, 0], [0.25, -0.25, 0]])
    cam = corners @ R.T + np.array([0, 1.98, 0]) - np.array([0, 1, 3.5])
    lt = e["tris"][-2:]
    np.testing.assert_allclose(e["verts"][lt[0]], cam[[0, 1, 3]], atol=2e-6)
    np.testing.assert_allclose(e["verts"][lt[1]], cam[[2, 1, 3]], atol=2e-6)
    assert list(e["attr"][-2:]) == [5, 5]
    assert e["materials"][5].type == 3 and list(e["materials"][5].emission) == [17.0, 12.0, 4.0]
    # the metal box is the first object (sorted keys): attr 0 is the GGX Cook-Torrance material
    assert e["materials"][0].type == 1 and e["materials"][0].distribution == 1
    # normals are per corner and faceforwarding happens at hit time, so only directions matter
    n = e["normals"].reshape(-1, 3)
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-5)


def test_transform_composition_is_post_multiplied(tmp_path):
    """transforms [translate(1,2,3), rotateY(90), scale(2,1,1)] map p to T*R*S*p (transform.cpp:44-104)."""
    from aurora_amd import scenes
    spec = scenes.box_room()
    spec.objects = {"tri": {"type": "geometry", "name": "tri", "path": "meshes/tri.obj", "material": "white",
                            "subdlevels": 0,
                            "transforms": [{"translate": [1.0, 2.0, 3.0]}, {"rotateY": 90.0}, {"scale": [2.0, 1.0, 1.0]}]}}
    spec.meshes = {"tri.obj": ([[1, 0, 0], [0, 1, 0], [0, 0, 1]], [[0, 1, 2]], None)}
    spec.camera = {"fov": 0.5, "transforms": []}
    path = spec.write(str(tmp_path))
    r = host_renderer()
    r.load(path, str(tmp_path / "config.json"))
    e = r.export()
    r.close()
    T = np.eye(4); T[:3, 3] = [1, 2, 3]
    a = math.radians(90.0)
    Ry = np.array([[math.cos(a), 0, math.sin(a), 0], [0, 1, 0, 0], [-math.sin(a), 0, math.cos(a), 0], [0, 0, 0, 1]])
    S = np.diag([2.0, 1, 1, 1])
    M = T @ Ry @ S
    p = np.array([[1, 0, 0, 1], [0, 1, 0, 1], [0, 0, 1, 1]], np.float64)
    np.testing.assert_allclose(e["verts"], (p @ M.T)[:, :3], atol=1e-6)


def test_config_templates_and_frame_numbering(tmp_path):
    from aurora_amd import scenes
    spec = scenes.box_room()
    spec.options["filename"] = "$RENDER//out.####.exr"
    path = spec.write(str(tmp_path))
    cfg = json.load(open(tmp_path / "config.json"))
    render_dir = cfg["basepath"] + "render/"
    os.makedirs(render_dir, exist_ok=True)
    open(render_dir + "out.0001.exr", "w").close()
    r = host_renderer()
    r.load(path, str(tmp_path / "config.json"))  # resolves without error
    assert r.info()["width"] == 256
    r.close()


def _expect_error(fn, code):
    from aurora_amd._lib import AuroraError
    with pytest.raises(AuroraError) as ei:
        fn()
    assert ei.value.code == code, str(ei.value)


def test_error_codes(tmp_path):
    from aurora_amd import scenes
    r = host_renderer()
    _expect_error(lambda: r.load(str(tmp_path / "missing.asc")), -2)
    bad = tmp_path / "bad.asc"
    bad.write_text("{ \"options\": { \"pixelsamples\": 4, ")
    _expect_error(lambda: r.load(str(bad)), -3)
    spec = scenes.box_room()
    spec.objects["walls_white"]["material"] = "nonexistent"
    p = spec.write(str(tmp_path / "s1"))
    _expect_error(lambda: r.load(p), -3)
    spec = scenes.box_room()
    spec.meshes.pop("wall_left.obj")
    spec.write(str(tmp_path / "s2"))
    os.remove(str(tmp_path / "s2" / "meshes" / "wall_left.obj")) if os.path.exists(
        str(tmp_path / "s2" / "meshes" / "wall_left.obj")) else None
    _expect_error(lambda: r.load(str(tmp_path / "s2" / "scene.asc")), -2)
    # GPU work on a host-only context is refused, not crashed
    good = scenes.box_room().write(str(tmp_path / "s3"))
    r.load(good)
    _expect_error(lambda: r.build(), -5)
    _expect_error(lambda: r.render_passes(0, 1), -5)
    r.close()


def test_non_finite_radiance_is_refused(tmp_path, box_room_export):
    """color * 2^exposure overflows float at exposure >= 128 (squareLight.h:84-86),
    a texel above 65504 is +inf as FP16: both would give infinite radiance, so
    loading or uploading such a scene fails with AUR_EINVAL."""
    from aurora_amd import scenes
    r = host_renderer()
    spec = scenes.box_room()
    light = next(k for k, o in spec.objects.items() if o["type"] == "light")
    spec.objects[light]["exposure"] = 200.0
    _expect_error(lambda: r.load(spec.write(str(tmp_path / "exp"))), -1)
    spec.objects[light]["exposure"] = 100.0  # 17 * 2^100 is finite
    r.load(spec.write(str(tmp_path / "exp_ok")))
    e = dict(box_room_export)
    lights = [type(l).from_buffer_copy(l) for l in e["lights"]]
    lights[0].exposure = 130.0
    e["lights"] = lights
    _expect_error(lambda: r.upload(e), -1)
    r.close()


def test_upload_rejects_bad_indices(box_room_export):
    e = dict(box_room_export)
    e["tris"] = e["tris"].copy()
    e["tris"][0, 0] = 10 ** 6
    r = host_renderer()
    _expect_error(lambda: r.upload(e), -1)
    e["tris"] = box_room_export["tris"]
    e["attr"] = box_room_export["attr"].copy()
    e["attr"][3] = 99
    _expect_error(lambda: r.upload(e), -1)
    r.close()


def test_json_comments_and_sorted_members(tmp_path):
    from aurora_amd import scenes
    spec = scenes.box_room()
    path = spec.write(str(tmp_path))
    text = open(path).read()
    open(path, "w").write("// AURORA SCENE FILE\n// generated by a test\n" + text)
    r = host_renderer()
    r.load(path)
    e = r.export()
    r.close()
    assert e["verts"].shape[0] > 0 and len(e["materials"]) == 6


def test_launch_options_validation():
    r = host_renderer()
    d = r.get_launch()
    assert d == {"trace_block": 64, "trace_waves_per_cu": 32, "pool_paths": 128 << 20, "shade_list_tiles": 4096}
    r.set_launch(trace_block=256, pool_paths=1 << 16)
    assert r.get_launch() == {"trace_block": 256, "trace_waves_per_cu": 32, "pool_paths": 1 << 16,
                              "shade_list_tiles": 4096}
    r.set_launch(shade_list_tiles=1)
    assert r.get_launch()["shade_list_tiles"] == 1
    _expect_error(lambda: r.set_launch(trace_block=100), -1)
    _expect_error(lambda: r.set_launch(trace_waves_per_cu=65), -1)
    _expect_error(lambda: r.set_launch(pool_paths=10), -1)
    # shadow records are indexed 3 * (2 * pool) in 32-bit ints
    _expect_error(lambda: r.set_launch(pool_paths=0x7FFFFFFF // 6 + 1), -1)
    r.set_launch(pool_paths=0x7FFFFFFF // 6)
    assert r.get_launch()["trace_block"] == 256
    r.close()


def test_batch_submit_errors_without_gpu(box_room_dir):
    import ctypes as C
    from aurora_amd import _lib as L
    from aurora_amd.renderer import batch_submit
    r = host_renderer()
    r.load(os.path.join(box_room_dir, "scene.asc"), os.path.join(box_room_dir, "config.json"))
    _expect_error(lambda: batch_submit(r, []), -1)
    _expect_error(lambda: batch_submit(r, [(r, 0, 1)]), -5)   # host-only context
    res = (L.JobResult * 1)()
    assert r.lib.aur_batch_submit(r._ctx, None, 1, None, res, None) == -1
    _expect_error(lambda: r.framebuffer(), -5)
    r.close()
    del C


def test_batch_workload_generator_ranges():
    from aurora_amd import scenes
    tris, res, spp, depth = [], set(), [], []
    for k in range(64):
        s = scenes.batch_job(3, k)
        n = 0
        for o in s.objects.values():
            if o.get("type") != "geometry":
                continue
            if "icosahedron" in o["path"]:
                n += 20 * 4 ** o["subdlevels"]
            elif "room" in o["path"]:
                n += 10
            else:
                n += 12
        tris.append(n + 2)  # + the light quad
        res.add(tuple(s.options["resolution"]))
        spp.append(s.options["pixelsamples"])
        depth.append(s.options["maxdepth"])
        assert s.options["mindepth"] == s.options["maxdepth"]
    assert res <= {(128, 128), (256, 256), (512, 512), (1024, 1024)} and len(res) == 4
    assert 4 <= min(spp) and max(spp) <= 256 and 3 <= min(depth) and max(depth) <= 12
    assert 900 <= min(tris) and max(tris) <= 700000
    assert scenes.batch_job(3, 7).options == scenes.batch_job(3, 7).options
