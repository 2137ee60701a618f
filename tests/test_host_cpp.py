"""The C++ host side (aurora_amd/host): the Aurora command line, the
counterpart of Aurora/main.cpp:36-52, and the header-only classes of
aurora_gpu.hpp driven from a C++ program (tests/cpp/host_api_check.cpp).

CPU tests cover the argument handling, which touches no device; GPU tests
render through the command line and trace through the C++ classes."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, random_rays

AURORA = os.path.join(ROOT, "aurora_amd", "host", "aurora")
CHECK = os.path.join(ROOT, "tests", "cpp", "host_api_check")


def _run(args, timeout=120):
    if not os.path.exists(args[0]):
        raise FileNotFoundError("%s is not built; run `python -m aurora_amd.build`" % args[0])
    return subprocess.run(args, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=timeout)


def test_command_line_usage():
    """A wrong argument count prints the usage line (main.cpp:40-41) and
    exits 1 before any device is touched."""
    for args in ([], ["a", "b", "c"]):
        r = _run([AURORA] + args)
        assert r.returncode == 1
        assert r.stderr.startswith("usage: ") and "<filename>" in r.stderr


def test_command_line_missing_file(tmp_path):
    """A scene file that does not open is reported (main.cpp:45-47), exit 1."""
    missing = str(tmp_path / "nope.asc")
    r = _run([AURORA, missing])
    assert r.returncode == 1
    assert r.stderr.strip() == "Could not open file %s." % missing


@pytest.fixture(scope="module")
def room(tmp_path_factory):
    from aurora_amd import scenes
    d = str(tmp_path_factory.mktemp("host_cpp"))
    path = scenes.box_room(res=48, spp=5, maxdepth=3, mindepth=3).write(d)
    os.makedirs(os.path.join(d, "render"), exist_ok=True)
    return path, os.path.join(d, "config.json"), d


@pytest.mark.gpu
def test_command_line_renders_the_scene(room):
    """`aurora <scene> <config>` writes the scene's options filename with the
    FP16 display of all passes, bit for bit what the Python Renderer's display
    holds after the same passes, and prints the build time and stats."""
    from aurora_amd import exr
    from aurora_amd.renderer import Renderer
    scene_file, config, d = room
    print("host command line: box room 48x48, 5 spp, depth 3", flush=True)
    r = _run([AURORA, scene_file, config])
    assert r.returncode == 0, r.stderr
    assert "Total time building render env:" in r.stdout
    assert "Rendered %d paths" % (48 * 48 * 5) in r.stdout
    out = os.path.join(d, "render", "render.exr")
    img = exr.read_exr(out)
    with Renderer(scene_file, config) as ren:
        ren.build()
        ren.render_passes(0, 5, fetch=False)
        want = ren.display()
    assert np.array_equal(img.astype(np.float16).view(np.uint16), want.view(np.uint16))


@pytest.mark.gpu
def test_command_line_reports_library_errors(tmp_path):
    """A scene the loader refuses ends the program with exit 2 and the
    library's message and code."""
    bad = tmp_path / "bad.asc"
    bad.write_text('{"options": {"resolution": [4, 4]}, "cameras": [')
    print("host command line: malformed scene", flush=True)
    r = _run([AURORA, str(bad)])
    assert r.returncode == 2
    assert "aur_scene_load" in r.stderr and "(code -3)" in r.stderr


@pytest.mark.gpu
def test_cpp_acceleration_structure_matches_oracle(gpu_box_room, box_room_dir, oracle, tmp_path):
    """AccelerationStructure::intersect / intersectBinary of aurora_gpu.hpp,
    called from C++: a query before the build throws AUR_ESTATE; after it the
    closest hits equal the oracle's brute force bit for bit and any-hit says
    blocked exactly where a closest hit exists."""
    s = oracle.Scene(gpu_box_room.export())
    rng = np.random.default_rng(5)
    rays = random_rays(rng, 20000, [-0.99, -0.99, -4.49], [0.99, 0.99, -2.51]).astype(np.float32)
    rays[::3, 7] = rng.uniform(0.05, 1.0, len(rays[::3])).astype(np.float32)  # short segments
    rin, rout = tmp_path / "rays.bin", tmp_path / "hits.bin"
    rays.tofile(str(rin))
    print("host api check: box room, 20000 rays", flush=True)
    r = _run([CHECK, os.path.join(box_room_dir, "scene.asc"), os.path.join(box_room_dir, "config.json"),
              str(rin), str(rout)])
    assert r.returncode == 0, r.stderr
    assert r.stdout.startswith("state error: ")
    raw = np.fromfile(str(rout), np.uint8)
    n = len(rays)
    assert raw.size == n * 16 + n
    hits = raw[:n * 16].view([("t", np.float32), ("u", np.float32), ("v", np.float32), ("prim", np.int32)])
    blocked = raw[n * 16:].astype(bool)
    o = s.trace(rays)
    hit = o["prim"] >= 0
    assert 0.2 < hit.mean() < 1.0
    assert np.array_equal(hits["prim"], o["prim"])
    for k in ("t", "u", "v"):
        assert np.array_equal(hits[k][hit], o[k][hit]), k
    assert np.array_equal(blocked, hit)
