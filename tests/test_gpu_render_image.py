"""Whole renders through aur_render_image (Renderer::renderImage with its
progressive EXR writes, outputStats and postRenderCleanup, renderer.cpp:200-304)
and through the PyAurora counterpart's core.render."""
import os
import re
import struct

import numpy as np
import pytest

from test_pyaurora import author_box_room


def _string_attrs(path):
    """The string attributes of an EXR header (name, type "string", size, bytes)."""
    data = open(path, "rb").read(1 << 16)
    pos, out = 8, {}
    while True:
        end = data.index(b"\0", pos)
        name = data[pos:end].decode()
        pos = end + 1
        if not name:
            return out
        end = data.index(b"\0", pos)
        typ = data[pos:end].decode()
        pos = end + 1
        (size,) = struct.unpack_from("<i", data, pos)
        pos += 4
        if typ == "string":
            out[name] = data[pos:pos + size].decode()
        pos += size


@pytest.fixture(scope="module")
def room(tmp_path_factory):
    from aurora_amd import scenes
    d = str(tmp_path_factory.mktemp("render_image"))
    path = scenes.box_room(res=64, spp=6, maxdepth=3, mindepth=3).write(d)
    os.makedirs(os.path.join(d, "render"), exist_ok=True)
    return path, os.path.join(d, "config.json"), d


def _display_of_full_render(scene_file, config):
    from aurora_amd.renderer import Renderer
    r = Renderer(scene_file, config)
    r.build()
    r.render_passes(0, r.info()["pixelsamples"], fetch=False)
    d = r.display()
    r.close()
    return d


@pytest.mark.gpu
def test_render_image_writes_display_with_stats(room):
    """One write after pass 0 and the final one (a delay no pass reaches); the
    file holds the FP16 display of all passes bit for bit, the options
    filename ($RENDER/render.exr) is used when no path is given, and the final
    write carries renderTime and RaySpeed."""
    from aurora_amd import exr
    from aurora_amd.renderer import Renderer
    scene_file, config, d = room
    print("render_image: box room 64x64, 6 spp, depth 3", flush=True)
    r = Renderer(scene_file, config)
    r.build()
    st = r.render_image(None, draw_delay=1e6)
    r.close()
    out = os.path.join(d, "render", "render.exr")
    assert st["draws"] == 2
    assert st["paths"] == 64 * 64 * 6 and st["ext_rays"] > 0 and st["shadow_rays"] > 0
    want = _display_of_full_render(scene_file, config)
    img = exr.read_exr(out)
    assert np.array_equal(img.astype(np.float16).view(np.uint16), want.view(np.uint16))
    meta = _string_attrs(out)
    assert re.fullmatch(r"\d+ h \d+ min \d+ sec\.", meta["renderTime"]), meta
    m = re.fullmatch(r"(\d+)K rays/sec\.", meta["RaySpeed"])
    assert m and int(m.group(1)) > 0, meta
    assert "renderer" in meta
    assert not os.path.exists(out + "_tmp")


@pytest.mark.gpu
def test_render_image_rewrites_after_every_pass_past_the_delay(room, tmp_path):
    """With a delay every pass exceeds, each pass is traced alone and written
    (as the reference's per-pass check does), plus the final write; the final
    image does not depend on the chunking."""
    from aurora_amd import exr
    from aurora_amd.renderer import Renderer
    scene_file, config, _ = room
    out = str(tmp_path / "progressive.exr")
    r = Renderer(scene_file, config)
    r.build()
    st = r.render_image(out, draw_delay=1e-9)
    r.close()
    assert st["draws"] == 6 + 1
    want = _display_of_full_render(scene_file, config)
    assert np.array_equal(exr.read_exr(out).astype(np.float16).view(np.uint16), want.view(np.uint16))


@pytest.mark.gpu
def test_pyaurora_render_matches_library_render(room, tmp_path):
    """core.render writes the authored scene file and renders it: the image is
    the display of the same scene file rendered through the Renderer."""
    from aurora_amd import exr
    from aurora_amd.pyaurora import core
    scene_file, config, d = room
    scene = author_box_room(os.path.join(d, "meshes"), res=64, spp=6, maxdepth=3, mindepth=3)
    asc = str(tmp_path / "tmp.asc")
    out = str(tmp_path / "pyaurora.exr")
    st = core.render(scene, scene_file=asc, config=config, image=out)
    assert os.path.exists(asc) and st["draws"] >= 2
    want = _display_of_full_render(scene_file, config)
    assert np.array_equal(exr.read_exr(out).astype(np.float16).view(np.uint16), want.view(np.uint16))
