"""The Python 3 PyAurora counterpart (aurora_amd.pyaurora): the JSON its
classes serialise, checked key by key against the shape PyAurora's own
``serialize`` methods produce (PyAurora/core.py:15-51, lights.py:28-33,
geometry.py:32-39, shaders.py:11-18, 43-50), and that a scene authored with
it loads into exactly the scene the loader builds from the equivalent
hand-written file."""
import json
import math
import os

import numpy as np
import pytest

from conftest import host_renderer


def author_box_room(mesh_dir, res=256, spp=16, maxdepth=5, mindepth=5, filename="$RENDER/render.exr"):
    """The closed box room (aurora_amd.scenes.box_room) authored through the
    PyAurora API, over the OBJ files in ``mesh_dir``."""
    from aurora_amd.pyaurora import core, geometry, lights, shaders
    scene = core.Scene()
    scene.setOptions({"resolution": [res, res], "pixelsamples": spp, "lightsamples": 1,
                      "mindepth": mindepth, "maxdepth": maxdepth, "filename": filename})
    cam = core.Camera(fov=math.tan(math.radians(20.0)))
    cam.translate(0, 1, 3.5)
    scene.setCamera(cam)
    white = shaders.MatteMaterial("white", [0.73, 0.73, 0.73])
    red = shaders.MatteMaterial("red", [0.63, 0.12, 0.05])
    green = shaders.MatteMaterial("green", [0.14, 0.45, 0.09])
    boxmatte = shaders.MatteMaterial("boxmatte", [0.73, 0.73, 0.73])
    boxspec = shaders.MetalMaterial("boxspec", [0.04, 0.04, 0.04], 0.3, 0.3, 0.5, [1.0, 1.0, 1.0])
    boxspec.setDistribution("ggx")
    for m in (white, red, green, boxmatte, boxspec):
        scene.appendMaterial(m)
    for name, mat in (("walls_white", white), ("wall_left", red), ("wall_right", green),
                      ("box_tall", boxmatte), ("box_short", boxspec)):
        g = geometry.ObjMesh(name, 0, os.path.join(mesh_dir, name + ".obj"))
        g.setMaterial(mat if name != "wall_left" else "red")  # an instance or a name
        scene.append(g)
    light = lights.SquareLight("light_ceiling", 0.25, 0.25, [17, 12, 4], 0)
    light.translate(0, 1.98, 0)
    light.rotateX(-90)
    scene.append(light)
    return scene


def test_serialised_shape_matches_pyaurora():
    """Each key lands where PyAurora puts it: the camera's keys directly in
    renderCam, objects / materials / shaders as objects keyed by name, lights
    and meshes with their own "name"; an entry's parameters are written as
    they are (an empty light transform list stays a list) while the scene's
    and the camera's lists become objects (an empty one {})."""
    from aurora_amd.pyaurora import core, geometry, lights, shaders
    scene = core.Scene()
    scene.setOptions({"resolution": [4, 2], "filename": "x.exr"})
    cam = core.Camera()
    cam.rotateY(10)
    cam.translate(1, 2, 3)
    scene.setCamera(cam)
    tex = shaders.Texture2DShader("tex", "color", "a.exr")
    tri = shaders.TriplanarTextureShader("tri", "float", "b.exr", 2.5)
    scene.appendShader(tex)
    scene.appendShader(tri)
    kel = shaders.KelemenMaterial("kel", "tex", [1, 0.5, 0.25], "tri", 0.04)
    kel.setNormalMap("tex")
    scene.appendMaterial(kel)
    scene.appendMaterial(shaders.GlassMaterial("glass"))
    mesh = geometry.ObjMesh("m", 2, "m.obj")
    mesh.setMaterial("kel")
    mesh.scale(2, 2, 2)
    scene.append(mesh)
    scene.append(lights.EnvLight("sky", [1, 1, 1], -1, ""))
    got = json.loads(json.dumps(scene.toJson()))
    want = {
        "options": {"resolution": [4, 2], "filename": "x.exr"},
        "renderCam": {"transforms": [{"rotateY": 10.0}, {"translate": [1.0, 2.0, 3.0]}], "fov": math.pi / 4},
        "objects": {
            "m": {"path": "m.obj", "subdlevels": 2, "name": "m", "transforms": [{"scale": [2.0, 2.0, 2.0]}],
                  "material": "kel", "type": "geometry"},
            "sky": {"color": [1.0, 1.0, 1.0], "exposure": -1.0, "envmap": "", "transforms": [],
                    "type": "envlight", "name": "sky"},
        },
        "materials": {
            "kel": {"diffusecolor": "tex", "speccolor": [1.0, 0.5, 0.25], "roughness": "tri",
                    "reflectance": 0.04, "type": "kelemenMaterial", "normalmap": "tex"},
            "glass": {"specColor": [1.0, 1.0, 1.0], "transmitColor": [1.0, 1.0, 1.0], "reflectance": 1.0,
                      "ior": 1.333, "type": "glassMaterial"},
        },
        "shaders": {
            "tex": {"texturename": "a.exr", "type": "texture2DShader", "attributetype": "color"},
            "tri": {"texturename": "b.exr", "scale": 2.5, "type": "triplanarTextureShader",
                    "attributetype": "float"},
        },
    }
    assert got == want
    # key order inside each entry follows PyAurora's serialize lists
    assert list(got["objects"]["m"]) == ["path", "subdlevels", "name", "transforms", "material", "type"]
    assert list(got["materials"]["kel"])[-1] == "normalmap"
    with pytest.raises(ValueError):
        kel.setDistribution("phong")


def test_box_room_authored_with_pyaurora_loads_as_the_scene_file(box_room_dir, box_room_export, tmp_path):
    """The same box room, once as the hand-written scene file and once authored
    through the PyAurora API, flattens to identical geometry, materials and
    lights."""
    from aurora_amd.pyaurora import core
    scene = author_box_room(os.path.join(box_room_dir, "meshes"))
    path = core.write(scene, str(tmp_path / "tmp.asc"))
    r = host_renderer()
    r.load(path, os.path.join(box_room_dir, "config.json"))
    e = r.export()
    info = r.info()
    r.close()
    want = box_room_export
    for k in ("verts", "tris", "normals", "attr"):
        assert np.array_equal(e[k], want[k]), k
    assert len(e["materials"]) == len(want["materials"])
    for a, b in zip(e["materials"], want["materials"]):
        assert bytes(a) == bytes(b)
    assert [bytes(l) for l in e["lights"]] == [bytes(l) for l in want["lights"]]
    assert (info["width"], info["height"], info["pixelsamples"]) == (256, 256, 16)


def test_loader_refuses_what_the_reference_parser_cannot_read(tmp_path):
    """TriangleMesh (inline points, no OBJ path) and the commented-out car and
    transmit materials serialise as PyAurora writes them but do not load."""
    from aurora_amd import _lib as L
    from aurora_amd.pyaurora import core, geometry, shaders
    for bad in ("triangles", "car", "transmit"):
        scene = core.Scene()
        scene.setOptions({"resolution": [8, 8], "pixelsamples": 1, "lightsamples": 1, "mindepth": 1,
                          "maxdepth": 2, "filename": "x.exr"})
        scene.setCamera(core.Camera())
        if bad == "triangles":
            m = geometry.TriangleMesh("t", [0, 0, 0, 1, 0, 0, 0, 1, 0], [0, 1, 2])
            scene.appendMaterial(shaders.MatteMaterial("m"))
        else:
            mat = shaders.CarMaterial("m") if bad == "car" else shaders.TransmitMaterial("m")
            scene.appendMaterial(mat)
            m = geometry.ObjMesh("t", 0, "t.obj")
        m.setMaterial("m")
        scene.append(m)
        path = core.write(scene, str(tmp_path / (bad + ".asc")))
        r = host_renderer()
        with pytest.raises(L.AuroraError) as ei:
            r.load(path)
        r.close()
        msg = str(ei.value)
        assert ("missing json attribute" in msg) if bad == "triangles" else ("unsupported material type" in msg), msg
