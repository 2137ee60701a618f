"""Scene authoring, Python 3: serialisation, Scene, Camera and render().

Counterpart of PyAurora/core.py.  The classes keep PyAurora's names,
constructor arguments and the JSON shape its ``serialize`` produces, which is
the shape the scene loader (``aur_scene_load``, restating jsonParser.cpp)
reads.  ``render`` writes the scene file and renders it through
libaurora_hip.so instead of starting the Aurora executable.
"""
from __future__ import annotations

import json
import math
from typing import Optional


class Serializable:
    """A named bag of JSON parameters (core.py:15-51).

    Every key in ``serialize`` starts as an empty list in ``self._s``; the
    subclass overwrites the ones that hold a value.  ``serialize(root)``
    writes the keys into ``root``: a list of dicts (transforms) as the list,
    any other list as an object holding one entry per item (each item
    serialises itself into it, under its own name), and scalars, strings and
    dicts unchanged.  An empty list therefore becomes ``{}``, as in PyAurora.
    """

    def __init__(self, name, serialize):
        self._s = {key: [] for key in serialize}
        self._s["name"] = str(name)
        self._serialize = list(serialize)

    def getName(self) -> str:
        return self._s["name"]

    def serialize(self, root: dict) -> None:
        for key in self._serialize:
            root[key] = _value(self._s[key])


def _value(v):
    """The JSON value of one parameter of the scene or the camera
    (core.py:36-51)."""
    if isinstance(v, list):
        if any(isinstance(item, dict) for item in v):
            return v
        out = {}
        for item in v:
            item.serialize(out)
        return out
    return _entry_value(v)


def _entry_value(v):
    """The JSON value of one parameter of a mesh, light, shader or material:
    written as it is (lights.py:28-33, geometry.py:32-39, shaders.py:11-18)."""
    if isinstance(v, Serializable):
        # PyAurora nests the object's own entry here, which its parser cannot
        # read (getStringAttr, jsonParser.cpp:227); the loader wants the name
        return v.getName()
    return v


class _Transformable:
    """translate / scale / rotateX / rotateY / rotateZ appended to the
    "transforms" list in call order (core.py:82-95; the loader composes
    them as jsonParser.cpp:168-223 does, angles in degrees)."""

    def translate(self, x, y, z):
        self._s["transforms"].append({"translate": [float(x), float(y), float(z)]})

    def scale(self, x, y, z):
        self._s["transforms"].append({"scale": [float(x), float(y), float(z)]})

    def rotateX(self, x):
        self._s["transforms"].append({"rotateX": float(x)})

    def rotateY(self, x):
        self._s["transforms"].append({"rotateY": float(x)})

    def rotateZ(self, x):
        self._s["transforms"].append({"rotateZ": float(x)})


class Scene(Serializable):
    """The scene root: options, renderCam, objects, materials, shaders
    (core.py:59-78)."""

    def __init__(self):
        Serializable.__init__(self, "scene", ["options", "renderCam", "objects", "materials", "shaders"])
        self._s["options"] = {}

    def setOptions(self, options: dict) -> None:
        """The "options" object: resolution [w, h], filename, pixelsamples,
        lightsamples, mindepth, maxdepth (jsonParser.cpp:431-475)."""
        self._s["options"] = options

    def setCamera(self, camera: "Camera") -> None:
        self._s["renderCam"] = [camera]

    def append(self, obj) -> None:
        """A geometry.Mesh or a lights.Light."""
        self._s["objects"].append(obj)

    def appendShader(self, shader) -> None:
        self._s["shaders"].append(shader)

    def appendMaterial(self, material) -> None:
        self._s["materials"].append(material)

    def toJson(self) -> dict:
        """The serialised scene, as ``render`` writes it."""
        root = {}
        self.serialize(root)
        return root


class Camera(_Transformable, Serializable):
    """The render camera: field of view in radians plus transforms
    (core.py:80-95, jsonParser.cpp:477-505)."""

    def __init__(self, fov=math.pi / 4):
        Serializable.__init__(self, "camera", ["transforms", "fov"])
        self._s["fov"] = float(fov)


def write(scene: Scene, path: str) -> str:
    """Writes the scene file as PyAurora's ``render`` does (core.py:103-116:
    ``json.dumps(root, indent=4)``, no header comment)."""
    with open(path, "w") as fh:
        fh.write(json.dumps(scene.toJson(), indent=4))
    return path


def render(scene: Scene, scene_file: str = "tmp.asc", config: Optional[str] = None, device: int = 0,
           draw_delay: float = 1.0, image: Optional[str] = None) -> dict:
    """PyAurora's ``render(scene)`` (core.py:103-120): writes the scene file
    (``tmp.asc`` in the working directory unless ``scene_file`` is given),
    then loads, builds and renders it as the Aurora executable would
    (Renderer::render: parse, build the render environment, renderImage with
    its progressive EXR writes, outputStats, postRenderCleanup).  ``config``
    is the config.json that resolves $BASE/$RENDER/... in paths; ``image``
    overrides the options filename.  Returns the render's counters plus the
    number of image writes (``draws``)."""
    from .. import tuning
    from ..renderer import Renderer

    write(scene, scene_file)
    with Renderer(scene_file, config=config, device=device) as r:
        cfg = tuning.choose(r)
        r.build(width=cfg["bvh_width"], max_leaf=cfg["max_leaf"])
        tuning.apply_launch(r, cfg)
        return r.render_image(image, draw_delay=draw_delay)
