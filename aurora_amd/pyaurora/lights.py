"""Lights, Python 3 counterpart of PyAurora/lights.py.

A light serialises into the scene's "objects" under its name, with its
"name", "type" and "transforms" (lights.py:3-33); the loader reads type
"light" as a SquareLight and "envlight" as an InfiniteAreaLight
(jsonParser.cpp:243-282, 634-650).
"""
from __future__ import annotations

from .core import Serializable, _Transformable, _entry_value


def _floats(color):
    return [float(c) for c in color]


class Light(_Transformable, Serializable):
    def __init__(self, name, serialize):
        Serializable.__init__(self, name, list(serialize) + ["transforms", "type", "name"])
        self._s["type"] = "light"

    def serialize(self, root: dict) -> None:
        root[self.getName()] = {key: _entry_value(self._s[key]) for key in self._serialize}


class SquareLight(Light):
    """A quad of half extents scaleX x scaleY in its local xy plane,
    radiance color * 2^exposure (lights.py:36-47, squareLight.cpp)."""

    def __init__(self, name, scaleX, scaleY, color, exposure):
        Light.__init__(self, name, ["scaleX", "scaleY", "color", "exposure"])
        self._s["color"] = _floats(color)
        self._s["exposure"] = float(exposure)
        self._s["scaleX"] = float(scaleX)
        self._s["scaleY"] = float(scaleY)


class EnvLight(Light):
    """Environment light: color * 2^exposure, times the lat-long EXR
    ``envmap`` when the path is not empty (lights.py:49-58,
    infiniteAreaLight.cpp)."""

    def __init__(self, name, color, exposure, envmap):
        Light.__init__(self, name, ["color", "exposure", "envmap"])
        self._s["color"] = _floats(color)
        self._s["exposure"] = float(exposure)
        self._s["type"] = "envlight"
        self._s["envmap"] = envmap
