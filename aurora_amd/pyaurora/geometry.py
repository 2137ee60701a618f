"""Meshes, Python 3 counterpart of PyAurora/geometry.py.

A mesh serialises into the scene's "objects" under its name with "name",
"transforms", "material" and "type": "geometry" (geometry.py:3-39).
"""
from __future__ import annotations

from .core import Serializable, _Transformable, _entry_value


class Mesh(_Transformable, Serializable):
    def __init__(self, name, serialize):
        Serializable.__init__(self, name, list(serialize) + ["name", "transforms", "material", "type"])
        self._s["type"] = "geometry"

    def setMaterial(self, m) -> None:
        """The material's name, or the material itself (written by name)."""
        self._s["material"] = m

    def serialize(self, root: dict) -> None:
        root[self.getName()] = {key: _entry_value(self._s[key]) for key in self._serialize}


class ObjMesh(Mesh):
    """An OBJ file, Loop-subdivided ``subdlevels`` times at load
    (geometry.py:42-47, jsonParser.cpp:225-241, objTriangleMesh.cpp)."""

    def __init__(self, name, subdlevels, path):
        Mesh.__init__(self, name, ["path", "subdlevels"])
        self._s["path"] = path
        self._s["subdlevels"] = subdlevels


class TriangleMesh(Mesh):
    """Inline points and point indices (geometry.py:49-52).  The reference's
    parser reads geometry only from OBJ files (jsonParser.cpp:225-241), and
    so does the loader: a scene holding one is refused at load (its
    "subdlevels" and "path" are missing)."""

    def __init__(self, name, points, pointIndices):
        Mesh.__init__(self, name, ["points", "pointIndices"])
        self._s["points"] = points
        self._s["pointIndices"] = pointIndices
