"""Shaders and materials, Python 3 counterpart of PyAurora/shaders.py.

Each serialises under its name into the scene's "shaders" or "materials"
object (shaders.py:3-179).  A colour or float parameter given as a string
names a shader that drives it (jsonParser.cpp:286-314).
"""
from __future__ import annotations

from .core import Serializable, _entry_value


def _color(c):
    """Three floats, or a shader name left as it is."""
    return c if isinstance(c, str) else [float(c[0]), float(c[1]), float(c[2])]


class _Entry(Serializable):
    def serialize(self, root: dict) -> None:
        root[self.getName()] = {key: _entry_value(self._s[key]) for key in self._serialize}


class Texture2DShader(_Entry):
    """A uv-mapped EXR texture read as FP16 RGBA; attributetype "color" or
    "float" (shaders.py:3-18, texture2DShader.h)."""

    def __init__(self, name, attributetype, texturename):
        Serializable.__init__(self, name, ["texturename", "type", "attributetype"])
        self._s["texturename"] = texturename
        self._s["type"] = "texture2DShader"
        self._s["attributetype"] = attributetype


class TriplanarTextureShader(_Entry):
    """A texture projected along the object-space axes with ``scale``
    (shaders.py:20-35, triplanarTextureShader.h)."""

    def __init__(self, name, attributetype, texturename, scale):
        Serializable.__init__(self, name, ["texturename", "scale", "type", "attributetype"])
        self._s["texturename"] = texturename
        self._s["type"] = "triplanarTextureShader"
        self._s["scale"] = scale
        self._s["attributetype"] = attributetype


class Material(_Entry):
    def setNormalMap(self, normalmap) -> None:
        """A colour shader name (or a constant colour) perturbing the shading
        normal (shaders.py:37-40, material.cpp:27-40)."""
        self._s["normalmap"] = normalmap
        if "normalmap" not in self._serialize:
            self._serialize.append("normalmap")

    def setDistribution(self, distribution: str) -> None:
        """"beckmann" (the reference's microfacet distribution, the default)
        or "ggx" for the Cook-Torrance lobes of a metal or Kelemen material;
        an extension of the loader, without a PyAurora counterpart."""
        if distribution.lower() not in ("beckmann", "ggx"):
            raise ValueError("unknown microfacet distribution: " + distribution)
        self._s["distribution"] = distribution.lower()
        if "distribution" not in self._serialize:
            self._serialize.append("distribution")


class MatteMaterial(Material):
    def __init__(self, name, color=(1, 1, 1)):
        Serializable.__init__(self, name, ["color", "type"])
        self._s["color"] = _color(color)
        self._s["type"] = "matteMaterial"


class MetalMaterial(Material):
    """Two Cook-Torrance lobes (roughnessA, roughnessB) blended by ``mix``,
    tinted by Schlick's Fresnel of ``reflectance`` times ``gain``
    (shaders.py:61-75, metalMaterial.cpp:32-62)."""

    def __init__(self, name, reflectance, roughnessA, roughnessB, mix, gain):
        Serializable.__init__(self, name, ["reflectance", "roughnessA", "roughnessB", "mix", "gain", "type"])
        self._s["reflectance"] = _color(reflectance)
        self._s["roughnessA"] = roughnessA
        self._s["roughnessB"] = roughnessB
        self._s["mix"] = mix
        self._s["gain"] = gain
        self._s["type"] = "metalMaterial"


class TransmitMaterial(Material):
    """Serialised as PyAurora does (shaders.py:77-87); the reference's
    transmitMaterial is commented out and the loader refuses the type."""

    def __init__(self, name, color=(1, 1, 1), ior=1.333):
        Serializable.__init__(self, name, ["color", "ior", "type"])
        self._s["color"] = _color(color)
        self._s["ior"] = float(ior)
        self._s["type"] = "transmitMaterial"


class GlassMaterial(Material):
    """Specular reflection and transmission with constant colours
    (shaders.py:89-106, glassMaterial.cpp:33-56)."""

    def __init__(self, name, specCol=(1, 1, 1), transmitCol=(1, 1, 1), reflectance=1.0, ior=1.333):
        Serializable.__init__(self, name, ["specColor", "transmitColor", "reflectance", "ior", "type"])
        self._s["specColor"] = _color(specCol)
        self._s["transmitColor"] = _color(transmitCol)
        self._s["reflectance"] = float(reflectance)
        self._s["ior"] = float(ior)
        self._s["type"] = "glassMaterial"


class CarMaterial(Material):
    """Serialised as PyAurora does (shaders.py:108-154); the reference's
    carMaterial is commented out and the loader refuses the type."""

    def __init__(self, name, diffcolor=(1, 1, 1), clearcoatreflectance=0.1, clearcoatroughness=0.1,
                 toppaintcolor=(1, 1, 1), basespectopreflectance=0.2, basespectoproughness=0.2,
                 bottompaintcolor=(1, 1, 1), basespecbottomreflectance=0.3, basespecbottomroughness=0.3):
        Serializable.__init__(self, name, ["diffusecolor", "clearcoatreflectance", "clearcoatroughness",
                                           "toppaintcolor", "basespectopreflectance", "basespectoproughness",
                                           "bottompaintcolor", "basespecbottomreflectance",
                                           "basespecbottomroughness", "type"])
        self._s["diffusecolor"] = _color(diffcolor)
        self._s["clearcoatreflectance"] = float(clearcoatreflectance)
        self._s["clearcoatroughness"] = float(clearcoatroughness)
        self._s["toppaintcolor"] = _color(toppaintcolor)
        self._s["basespectopreflectance"] = float(basespectopreflectance)
        self._s["basespectoproughness"] = float(basespectoproughness)
        self._s["bottompaintcolor"] = _color(bottompaintcolor)
        self._s["basespecbottomreflectance"] = float(basespecbottomreflectance)
        self._s["basespecbottomroughness"] = float(basespecbottomroughness)
        self._s["type"] = "carMaterial"


class KelemenMaterial(Material):
    """Kelemen-Szirmay-Kalos diffuse + Cook-Torrance specular, weighted by the
    albedo table of ``reflectance`` (shaders.py:156-179,
    kelemenMaterial.cpp)."""

    def __init__(self, name, diffcolor=(1, 1, 1), speccolor=(1, 1, 1), roughness=50, reflectance=0.1):
        Serializable.__init__(self, name, ["diffusecolor", "speccolor", "roughness", "reflectance", "type"])
        self._s["diffusecolor"] = _color(diffcolor)
        self._s["speccolor"] = _color(speccolor)
        self._s["roughness"] = roughness
        self._s["reflectance"] = reflectance
        self._s["type"] = "kelemenMaterial"
