"""Scene authoring in Python 3, the counterpart of PyAurora.

    from aurora_amd.pyaurora import core, geometry, lights, shaders

mirrors PyAurora's ``import core, geometry, lights, shaders``: the same
classes, constructor arguments and JSON scene shape, with ``core.render``
rendering through libaurora_hip.so.
"""
from . import core, geometry, lights, shaders  # noqa: F401

__all__ = ["core", "geometry", "lights", "shaders"]
