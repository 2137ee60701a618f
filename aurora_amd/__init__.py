"""aurora_amd -- MI355X-native path-tracing hot path for Aurora scenes.

The renderer itself is libaurora_hip.so (HIP kernels for gfx950 behind the C
ABI in include/aurora_hip.h); this package holds its Python binding, the
PyAurora-style scene authoring API and the deterministic scene generators.
"""
from .renderer import Renderer  # noqa: F401

__all__ = ["Renderer"]
