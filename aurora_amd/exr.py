"""OpenEXR files through the library's codec (``aur_exr_write`` /
``aur_exr_read``): FP16 or FP32 RGBA, NONE / ZIPS / ZIP compression, string
metadata.  Images are (H, W, 4) float32 arrays, row 0 first."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

COMPRESSION = {"none": 0, "zips": 2, "zip": 3}


def _meta(meta):
    meta = meta or {}
    keys = (C.c_char_p * max(len(meta), 1))(*[k.encode() for k in meta])
    vals = (C.c_char_p * max(len(meta), 1))(*[str(v).encode() for v in meta.values()])
    return keys, vals, len(meta)


def write_exr(path: str, rgba, compression: str = "zip", half: bool = True, meta: dict | None = None) -> None:
    img = np.ascontiguousarray(rgba, dtype=np.float32)
    if img.ndim != 3 or img.shape[2] != 4:
        raise ValueError("rgba must be (H, W, 4)")
    keys, vals, n = _meta(meta)
    lib = L.load_library()
    L.check(lib.aur_exr_write(path.encode(), img.shape[1], img.shape[0], L.fptr(img), COMPRESSION[compression],
                              int(half), keys, vals, n))


def read_exr(path: str) -> np.ndarray:
    lib = L.load_library()
    w, h = C.c_int(), C.c_int()
    L.check(lib.aur_exr_read(path.encode(), C.byref(w), C.byref(h), None, 0))
    out = np.zeros((h.value, w.value, 4), np.float32)
    L.check(lib.aur_exr_read(path.encode(), C.byref(w), C.byref(h), L.fptr(out), out.size))
    return out
