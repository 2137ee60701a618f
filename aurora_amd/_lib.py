"""ctypes binding of libaurora_hip.so (the C ABI in include/aurora_hip.h).

The shared library is loaded from this package directory.  There is no
fallback: if the library is missing or a call fails, an error is raised.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libaurora_hip.so")

AUR_MAT_MATTE, AUR_MAT_METAL, AUR_MAT_KELEMEN, AUR_MAT_LIGHT, AUR_MAT_GLASS = 0, 1, 2, 3, 4
AUR_DIST_BECKMANN, AUR_DIST_GGX = 0, 1
AUR_LIGHT_AREA, AUR_LIGHT_ENV = 0, 1
AUR_SLOT_DIFFUSE, AUR_SLOT_SPEC, AUR_SLOT_GAIN, AUR_SLOT_ROUGH, AUR_SLOT_ROUGHB, AUR_SLOT_MIX, AUR_SLOT_NORMAL = range(7)
AUR_N_SLOTS = 7
AUR_SHADER_TEXTURE2D, AUR_SHADER_TRIPLANAR, AUR_SHADER_CONSTANT = 0, 1, 2
AUR_SHADER_COLOR, AUR_SHADER_FLOAT = 0, 1
AUR_SAMPLER_HASH, AUR_SAMPLER_HALTON = 0, 1

ERROR_NAMES = {-1: "AUR_EINVAL", -2: "AUR_EIO", -3: "AUR_EPARSE", -4: "AUR_EHIP", -5: "AUR_ESTATE",
               -6: "AUR_ECAPACITY"}


class AuroraError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__("%s (%d): %s" % (ERROR_NAMES.get(code, "error"), code, msg))
        self.code = code


class Ray(C.Structure):
    _fields_ = [("o", C.c_float * 3), ("tmin", C.c_float), ("d", C.c_float * 3), ("tmax", C.c_float)]


class Hit(C.Structure):
    _fields_ = [("t", C.c_float), ("u", C.c_float), ("v", C.c_float), ("prim", C.c_int)]


class Material(C.Structure):
    _fields_ = [("type", C.c_int), ("distribution", C.c_int), ("diffuse", C.c_float * 3),
                ("spec", C.c_float * 3), ("gain", C.c_float * 3), ("roughness", C.c_float),
                ("roughnessB", C.c_float), ("mix", C.c_float), ("reflectance", C.c_float),
                ("emission", C.c_float * 3), ("shader", C.c_int * AUR_N_SLOTS), ("c2o", C.c_float * 16),
                ("transmit", C.c_float * 3), ("ior", C.c_float)]

    def __init__(self, *args, **kw):
        super().__init__(*args, **kw)
        if "shader" not in kw and len(args) < 11:
            for k in range(AUR_N_SLOTS):
                self.shader[k] = -1
        if "c2o" not in kw and len(args) < 12:
            for k in range(4):
                self.c2o[5 * k] = 1.0


class Light(C.Structure):
    _fields_ = [("type", C.c_int), ("m", C.c_float * 16), ("sx", C.c_float), ("sy", C.c_float),
                ("color", C.c_float * 3), ("exposure", C.c_float), ("envmap", C.c_int)]

    def __init__(self, *args, **kw):
        super().__init__(*args, **kw)
        if "envmap" not in kw and len(args) < 7:
            self.envmap = -1


class Shader(C.Structure):
    _fields_ = [("type", C.c_int), ("kind", C.c_int), ("texture", C.c_int), ("scale", C.c_float),
                ("value", C.c_float * 3)]


class Texture(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("offset", C.c_int64)]


class SceneDesc(C.Structure):
    _fields_ = [("n_verts", C.c_int), ("n_tris", C.c_int), ("verts", C.POINTER(C.c_float)),
                ("tris", C.POINTER(C.c_int)), ("normals", C.POINTER(C.c_float)), ("attr", C.POINTER(C.c_int)),
                ("n_attrs", C.c_int), ("mats", C.POINTER(Material)), ("n_lights", C.c_int),
                ("lights", C.POINTER(Light)), ("fov", C.c_float), ("width", C.c_int), ("height", C.c_int),
                ("pixelsamples", C.c_int), ("lightsamples", C.c_int), ("mindepth", C.c_int),
                ("maxdepth", C.c_int), ("uvs", C.POINTER(C.c_float)), ("n_shaders", C.c_int),
                ("shaders", C.POINTER(Shader)), ("n_textures", C.c_int), ("textures", C.POINTER(Texture)),
                ("n_texels", C.c_int64), ("texels", C.POINTER(C.c_float)), ("cam_to_world", C.c_float * 16),
                ("world_to_cam", C.c_float * 16)]


class SceneInfo(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("n_verts", "n_tris", "n_attrs", "n_lights", "n_objects", "width", "height",
                                       "pixelsamples", "lightsamples", "mindepth", "maxdepth")] + \
               [("fov", C.c_float)] + \
               [(n, C.c_int) for n in ("bvh_width", "bvh_nodes", "bvh_depth", "bvh_max_leaf", "n_shaders",
                                       "n_textures")] + \
               [("n_texels", C.c_int64)]


class AccelOpts(C.Structure):
    _fields_ = [("width", C.c_int), ("max_leaf", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("paths", C.c_int64), ("ext_rays", C.c_int64), ("hit_rays", C.c_int64),
                ("shadow_rays", C.c_int64), ("ms", C.c_double)]

    def as_dict(self):
        return {"paths": self.paths, "ext_rays": self.ext_rays, "hit_rays": self.hit_rays,
                "shadow_rays": self.shadow_rays, "ms": self.ms}


class ParamGrads(C.Structure):
    _fields_ = [("albedo_tri", C.POINTER(C.c_float)), ("roughness", C.POINTER(C.c_float)),
                ("spec_weight", C.POINTER(C.c_float)), ("albedo_texture", C.POINTER(C.c_float))]


class KernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_int64), ("total_ms", C.c_double)]


class BsdfQuery(C.Structure):
    _fields_ = [("kind", C.c_int), ("distribution", C.c_int), ("color", C.c_float * 3), ("roughness", C.c_float),
                ("V", C.c_float * 3), ("N", C.c_float * 3), ("L", C.c_float * 3), ("u1", C.c_float),
                ("u2", C.c_float)]


class BsdfResult(C.Structure):
    _fields_ = [("f", C.c_float * 3), ("pdf", C.c_float), ("sample_dir", C.c_float * 3), ("sample_pdf", C.c_float),
                ("sample_color", C.c_float * 3)]


class LightQuery(C.Structure):
    _fields_ = [("light", C.c_int), ("P", C.c_float * 3), ("N", C.c_float * 3), ("u1", C.c_float),
                ("u2", C.c_float), ("D", C.c_float * 3), ("tmax", C.c_float)]


class LightResult(C.Structure):
    _fields_ = [("sample_dir", C.c_float * 3), ("sample_tmax", C.c_float), ("sample_pdf", C.c_float),
                ("sample_color", C.c_float * 3), ("Le", C.c_float * 3), ("pdf", C.c_float), ("pdf_tmax", C.c_float),
                ("visible", C.c_int)]


class Job(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("first_pass", C.c_int), ("n_passes", C.c_int)]


class JobResult(C.Structure):
    _fields_ = [("paths", C.c_int64), ("ext_rays", C.c_int64), ("hit_rays", C.c_int64), ("shadow_rays", C.c_int64),
                ("latency_ms", C.c_double), ("order", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchOpts(C.Structure):
    _fields_ = [("order", C.c_int), ("use_graph", C.c_int), ("iters_per_poll", C.c_int)]


class LaunchOpts(C.Structure):
    _fields_ = [("trace_block", C.c_int), ("trace_waves_per_cu", C.c_int), ("pool_paths", C.c_int),
                ("shade_list_tiles", C.c_int)]


_P = C.c_void_p
_FP = C.POINTER(C.c_float)
_IP = C.POINTER(C.c_int)

# name -> (restype, argtypes)
SIGNATURES = {
    "aur_create": (C.c_int, [C.POINTER(_P), C.c_int]),
    "aur_destroy": (None, [_P]),
    "aur_last_error": (C.c_char_p, []),
    "aur_set_stream": (C.c_int, [_P, _P]),
    "aur_set_seed": (C.c_int, [_P, C.c_uint32]),
    "aur_set_sampler": (C.c_int, [_P, C.c_int]),
    "aur_scene_load": (C.c_int, [_P, C.c_char_p, C.c_char_p]),
    "aur_scene_upload": (C.c_int, [_P, C.POINTER(SceneDesc)]),
    "aur_set_tri_albedo": (C.c_int, [_P, _FP]),
    "aur_display": (C.c_int, [_P, C.POINTER(C.c_uint16)]),
    "aur_set_albedo_texture": (C.c_int, [_P, C.c_int, C.c_int, _FP, _IP, C.c_int]),
    "aur_build_accel": (C.c_int, [_P, C.POINTER(AccelOpts)]),
    "aur_scene_info_get": (C.c_int, [_P, C.POINTER(SceneInfo)]),
    "aur_bvh_export": (C.c_int, [_P, C.c_void_p, _FP]),
    "aur_scene_export": (C.c_int, [_P, _FP, _IP, _FP, _IP, C.POINTER(Material), C.POINTER(Light)]),
    "aur_scene_export_shading": (C.c_int, [_P, _FP, C.POINTER(Shader), C.POINTER(Texture), _FP, _FP]),
    "aur_render": (C.c_int, [_P, C.c_int, C.c_int, _FP, C.POINTER(Stats)]),
    "aur_render_paths": (C.c_int, [_P, _IP, _IP, C.c_int, _FP, _FP, C.POINTER(Stats)]),
    "aur_framebuffer": (C.c_int, [_P, _FP]),
    "aur_write_image": (C.c_int, [_P, C.c_char_p, C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_char_p), C.c_int]),
    "aur_render_image": (C.c_int, [_P, C.c_char_p, C.c_double, C.POINTER(Stats), C.POINTER(C.c_int)]),
    "aur_exr_write": (C.c_int, [C.c_char_p, C.c_int, C.c_int, _FP, C.c_int, C.c_int, C.POINTER(C.c_char_p),
                                C.POINTER(C.c_char_p), C.c_int]),
    "aur_exr_read": (C.c_int, [C.c_char_p, _IP, _IP, _FP, C.c_int64]),
    "aur_batch_submit": (C.c_int, [_P, C.POINTER(Job), C.c_int, C.POINTER(BatchOpts), C.POINTER(JobResult),
                                   C.POINTER(Stats)]),
    "aur_set_launch": (C.c_int, [_P, C.POINTER(LaunchOpts)]),
    "aur_get_launch": (C.c_int, [_P, C.POINTER(LaunchOpts)]),
    "aur_set_material_params": (C.c_int, [_P, _FP, _FP]),
    "aur_adjoint": (C.c_int, [_P, C.c_int, C.c_int, _FP, C.POINTER(ParamGrads), C.POINTER(Stats)]),
    "aur_keep_path_records": (C.c_int, [_P, C.c_int, C.c_int]),
    "aur_set_profiling": (C.c_int, [_P, C.c_int]),
    "aur_kernel_stats": (C.c_int, [_P, C.POINTER(KernelStat), C.c_int, _IP]),
    "aur_trace": (C.c_int, [_P, C.POINTER(Ray), C.c_int, C.POINTER(Hit), C.c_int]),
    "aur_trace_noalloc": (C.c_int, [_P, C.c_void_p, C.c_int, C.c_void_p, C.c_int]),
    "aur_bsdf": (C.c_int, [_P, C.POINTER(BsdfQuery), C.c_int, C.POINTER(BsdfResult)]),
    "aur_light_queries": (C.c_int, [_P, C.POINTER(LightQuery), C.c_int, C.POINTER(LightResult)]),
    "aur_kelemen_table": (C.c_int, [_P, C.c_float, C.c_int, C.c_int, C.c_uint32, _FP]),
    "aur_loop_subdivide": (C.c_int, [_P, C.c_int, C.c_int, _FP, _IP, C.c_int, C.c_float, C.c_float, C.c_int,
                                     C.c_uint32, _FP, _IP, _FP, C.c_int, C.c_int, _IP, _IP]),
}

_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Loads libaurora_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError("libaurora_hip.so not found at %s; run `python -m aurora_amd.build`" % path)
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(code: int) -> None:
    if code != 0:
        msg = load_library().aur_last_error()
        raise AuroraError(code, msg.decode() if msg else "")


def fptr(a: np.ndarray):
    return a.ctypes.data_as(_FP)


def iptr(a: np.ndarray):
    return a.ctypes.data_as(_IP)
