"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, never by the product.  ``build()`` compiles the
C restatement with the Makefile next to this file.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

# samples per albedo-table entry: SPECALBEDO_SAMPLES (constants.h:26)
KELEMEN_SAMPLES = 250000
KELEMEN_SEED = 0x4B454C


def build() -> str:
    r = subprocess.run(["make", "-s", "-C", HERE], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout)
    return LIB_PATH


class OMaterial(C.Structure):
    _fields_ = [("type", C.c_int), ("distribution", C.c_int), ("diffuse", C.c_float * 3), ("spec", C.c_float * 3),
                ("gain", C.c_float * 3), ("roughness", C.c_float), ("roughnessB", C.c_float), ("mix", C.c_float),
                ("reflectance", C.c_float), ("emission", C.c_float * 3), ("table", C.POINTER(C.c_float)),
                ("shader", C.c_int * 7), ("c2o", C.c_float * 16), ("transmit", C.c_float * 3), ("ior", C.c_float)]


class OLight(C.Structure):
    _fields_ = [("type", C.c_int), ("m", C.c_float * 16), ("sx", C.c_float), ("sy", C.c_float),
                ("color", C.c_float * 3), ("exposure", C.c_float), ("intensity", C.c_float), ("N", C.c_float * 3),
                ("pcam", (C.c_float * 3) * 4), ("tri", (C.c_float * 9) * 2), ("envmap", C.c_int),
                ("c2w", C.c_float * 16), ("w2c", C.c_float * 16)]


class OShader(C.Structure):
    _fields_ = [("type", C.c_int), ("kind", C.c_int), ("texture", C.c_int), ("scale", C.c_float),
                ("value", C.c_float * 3)]


class OTexture(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("rgba", C.POINTER(C.c_float)), ("pdf_norm", C.c_float)]


class OScene(C.Structure):
    _fields_ = [("n_verts", C.c_int), ("n_tris", C.c_int), ("verts", C.POINTER(C.c_float)),
                ("tris", C.POINTER(C.c_int)), ("normals", C.POINTER(C.c_float)), ("attr", C.POINTER(C.c_int)),
                ("tri_albedo", C.POINTER(C.c_float)), ("n_attrs", C.c_int), ("mats", C.POINTER(OMaterial)),
                ("n_lights", C.c_int), ("lights", C.POINTER(OLight)), ("fov", C.c_float), ("width", C.c_int),
                ("height", C.c_int), ("maxdepth", C.c_int), ("mindepth", C.c_int), ("uvs", C.POINTER(C.c_float)),
                ("n_shaders", C.c_int), ("shaders", C.POINTER(OShader)), ("n_textures", C.c_int),
                ("textures", C.POINTER(OTexture)), ("sampler", C.c_int), ("lightsamples", C.c_int),
                ("ptex", C.POINTER(C.c_float)), ("ptex_w", C.c_int), ("ptex_h", C.c_int),
                ("ptex_mat", C.POINTER(C.c_int))]


class ORay(C.Structure):
    _fields_ = [("o", C.c_float * 3), ("d", C.c_float * 3), ("tmin", C.c_float), ("tmax", C.c_float)]


class OHit(C.Structure):
    _fields_ = [("t", C.c_float), ("u", C.c_float), ("v", C.c_float), ("prim", C.c_int)]


class OCounters(C.Structure):
    _fields_ = [("paths", C.c_int64), ("ext", C.c_int64), ("hit", C.c_int64), ("shadow", C.c_int64)]

    def as_dict(self):
        return {"paths": self.paths, "ext_rays": self.ext, "hit_rays": self.hit, "shadow_rays": self.shadow}


_FP = C.POINTER(C.c_float)
_IP = C.POINTER(C.c_int)
_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    sig = {
        "orc_intersect_tri": (C.c_int, [_FP, _FP, _FP, C.POINTER(ORay), _FP, _FP, _FP]),
        "orc_trace_brute": (None, [C.POINTER(OScene), C.POINTER(ORay), C.c_int, C.c_int, C.POINTER(OHit)]),
        "orc_bvh_build": (C.c_void_p, [C.POINTER(OScene), C.c_int]),
        "orc_bvh_free": (None, [C.c_void_p]),
        "orc_trace_bvh": (None, [C.POINTER(OScene), C.c_void_p, C.POINTER(ORay), C.c_int, C.c_int, C.POINTER(OHit)]),
        "orc_light_init": (None, [C.POINTER(OLight)]),
        "orc_texture_init_pdf": (None, [C.POINTER(OTexture)]),
        "orc_halton_sample": (C.c_float, [C.c_uint, C.c_uint]),
        "orc_halton_enum": (None, [C.c_uint, C.c_uint] + [C.POINTER(C.c_uint)] * 5),
        "orc_halton_index": (C.c_uint, [C.c_uint] * 5),
        "orc_halton_slot": (C.c_float, [C.c_uint] * 6),
        "orc_half_round": (C.c_float, [C.c_float]),
        "orc_texture_read": (None, [C.POINTER(OTexture), C.c_float, C.c_float, _FP]),
        "orc_shader_eval": (None, [C.POINTER(OScene), C.c_int, _FP, _FP, C.c_float, C.c_float, C.c_int, _FP]),
        "orc_env_eval": (None, [C.POINTER(OScene), C.POINTER(OLight), _FP, _FP]),
        "orc_bsdf_eval": (None, [C.c_int, _FP, _FP, _FP, _FP, _FP, _FP]),
        "orc_bsdf_sample": (None, [C.c_int, _FP, _FP, _FP, C.c_float, C.c_float, _FP, _FP, _FP]),
        "orc_kelemen_table": (None, [C.c_float, C.c_int, C.c_int, C.c_uint32, _FP]),
        "orc_kelemen_albedo": (C.c_float, [_FP, C.c_float, C.c_float]),
        "orc_path_key": (C.c_uint32, [C.c_uint32, C.c_uint32, C.c_uint32]),
        "orc_slot_uniform": (C.c_float, [C.c_uint32, C.c_uint32]),
        "orc_path": (None, [C.POINTER(OScene), C.c_void_p, C.c_int, C.c_int, C.c_uint32, C.c_uint32, _FP, _FP,
                            C.POINTER(OCounters)]),
        "orc_render": (None, [C.POINTER(OScene), C.c_void_p, C.c_uint32, C.c_int, C.c_int, _FP, C.POINTER(OCounters),
                              C.c_int]),
        "orc_render_paths": (None, [C.POINTER(OScene), C.c_void_p, C.c_uint32, _IP, _IP, C.c_int, _FP, _FP,
                                    C.POINTER(OCounters), C.c_int]),
        "orc_loop_subdivide": (C.c_int, [C.c_int, C.c_int, _FP, _IP, C.c_int, _FP, _IP, _FP, C.c_int, C.c_int, _IP,
                                         _IP, _FP, _FP, C.c_int]),
        "orc_displace_radial": (None, [C.c_int, _FP, C.c_float, C.c_float, C.c_int, C.c_uint32]),
        "orc_vertex_normals": (None, [C.c_int, C.c_int, _FP, _IP, _FP]),
        "orc_camera_ray": (None, [C.POINTER(OScene), C.c_int, C.c_int, C.c_float, C.c_float, C.POINTER(ORay)]),
        "orc_light_sample": (None, [C.POINTER(OLight), _FP, _FP, C.c_float, C.c_float, C.POINTER(ORay), _FP, _FP]),
        "orc_light_pdf": (C.c_float, [C.POINTER(OLight), C.POINTER(ORay)]),
        "orc_light_sample_in": (None, [C.POINTER(OScene), C.POINTER(OLight), _FP, _FP, C.c_float, C.c_float,
                                       C.POINTER(ORay), _FP, _FP]),
        "orc_light_pdf_in": (C.c_float, [C.POINTER(OScene), C.POINTER(OLight), C.POINTER(ORay)]),
        "orc_light_visible": (C.c_int, [C.POINTER(OLight), _FP, _FP]),
        "orc_adjoint": (None, [C.POINTER(OScene), C.c_void_p, C.c_uint32, C.c_int, C.c_int, _FP, _FP, _FP, _FP, _FP, _FP,
                               C.c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def fp(a):
    return a.ctypes.data_as(_FP)


def ip(a):
    return a.ctypes.data_as(_IP)


_table_cache = {}


def kelemen_table(reflectance: float, distribution: int, n_samples: int = KELEMEN_SAMPLES,
                  seed: int = KELEMEN_SEED) -> np.ndarray:
    key = (float(np.float32(reflectance)), distribution, n_samples, seed)
    if key not in _table_cache:
        out = np.zeros(64 * 32, np.float32)
        lib().orc_kelemen_table(reflectance, distribution, n_samples, seed, fp(out))
        _table_cache[key] = out
    return _table_cache[key]


def scene_kelemen_table(reflectance: float, distribution: int) -> np.ndarray:
    """The table a scene's Kelemen material reads: the estimate at
    SPECALBEDO_SAMPLES per entry, written through the FP16 display and read
    back as Imf::HALF (kelemenMaterial.cpp:53-76, 170-177): each entry
    rounded to FP16, to nearest even."""
    key = ("half", float(np.float32(reflectance)), distribution)
    if key not in _table_cache:
        _table_cache[key] = kelemen_table(reflectance, distribution).astype(np.float16).astype(np.float32)
    return _table_cache[key]


class Scene:
    """An oracle scene built from the renderer's flattened export."""

    def __init__(self, exp: dict, tri_albedo=None, bvh: bool = False, leaf: int = 4, ptex=None, ptex_attrs=()):
        """ptex: an (h, w, 3) FP32 albedo texture (row 0 at the top) replacing
        the diffuse colour of the attributes in ptex_attrs, read at the hit's
        uv as aur_set_albedo_texture's texture is."""
        self.exp = exp
        self.verts = np.ascontiguousarray(exp["verts"], np.float32).reshape(-1)
        self.tris = np.ascontiguousarray(exp["tris"], np.int32).reshape(-1)
        self.normals = np.ascontiguousarray(exp["normals"], np.float32).reshape(-1)
        self.attr = np.ascontiguousarray(exp["attr"], np.int32).reshape(-1)
        self.albedo = None if tri_albedo is None else np.ascontiguousarray(tri_albedo, np.float32).reshape(-1)
        mats = exp["materials"]
        self.mats = (OMaterial * max(len(mats), 1))()
        self.tables = []
        for k, m in enumerate(mats):
            o = self.mats[k]
            o.type, o.distribution = m.type, m.distribution
            for j in range(3):
                o.diffuse[j], o.spec[j], o.gain[j], o.emission[j] = m.diffuse[j], m.spec[j], m.gain[j], m.emission[j]
            o.roughness, o.roughnessB, o.mix, o.reflectance = m.roughness, m.roughnessB, m.mix, m.reflectance
            for j in range(7):
                o.shader[j] = getattr(m, "shader", [-1] * 7)[j]
            for j in range(16):
                o.c2o[j] = getattr(m, "c2o", np.eye(4, dtype=np.float32).reshape(-1))[j]
            for j in range(3):
                o.transmit[j] = getattr(m, "transmit", [0.0] * 3)[j]
            o.ior = getattr(m, "ior", 1.0)
            if m.type == 2:
                # roundf(): halves away from zero, like the reference (kelemenMaterial.cpp:50)
                q = min(max(int(np.floor(float(np.float32(m.reflectance)) * 20.0 + 0.5)), 0), 20)
                t = scene_kelemen_table(np.float32(q / 20.0), m.distribution)
                self.tables.append(t)
                o.table = fp(t)
        cam = np.asarray(exp.get("camera", np.concatenate([np.eye(4), np.eye(4)]).reshape(-1)), np.float32)
        # textures: private copies (initPdf rewrites alpha); uvs; shaders
        self.tex_rgba = []
        texs = exp.get("textures") or []
        self.textures = (OTexture * max(len(texs), 1))()
        texels = np.asarray(exp.get("texels", np.zeros(0)), np.float32).reshape(-1)
        for k, t in enumerate(texs):
            n = t.width * t.height
            rgba = np.ascontiguousarray(texels[4 * t.offset:4 * (t.offset + n)].copy())
            self.tex_rgba.append(rgba)
            self.textures[k].width, self.textures[k].height = t.width, t.height
            self.textures[k].rgba = fp(rgba)
            self.textures[k].pdf_norm = 1.0
        shs = exp.get("shaders") or []
        self.shaders = (OShader * max(len(shs), 1))()
        for k, sh in enumerate(shs):
            o = self.shaders[k]
            o.type, o.kind, o.texture, o.scale = sh.type, sh.kind, sh.texture, sh.scale
            for j in range(3):
                o.value[j] = sh.value[j]
        self.uvs = None
        if exp.get("uvs") is not None:
            self.uvs = np.ascontiguousarray(exp["uvs"], np.float32).reshape(-1)
        lights = exp["lights"]
        self.lights = (OLight * max(len(lights), 1))()
        for k, l in enumerate(lights):
            o = self.lights[k]
            o.type = l.type
            for j in range(16):
                o.m[j] = l.m[j]
            o.sx, o.sy, o.exposure = l.sx, l.sy, l.exposure
            for j in range(3):
                o.color[j] = l.color[j]
            o.envmap = getattr(l, "envmap", -1) if l.type == 1 else -1
            for j in range(16):
                o.c2w[j], o.w2c[j] = cam[j], cam[16 + j]
            lib().orc_light_init(C.byref(o))
            if o.envmap >= 0:
                lib().orc_texture_init_pdf(C.byref(self.textures[o.envmap]))
        s = OScene()
        s.n_verts = self.verts.size // 3
        s.n_tris = self.tris.size // 3
        s.verts, s.tris, s.normals, s.attr = fp(self.verts), ip(self.tris), fp(self.normals), ip(self.attr)
        s.tri_albedo = fp(self.albedo) if self.albedo is not None else None
        s.n_attrs = len(mats)
        s.mats = C.cast(self.mats, C.POINTER(OMaterial))
        s.n_lights = len(lights)
        s.lights = C.cast(self.lights, C.POINTER(OLight))
        s.fov = exp["fov"]
        s.width, s.height = exp["width"], exp["height"]
        s.maxdepth, s.mindepth = exp["maxdepth"], exp["mindepth"]
        s.uvs = fp(self.uvs) if self.uvs is not None else None
        s.n_shaders, s.shaders = len(shs), C.cast(self.shaders, C.POINTER(OShader))
        s.n_textures, s.textures = len(texs), C.cast(self.textures, C.POINTER(OTexture))
        s.sampler = 1 if exp.get("sampler") == "halton" else 0
        s.lightsamples = int(exp.get("lightsamples", 1))
        self.ptex = None
        if ptex is not None:
            self.ptex = np.ascontiguousarray(ptex, np.float32)
            assert self.ptex.ndim == 3 and self.ptex.shape[2] == 3
            self.ptex_mat = np.zeros(max(len(mats), 1), np.int32)
            self.ptex_mat[list(ptex_attrs)] = 1
            s.ptex, s.ptex_h, s.ptex_w = fp(self.ptex), self.ptex.shape[0], self.ptex.shape[1]
            s.ptex_mat = ip(self.ptex_mat)
        self.s = s
        self.bvh = lib().orc_bvh_build(C.byref(s), leaf) if bvh else None

    def __del__(self):
        try:
            if self.bvh:
                lib().orc_bvh_free(self.bvh)
        except Exception:
            pass

    def camera_rays(self, pixel, j0: float = 0.5, j1: float = 0.5) -> np.ndarray:
        """Camera rays (camera.cpp:34-62) as (n, 8) rows like aur_ray."""
        out = np.zeros((len(pixel), 8), np.float32)
        r = ORay()
        W = self.exp["width"]
        for k, p in enumerate(pixel):
            lib().orc_camera_ray(C.byref(self.s), int(p) % W, int(p) // W, j0, j1, C.byref(r))
            out[k] = [r.o[0], r.o[1], r.o[2], r.tmin, r.d[0], r.d[1], r.d[2], r.tmax]
        return out

    # -- textures, shaders, env maps (single routines) --------------------
    def texture_read(self, tex: int, s: float, t: float) -> np.ndarray:
        out = np.zeros(3, np.float32)
        lib().orc_texture_read(C.byref(self.textures[tex]), s, t, fp(out))
        return out

    def shader_eval(self, shader: int, P, Ng, s: float, t: float, attr: int) -> np.ndarray:
        out = np.zeros(3, np.float32)
        lib().orc_shader_eval(C.byref(self.s), shader, fp(np.asarray(P, np.float32)), fp(np.asarray(Ng, np.float32)),
                              s, t, attr, fp(out))
        return out

    def env_eval(self, light: int, d) -> np.ndarray:
        out = np.zeros(3, np.float32)
        lib().orc_env_eval(C.byref(self.s), C.byref(self.lights[light]), fp(np.asarray(d, np.float32)), fp(out))
        return out

    def trace(self, rays: np.ndarray, any_hit: bool = False, use_bvh: bool = False):
        """rays: (n, 8) rows (ox, oy, oz, tmin, dx, dy, dz, tmax) like aur_ray."""
        rays = np.asarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        orays = (ORay * max(n, 1))()
        for i in range(n):
            r = rays[i]
            orays[i].o[0], orays[i].o[1], orays[i].o[2] = r[0], r[1], r[2]
            orays[i].d[0], orays[i].d[1], orays[i].d[2] = r[4], r[5], r[6]
            orays[i].tmin, orays[i].tmax = r[3], r[7]
        hits = (OHit * max(n, 1))()
        if use_bvh:
            lib().orc_trace_bvh(C.byref(self.s), self.bvh, orays, n, int(any_hit), hits)
        else:
            lib().orc_trace_brute(C.byref(self.s), orays, n, int(any_hit), hits)
        out = np.zeros(n, dtype=[("t", np.float32), ("u", np.float32), ("v", np.float32), ("prim", np.int32)])
        for i in range(n):
            out[i] = (hits[i].t, hits[i].u, hits[i].v, hits[i].prim)
        return out

    def render(self, seed: int, first_pass: int, n_passes: int, threads: int = 0):
        acc = np.zeros((self.exp["height"], self.exp["width"], 4), np.float32)
        cnt = OCounters()
        lib().orc_render(C.byref(self.s), self.bvh, seed, first_pass, n_passes, fp(acc), C.byref(cnt), threads)
        return acc, cnt.as_dict()

    def set_material_params(self, roughness=None, spec_weight=None):
        """Mirrors aur_set_material_params (values clamped to >= 1e-6)."""
        for k in range(len(self.exp["materials"])):
            o = self.mats[k]
            if roughness is not None and o.type in (1, 2):
                o.roughness = max(float(roughness[k]), 1e-6)
                if o.type == 1:
                    o.roughnessB = o.roughness
            if spec_weight is not None and o.type == 2:
                w = max(float(spec_weight[k]), 1e-6)
                o.spec[0] = o.spec[1] = o.spec[2] = w

    def adjoint(self, seed: int, first_pass: int, n_passes: int, dldi, threads: int = 0):
        """CPU reference gradient: (albedo grads (n_tris, 3) or None, roughness grads, spec grads), and the
        albedo texture's gradient (h, w, 3) as a fourth item when the scene has one."""
        dl = np.ascontiguousarray(dldi, np.float32).reshape(-1)
        nt = self.tris.size // 3
        nm = len(self.exp["materials"])
        ga = np.zeros(3 * nt, np.float32)
        gr = np.zeros(nm, np.float32)
        gs = np.zeros(nm, np.float32)
        gt = np.zeros(self.ptex.shape, np.float32) if self.ptex is not None else None
        lib().orc_adjoint(C.byref(self.s), self.bvh, seed, first_pass, n_passes, fp(dl),
                          fp(self.albedo) if self.albedo is not None else None, fp(ga), fp(gr), fp(gs),
                          fp(gt) if gt is not None else None, threads)
        out = ((ga.reshape(-1, 3) if self.albedo is not None else None), gr, gs)
        return out + (gt,) if gt is not None else out

    def render_paths(self, seed: int, pixel, sample, slots=None, threads: int = 0):
        pixel = np.ascontiguousarray(pixel, np.int32)
        sample = np.ascontiguousarray(sample, np.int32)
        n = pixel.size
        out = np.zeros((n, 4), np.float32)
        sl = None if slots is None else np.ascontiguousarray(slots, np.float32)
        cnt = OCounters()
        lib().orc_render_paths(C.byref(self.s), self.bvh, seed, ip(pixel), ip(sample), n,
                               fp(sl) if sl is not None else None, fp(out), C.byref(cnt), threads)
        return out, cnt.as_dict()


def bsdf_eval(kind, color, rough, dist, V, N, L):
    p = np.array(list(color) + [rough, dist], np.float32)
    f = np.zeros(3, np.float32)
    pdf = np.zeros(1, np.float32)
    lib().orc_bsdf_eval(kind, fp(p), fp(np.asarray(V, np.float32)), fp(np.asarray(N, np.float32)),
                        fp(np.asarray(L, np.float32)), fp(f), fp(pdf))
    return f, float(pdf[0])


def bsdf_sample(kind, color, rough, dist, V, N, u1, u2):
    p = np.array(list(color) + [rough, dist], np.float32)
    d = np.zeros(3, np.float32)
    pdf = np.zeros(1, np.float32)
    col = np.zeros(3, np.float32)
    lib().orc_bsdf_sample(kind, fp(p), fp(np.asarray(V, np.float32)), fp(np.asarray(N, np.float32)), u1, u2, fp(d),
                          fp(pdf), fp(col))
    return d, float(pdf[0]), col


def loop_subdivide(P, F, levels: int, amplitude: float = 0.0, frequency: float = 1.0, octaves: int = 0,
                   seed: int = 0, uv=None, limit: bool = True):
    """Oracle Loop subdivision; with amplitude != 0 the radial displacement and
    area-weighted normals follow.  Returns (P, F, N), and the output faces'
    corner uvs (nf_out, 3, 2) as a fourth item when uv (nf, 3, 2 corner uvs
    of the input faces) is given."""
    P = np.ascontiguousarray(P, np.float32).reshape(-1)
    F = np.ascontiguousarray(F, np.int32).reshape(-1)
    nv, nf = P.size // 3, F.size // 3
    cap_f = nf * 4 ** levels
    cap_v = nv + 3 * cap_f
    Po = np.zeros(3 * cap_v, np.float32)
    Fo = np.zeros(3 * cap_f, np.int32)
    No = np.zeros(3 * cap_v, np.float32)
    onv, onf = C.c_int(), C.c_int()
    uvi = None if uv is None else np.ascontiguousarray(uv, np.float32).reshape(-1)
    if uvi is not None and uvi.size != 6 * nf:
        raise ValueError("uv needs 6 values per input face")
    uvo = np.zeros(6 * cap_f, np.float32) if uvi is not None else None
    rc = lib().orc_loop_subdivide(nv, nf, fp(P), ip(F), levels, fp(Po), ip(Fo), fp(No), cap_v, cap_f, C.byref(onv),
                                  C.byref(onf), fp(uvi) if uvi is not None else None,
                                  fp(uvo) if uvo is not None else None, int(limit))
    if rc != 0:
        raise RuntimeError("oracle loop subdivision failed: %d" % rc)
    Po, Fo, No = Po[:3 * onv.value].copy(), Fo[:3 * onf.value].copy(), No[:3 * onv.value].copy()
    if amplitude != 0.0:
        lib().orc_displace_radial(onv.value, fp(Po), amplitude, frequency, octaves, seed)
        lib().orc_vertex_normals(onv.value, onf.value, fp(Po), ip(Fo), fp(No))
    out = (Po.reshape(-1, 3), Fo.reshape(-1, 3), No.reshape(-1, 3))
    return out + (uvo[:6 * onf.value].reshape(-1, 3, 2),) if uvo is not None else out


def make_light(m16, sx, sy, color, exposure=0.0, kind=0) -> OLight:
    o = OLight()
    o.type = kind
    for j in range(16):
        o.m[j] = m16[j]
    o.sx, o.sy, o.exposure = sx, sy, exposure
    for j in range(3):
        o.color[j] = color[j]
    lib().orc_light_init(C.byref(o))
    return o


def halton_sample(dim: int, index: int) -> float:
    return float(lib().orc_halton_sample(dim, index & 0xFFFFFFFF))


def halton_enum(w: int, h: int) -> dict:
    v = [C.c_uint() for _ in range(5)]
    lib().orc_halton_enum(w, h, *[C.byref(x) for x in v])
    return dict(zip(("p2", "p3", "mx", "my", "inc"), [x.value for x in v]))


def halton_index(w: int, h: int, i: int, x: int, y: int) -> int:
    return int(lib().orc_halton_index(w, h, i, x, y))


def halton_slot(w: int, h: int, i: int, x: int, y: int, slot: int) -> float:
    return float(lib().orc_halton_slot(w, h, i, x, y, slot))


def half_round(x: float) -> float:
    return float(lib().orc_half_round(float(x)))


def light_sample(light: OLight, P, N, x, y, scene=None):
    """generateSample; pass the oracle Scene for lights that read its textures."""
    ray = ORay()
    pdf = np.zeros(1, np.float32)
    col = np.zeros(3, np.float32)
    P, N = fp(np.asarray(P, np.float32)), fp(np.asarray(N, np.float32))
    if scene is None:
        lib().orc_light_sample(C.byref(light), P, N, x, y, C.byref(ray), fp(pdf), fp(col))
    else:
        lib().orc_light_sample_in(C.byref(scene.s), C.byref(light), P, N, x, y, C.byref(ray), fp(pdf), fp(col))
    return np.array(ray.d[:], np.float32), float(ray.tmax), float(pdf[0]), col


def light_pdf(light: OLight, o, d, tmin=0.0005, tmax=1e7, scene=None):
    ray = ORay()
    for j in range(3):
        ray.o[j], ray.d[j] = o[j], d[j]
    ray.tmin, ray.tmax = tmin, tmax
    if scene is None:
        p = lib().orc_light_pdf(C.byref(light), C.byref(ray))
    else:
        p = lib().orc_light_pdf_in(C.byref(scene.s), C.byref(light), C.byref(ray))
    return float(p), float(ray.tmax)


def intersect_tri(p, e1, e2, o, d, tmin, tmax):
    ray = ORay()
    for j in range(3):
        ray.o[j], ray.d[j] = o[j], d[j]
    ray.tmin, ray.tmax = tmin, tmax
    t, b1, b2 = np.zeros(1, np.float32), np.zeros(1, np.float32), np.zeros(1, np.float32)
    hit = lib().orc_intersect_tri(fp(np.asarray(p, np.float32)), fp(np.asarray(e1, np.float32)),
                                  fp(np.asarray(e2, np.float32)), C.byref(ray), fp(t), fp(b1), fp(b2))
    return bool(hit), float(t[0]), float(b1[0]), float(b2[0])
