"""Host-side mirror of Aurora's ``Renderer`` (src/frontend/renderer.cpp) over the C ABI.

``Renderer(scene_file).render()`` follows the reference's sequence
(renderer.cpp:39-63): parse the scene description, build the render
environment (acceleration structure), render ``pixelsamples`` passes and
return the image.  The per-pass running mean of ``OpenexrDisplay::appendValue``
(openexrDisplay.cpp:61-72) is kept as an FP32 sum divided by the pass count.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib as L


class Renderer:
    """One GPU rendering context (``aur_ctx``)."""

    def __init__(self, scene_file: Optional[str] = None, config: Optional[str] = None, device: int = 0):
        self.lib = L.load_library()
        self._ctx = C.c_void_p()
        L.check(self.lib.aur_create(C.byref(self._ctx), device))
        self.passes_done = 0
        if scene_file is not None:
            self.load(scene_file, config)

    # -- lifetime ------------------------------------------------------------
    def close(self) -> None:
        if self._ctx:
            self.lib.aur_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- scene ---------------------------------------------------------------
    def load(self, scene_file: str, config: Optional[str] = None) -> None:
        """JsonParser::parseScene + ObjTriangleMesh (jsonParser.cpp:30-59)."""
        L.check(self.lib.aur_scene_load(self._ctx, scene_file.encode(), config.encode() if config else None))
        self.passes_done = 0

    def upload(self, scene: dict) -> None:
        """Uploads an already-flattened camera-space scene (see ``export``)."""
        self._keep = keep = {}
        verts = keep["verts"] = np.ascontiguousarray(scene["verts"], dtype=np.float32).reshape(-1)
        tris = keep["tris"] = np.ascontiguousarray(scene["tris"], dtype=np.int32).reshape(-1)
        normals = keep["normals"] = np.ascontiguousarray(scene["normals"], dtype=np.float32).reshape(-1)
        attr = keep["attr"] = np.ascontiguousarray(scene["attr"], dtype=np.int32).reshape(-1)
        mats = keep["mats"] = (L.Material * max(len(scene["materials"]), 1))(*scene["materials"])
        lights = keep["lights"] = (L.Light * max(len(scene["lights"]), 1))(*scene["lights"])
        d = L.SceneDesc()
        d.n_verts = verts.size // 3
        d.n_tris = tris.size // 3
        d.verts, d.tris, d.normals, d.attr = L.fptr(verts), L.iptr(tris), L.fptr(normals), L.iptr(attr)
        d.n_attrs = len(scene["materials"])
        d.mats = C.cast(mats, C.POINTER(L.Material))
        d.n_lights = len(scene["lights"])
        d.lights = C.cast(lights, C.POINTER(L.Light))
        d.fov = scene["fov"]
        d.width, d.height = scene["width"], scene["height"]
        d.pixelsamples = scene.get("pixelsamples", 1)
        d.lightsamples = scene.get("lightsamples", 1)
        d.mindepth, d.maxdepth = scene["mindepth"], scene["maxdepth"]
        if scene.get("uvs") is not None:
            uvs = keep["uvs"] = np.ascontiguousarray(scene["uvs"], dtype=np.float32).reshape(-1)
            d.uvs = L.fptr(uvs)
        shaders = scene.get("shaders") or []
        if shaders:
            sh = keep["shaders"] = (L.Shader * len(shaders))(*shaders)
            d.n_shaders, d.shaders = len(shaders), C.cast(sh, C.POINTER(L.Shader))
        textures = scene.get("textures") or []
        if textures:
            tx = keep["textures"] = (L.Texture * len(textures))(*textures)
            texels = keep["texels"] = np.ascontiguousarray(scene["texels"], dtype=np.float32).reshape(-1)
            d.n_textures, d.textures = len(textures), C.cast(tx, C.POINTER(L.Texture))
            d.n_texels, d.texels = texels.size // 4, L.fptr(texels)
        cam = np.asarray(scene.get("camera", np.concatenate([np.eye(4), np.eye(4)]).reshape(-1)), np.float32)
        for k in range(16):
            d.cam_to_world[k], d.world_to_cam[k] = float(cam[k]), float(cam[16 + k])
        L.check(self.lib.aur_scene_upload(self._ctx, C.byref(d)))
        self.passes_done = 0

    def build(self, width: int = 4, max_leaf: int = 2) -> None:
        """Loop subdivision, flattening and the wide BVH, on the GPU."""
        o = L.AccelOpts(width, max_leaf)
        L.check(self.lib.aur_build_accel(self._ctx, C.byref(o)))

    def set_seed(self, seed: int) -> None:
        L.check(self.lib.aur_set_seed(self._ctx, seed & 0xFFFFFFFF))

    def set_sampler(self, sampler: str) -> None:
        """"hash" (default) or "halton" (Faure-permuted Halton QMC, per-pixel enumeration)."""
        L.check(self.lib.aur_set_sampler(self._ctx, {"hash": L.AUR_SAMPLER_HASH, "halton": L.AUR_SAMPLER_HALTON}[sampler]))

    def set_tri_albedo(self, albedo: Optional[np.ndarray]) -> None:
        if albedo is None:
            L.check(self.lib.aur_set_tri_albedo(self._ctx, None))
            return
        self._albedo = np.ascontiguousarray(albedo, dtype=np.float32).reshape(-1)
        L.check(self.lib.aur_set_tri_albedo(self._ctx, L.fptr(self._albedo)))

    def set_albedo_texture(self, rgb: Optional[np.ndarray], materials=None) -> None:
        """An FP32 RGB albedo texture (height, width, 3), row 0 at the top,
        replacing the diffuse colour of the attributes `materials` (indices:
        geometry objects in sorted name order, then lights) at their
        triangles' uvs; None removes it."""
        if rgb is None:
            L.check(self.lib.aur_set_albedo_texture(self._ctx, 0, 0, None, None, 0))
            self._ptex = None
            return
        tex = np.ascontiguousarray(rgb, dtype=np.float32)
        if tex.ndim != 3 or tex.shape[2] != 3:
            raise ValueError("albedo texture must be (height, width, 3)")
        idx = np.array([int(m) for m in materials], np.int32)
        self._ptex = tex
        self._ptex_attrs = idx
        L.check(self.lib.aur_set_albedo_texture(self._ctx, tex.shape[1], tex.shape[0], L.fptr(tex),
                                                idx.ctypes.data_as(C.POINTER(C.c_int)), len(idx)))

    def bvh(self) -> dict:
        """The built wide BVH: per node the child boxes (n, 6, W) as lo x, hi x,
        lo y, hi y, lo z, hi z rows, the child codes (n, W), the index of
        each node's first inner child (n,) and the triangle records in BVH
        order (n_tris, 3, 4): v0 + global id bits, e1, e2."""
        i = self.info()
        n, w = i["bvh_nodes"], i["bvh_width"]
        raw = np.zeros(n * 8 * w, np.int32)
        tri = np.zeros((i["n_tris"], 3, 4), np.float32)
        L.check(self.lib.aur_bvh_export(self._ctx, raw.ctypes.data_as(C.c_void_p), L.fptr(tri)))
        raw = raw.reshape(n, 8, w)
        return {"width": w, "boxes": raw[:, :6].view(np.float32), "child": raw[:, 6].copy(),
                "first_inner": raw[:, 7, 0].copy(), "tris": tri}

    def info(self) -> dict:
        i = L.SceneInfo()
        L.check(self.lib.aur_scene_info_get(self._ctx, C.byref(i)))
        return {n: getattr(i, n) for n, _ in L.SceneInfo._fields_}

    def export(self) -> dict:
        """The flattened camera-space scene as numpy arrays + structs."""
        i = self.info()
        nv, nt = i["n_verts"], i["n_tris"]
        verts = np.zeros(3 * nv, np.float32)
        tris = np.zeros(3 * nt, np.int32)
        normals = np.zeros(9 * nt, np.float32)
        attr = np.zeros(nt, np.int32)
        mats = (L.Material * max(i["n_attrs"], 1))()
        lights = (L.Light * max(i["n_lights"], 1))()
        L.check(self.lib.aur_scene_export(self._ctx, L.fptr(verts), L.iptr(tris), L.fptr(normals), L.iptr(attr),
                                          mats, lights))
        uvs = np.zeros(6 * nt, np.float32)
        shaders = (L.Shader * max(i["n_shaders"], 1))()
        textures = (L.Texture * max(i["n_textures"], 1))()
        texels = np.zeros(4 * max(i["n_texels"], 1), np.float32)
        cam = np.zeros(32, np.float32)
        L.check(self.lib.aur_scene_export_shading(self._ctx, L.fptr(uvs), shaders, textures, L.fptr(texels),
                                                  L.fptr(cam)))
        return {"verts": verts.reshape(-1, 3), "tris": tris.reshape(-1, 3), "normals": normals.reshape(-1, 9),
                "attr": attr, "materials": [mats[k] for k in range(i["n_attrs"])],
                "lights": [lights[k] for k in range(i["n_lights"])], "fov": i["fov"], "width": i["width"],
                "height": i["height"], "maxdepth": i["maxdepth"], "mindepth": i["mindepth"],
                "pixelsamples": i["pixelsamples"], "lightsamples": i["lightsamples"], "uvs": uvs.reshape(-1, 6),
                "shaders": [shaders[k] for k in range(i["n_shaders"])],
                "textures": [textures[k] for k in range(i["n_textures"])], "texels": texels[:4 * i["n_texels"]],
                "camera": cam}

    # -- rendering -----------------------------------------------------------
    def render_passes(self, first_pass: int, n_passes: int, fetch: bool = True):
        """Renders passes into the device accumulation (sum of per-pass RGBA)."""
        i = self.info()
        out = np.zeros((i["height"], i["width"], 4), np.float32) if fetch else None
        st = L.Stats()
        L.check(self.lib.aur_render(self._ctx, first_pass, n_passes, L.fptr(out) if fetch else None, C.byref(st)))
        self.passes_done = first_pass + n_passes
        return out, st.as_dict()

    def render(self, n_passes: Optional[int] = None) -> np.ndarray:
        """Renderer::renderImage: all passes, returns the mean image (H, W, 4)."""
        n = n_passes if n_passes is not None else self.info()["pixelsamples"]
        acc, self.last_stats = self.render_passes(0, n)
        return acc / max(n, 1)

    def framebuffer(self) -> np.ndarray:
        """The accumulation buffer (sum of per-pass RGBA, (H, W, 4)) of the last render or batch."""
        i = self.info()
        out = np.zeros((i["height"], i["width"], 4), np.float32)
        L.check(self.lib.aur_framebuffer(self._ctx, L.fptr(out)))
        return out

    def display(self) -> np.ndarray:
        """The reference display's FP16 running mean of the passes rendered
        (OpenexrDisplay::appendValue), (H, W, 4) float16."""
        i = self.info()
        out = np.zeros((i["height"], i["width"], 4), np.float16)
        L.check(self.lib.aur_display(self._ctx, out.ctypes.data_as(C.POINTER(C.c_uint16))))
        return out

    def write_image(self, path: str, n_passes: int, meta: dict | None = None) -> None:
        """The mean image (framebuffer / n_passes; n_passes = 0 writes the
        reference display's FP16 running mean instead) as an FP16 RGBA ZIP EXR
        with string metadata, written to <path>_tmp and renamed
        (OpenexrDisplay::draw)."""
        from .exr import _meta
        keys, vals, n = _meta(meta)
        L.check(self.lib.aur_write_image(self._ctx, path.encode(), int(n_passes), keys, vals, n))

    def render_image(self, path: Optional[str] = None, draw_delay: float = 1.0) -> dict:
        """Renderer::renderImage + outputStats + postRenderCleanup
        (renderer.cpp:200-304): all passes, the EXR rewritten after pass 0 and
        every ``draw_delay`` seconds, then a final write carrying renderTime
        and RaySpeed.  ``path`` None writes the scene's options filename.
        Returns the counters plus ``draws`` (files written)."""
        st = L.Stats()
        draws = C.c_int(0)
        L.check(self.lib.aur_render_image(self._ctx, path.encode() if path else None, float(draw_delay),
                                          C.byref(st), C.byref(draws)))
        self.passes_done = self.info()["pixelsamples"]
        self.last_stats = st.as_dict()
        return dict(self.last_stats, draws=draws.value)

    def set_launch(self, trace_block: int = 0, trace_waves_per_cu: int = 0, pool_paths: int = 0,
                   shade_list_tiles: int = 0) -> None:
        """Launch shape of the tracer kernels; 0 keeps a field."""
        o = L.LaunchOpts(int(trace_block), int(trace_waves_per_cu), int(pool_paths), int(shade_list_tiles))
        L.check(self.lib.aur_set_launch(self._ctx, C.byref(o)))

    def get_launch(self) -> dict:
        o = L.LaunchOpts()
        L.check(self.lib.aur_get_launch(self._ctx, C.byref(o)))
        return {"trace_block": o.trace_block, "trace_waves_per_cu": o.trace_waves_per_cu, "pool_paths": o.pool_paths,
                "shade_list_tiles": o.shade_list_tiles}

    def render_paths(self, pixel, sample, slots=None):
        """Per-path RGBA for explicit (pixel, sample) pairs; optional injected uniforms."""
        pixel = np.ascontiguousarray(pixel, dtype=np.int32)
        sample = np.ascontiguousarray(sample, dtype=np.int32)
        n = pixel.size
        out = np.zeros((n, 4), np.float32)
        sl = None
        if slots is not None:
            sl = np.ascontiguousarray(slots, dtype=np.float32)
        st = L.Stats()
        L.check(self.lib.aur_render_paths(self._ctx, L.iptr(pixel), L.iptr(sample), n,
                                          L.fptr(sl) if sl is not None else None, L.fptr(out), C.byref(st)))
        return out, st.as_dict()

    # -- inverse rendering ---------------------------------------------------
    def set_material_params(self, roughness=None, spec_weight=None) -> None:
        """Per-material roughness / Kelemen spec weight (arrays of n_attrs)."""
        self._rough = None if roughness is None else np.ascontiguousarray(roughness, np.float32)
        self._specw = None if spec_weight is None else np.ascontiguousarray(spec_weight, np.float32)
        L.check(self.lib.aur_set_material_params(self._ctx, L.fptr(self._rough) if self._rough is not None else None,
                                                 L.fptr(self._specw) if self._specw is not None else None))

    def keep_path_records(self, first_pass: int, n_passes: int) -> None:
        """Later renders of passes [first_pass, first_pass + n_passes) keep each
        path's radiance record (16 bytes per path), so that an adjoint of those
        passes with unchanged inputs skips its first replay; n_passes = 0
        frees them."""
        L.check(self.lib.aur_keep_path_records(self._ctx, int(first_pass), int(n_passes)))

    def adjoint(self, first_pass: int, n_passes: int, dloss_dimage: np.ndarray, albedo: bool = True,
                texture: bool = False):
        """Path-replay adjoint: returns (albedo grads (n_tris, 3) or None, roughness grads, spec grads, stats),
        and with texture=True the albedo texture's gradient (height, width, 3) as a fifth element."""
        i = self.info()
        dl = np.zeros((i["height"], i["width"], 4), np.float32)
        g = np.asarray(dloss_dimage, np.float32)
        dl[..., :g.shape[-1]] = g.reshape(i["height"], i["width"], -1)[..., :4]
        ga = np.zeros(3 * i["n_tris"], np.float32) if albedo else None
        gr = np.zeros(i["n_attrs"], np.float32)
        gs = np.zeros(i["n_attrs"], np.float32)
        gt = None
        if texture:
            if getattr(self, "_ptex", None) is None:
                raise L.AuroraError(L.AUR_ESTATE, "texture gradients need set_albedo_texture")
            gt = np.zeros(self._ptex.shape, np.float32)
        pg = L.ParamGrads(L.fptr(ga) if albedo else None, L.fptr(gr), L.fptr(gs), L.fptr(gt) if texture else None)
        st = L.Stats()
        L.check(self.lib.aur_adjoint(self._ctx, first_pass, n_passes, L.fptr(dl), C.byref(pg), C.byref(st)))
        out = ((ga.reshape(-1, 3) if albedo else None), gr, gs, st.as_dict())
        return out + (gt,) if texture else out

    # -- instrumentation -----------------------------------------------------
    def set_profiling(self, on: bool) -> None:
        L.check(self.lib.aur_set_profiling(self._ctx, int(on)))

    def kernel_stats(self) -> dict:
        arr = (L.KernelStat * 16)()
        n = C.c_int()
        L.check(self.lib.aur_kernel_stats(self._ctx, arr, 16, C.byref(n)))
        return {arr[k].name.decode(): {"launches": arr[k].launches, "total_ms": arr[k].total_ms}
                for k in range(n.value)}

    # -- parity queries ------------------------------------------------------
    def trace(self, rays: np.ndarray, any_hit: bool = False, extend_kernel: bool = False) -> np.ndarray:
        """rays: (n, 8) float32 rows (ox, oy, oz, tmin, dx, dy, dz, tmax).
        Returns a structured array with t, u, v, prim.  extend_kernel traces
        closest hits through the renderer's extension kernel itself."""
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        n = rays.shape[0]
        hits = np.zeros(n, dtype=[("t", np.float32), ("u", np.float32), ("v", np.float32), ("prim", np.int32)])
        if extend_kernel and any_hit:
            raise ValueError("the extension kernel traces closest hits")
        mode = 2 if extend_kernel else (1 if any_hit else 0)
        L.check(self.lib.aur_trace(self._ctx, rays.ctypes.data_as(C.POINTER(L.Ray)), n,
                                   hits.ctypes.data_as(C.POINTER(L.Hit)), mode))
        return hits

    def light(self, queries) -> list:
        """Light::generateSample / eval / pdf / visible per aur_light_query."""
        arr = (L.LightQuery * len(queries))(*queries)
        res = (L.LightResult * len(queries))()
        L.check(self.lib.aur_light_queries(self._ctx, arr, len(queries), res))
        return [res[k] for k in range(len(queries))]

    def bsdf(self, queries) -> list:
        arr = (L.BsdfQuery * len(queries))(*queries)
        res = (L.BsdfResult * len(queries))()
        L.check(self.lib.aur_bsdf(self._ctx, arr, len(queries), res))
        return [res[k] for k in range(len(queries))]

    def loop_subdivide(self, P, F, levels: int, amplitude: float = 0.0, frequency: float = 1.0, octaves: int = 0,
                       seed: int = 0):
        """GPU Loop subdivision of an indexed mesh; returns (P, F, N) in object space."""
        P = np.ascontiguousarray(P, np.float32).reshape(-1)
        F = np.ascontiguousarray(F, np.int32).reshape(-1)
        nv, nf = P.size // 3, F.size // 3
        cap_f = nf * 4 ** levels
        cap_v = nv + 3 * cap_f  # generous: V_L <= V_0 + sum of edges
        Po = np.zeros(3 * cap_v, np.float32)
        Fo = np.zeros(3 * cap_f, np.int32)
        No = np.zeros(3 * cap_v, np.float32)
        onv, onf = C.c_int(), C.c_int()
        L.check(self.lib.aur_loop_subdivide(self._ctx, nv, nf, L.fptr(P), L.iptr(F), levels, amplitude, frequency,
                                            octaves, seed, L.fptr(Po), L.iptr(Fo), L.fptr(No), cap_v, cap_f,
                                            C.byref(onv), C.byref(onf)))
        return (Po[:3 * onv.value].reshape(-1, 3), Fo[:3 * onf.value].reshape(-1, 3),
                No[:3 * onv.value].reshape(-1, 3))

    def kelemen_table(self, reflectance: float, distribution: int, n_samples: int, seed: int) -> np.ndarray:
        out = np.zeros(64 * 32, np.float32)
        L.check(self.lib.aur_kelemen_table(self._ctx, reflectance, distribution, n_samples, seed, L.fptr(out)))
        return out.reshape(64, 32)


def batch_submit(exec_renderer: "Renderer", jobs, order: int = 0, use_graph: bool = True, iters_per_poll: int = 8):
    """Renders jobs [(renderer, first_pass, n_passes), ...] in one persistent
    launch on exec_renderer's stream.  Returns (per-job result dicts in job
    order, total stats dict).  Each job's image lands in its renderer's
    framebuffer (``Renderer.framebuffer``)."""
    n = len(jobs)
    arr = (L.Job * n)()
    for k, (r, fp, np_) in enumerate(jobs):
        arr[k].scene = r._ctx.value
        arr[k].first_pass = int(fp)
        arr[k].n_passes = int(np_)
    res = (L.JobResult * n)()
    st = L.Stats()
    opts = L.BatchOpts(int(order), int(bool(use_graph)), int(iters_per_poll))
    lib = exec_renderer.lib
    L.check(lib.aur_batch_submit(exec_renderer._ctx, arr, n, C.byref(opts), res, C.byref(st)))
    return [res[k].as_dict() for k in range(n)], st.as_dict()
