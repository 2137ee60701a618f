"""Deterministic scene generators for the benchmark workloads.

Every scene is written in Aurora's own input format: a JSON scene file (the
PyAurora export, parsed by the reference's ``JsonParser::parseScene``,
src/translator/jsonParser.cpp:30-59) next to Wavefront OBJ meshes (parsed by
``ObjParser::run``, src/geometry/objParser.cpp:58-202) and a ``config.json`` of
path templates (Aurora/config.json:1-7).

Build extensions use keys the reference ignores ("TODO: check for
invalid/unknown keys", jsonParser.cpp:47):
  * material ``"distribution": "ggx"`` selects the GGX microfacet distribution
    for Cook-Torrance lobes (Beckmann otherwise);
  * geometry ``"displace": {amplitude, octaves, frequency, seed}`` applies
    radial value-noise displacement after Loop subdivision.

Textures (the "shaders" section, env maps) are OpenEXR files written next to
the scene.

Random numbers come from SplitMix64 with an explicit ``(x >> 40) * 2**-24``
float conversion, so the scenes are identical on every platform.
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass, field

import numpy as np


class SplitMix64:
    """SplitMix64 stream; ``uniform()`` returns floats in [0, 1)."""

    def __init__(self, seed: int):
        self.state = seed & 0xFFFFFFFFFFFFFFFF

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 40) * (1.0 / 16777216.0))

    def log_uniform(self, lo: float, hi: float) -> float:
        return math.exp(self.uniform(math.log(lo), math.log(hi)))

    def randint(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1)) % (hi - lo + 1)


# --------------------------------------------------------------------------
# OBJ writers
# --------------------------------------------------------------------------

def write_obj(path: str, verts, faces, normals=None, uvs=None) -> None:
    """Writes an OBJ with ``v``/``vn``/``f v//n`` lines (1-based indices).

    ``normals`` are per vertex (same count as ``verts``); when omitted each
    face gets its geometric normal, which ``ObjTriangleMesh`` copies per
    corner (objTriangleMesh.cpp:228-242).  ``uvs`` (per vertex) adds ``vt``
    lines and ``f v/t/n`` corners.
    """
    verts = np.asarray(verts, dtype=np.float64)
    faces = np.asarray(faces, dtype=np.int64)
    lines = ["# generated by aurora_amd.scenes"]
    lines += ["v %.9g %.9g %.9g" % tuple(v) for v in verts]
    if uvs is not None:
        uvs = np.asarray(uvs, dtype=np.float64)
        lines += ["vt %.9g %.9g" % tuple(t) for t in uvs]
        if normals is None:
            normals = np.zeros_like(verts)
            for f in faces:
                a, b, c = verts[f[0]], verts[f[1]], verts[f[2]]
                n = np.cross(b - a, c - a)
                normals[f] += n / (np.linalg.norm(n) or 1.0)
            normals /= np.maximum(np.linalg.norm(normals, axis=1, keepdims=True), 1e-12)
        lines += ["vn %.9g %.9g %.9g" % tuple(n) for n in np.asarray(normals, dtype=np.float64)]
        lines += ["f %d/%d/%d %d/%d/%d %d/%d/%d" % tuple(int(x) + 1 for x in (f[0], f[0], f[0], f[1], f[1], f[1],
                                                                             f[2], f[2], f[2]))
                  for f in faces]
        with open(path, "w") as fh:
            fh.write("\n".join(lines) + "\n")
        return
    if normals is None:
        fn = []
        for f in faces:
            a, b, c = verts[f[0]], verts[f[1]], verts[f[2]]
            n = np.cross(b - a, c - a)
            n = n / (np.linalg.norm(n) or 1.0)
            fn.append(n)
        lines += ["vn %.9g %.9g %.9g" % tuple(n) for n in fn]
        lines += ["f %d//%d %d//%d %d//%d" % (f[0] + 1, i + 1, f[1] + 1, i + 1, f[2] + 1, i + 1)
                  for i, f in enumerate(faces)]
    else:
        normals = np.asarray(normals, dtype=np.float64)
        lines += ["vn %.9g %.9g %.9g" % tuple(n) for n in normals]
        lines += ["f %d//%d %d//%d %d//%d" % (f[0] + 1, f[0] + 1, f[1] + 1, f[1] + 1, f[2] + 1, f[2] + 1)
                  for f in faces]
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def write_grid_obj(path: str, verts, faces, normals, uvs) -> None:
    """Writes a large indexed mesh whose vertices carry their own uv and normal
    (``f a/a/a b/b/b c/c/c``), with vectorised formatting."""
    faces = np.asarray(faces, dtype=np.int64) + 1
    with open(path, "w") as fh:
        fh.write("# generated by aurora_amd.scenes\n")
        np.savetxt(fh, np.asarray(verts, np.float64), fmt="v %.7g %.7g %.7g")
        np.savetxt(fh, np.asarray(uvs, np.float64), fmt="vt %.7g %.7g")
        np.savetxt(fh, np.asarray(normals, np.float64), fmt="vn %.6g %.6g %.6g")
        np.savetxt(fh, np.repeat(faces, 3, axis=1), fmt="f %d/%d/%d %d/%d/%d %d/%d/%d")


def _write_grid(job):
    path, mesh = job
    if "terrain" in mesh:
        mesh = terrain_patch(**mesh["terrain"])
    write_grid_obj(path, **mesh)
    return path


def _write_grids(jobs):
    """Generates (terrain descriptions) and writes large meshes, in parallel
    worker processes when there are several."""
    if len(jobs) <= 1:
        for j in jobs:
            _write_grid(j)
        return
    import concurrent.futures as cf
    import multiprocessing as mp
    workers = max(1, min(len(jobs), os.cpu_count() or 1, 16))
    with cf.ProcessPoolExecutor(workers, mp_context=mp.get_context("fork")) as ex:
        list(ex.map(_write_grid, jobs))


def _hash2(ix, iz, seed):
    """Integer lattice hash to [0, 1) (vectorised, uint32 arithmetic that wraps)."""
    with np.errstate(over="ignore"):
        h = (np.asarray(ix).astype(np.uint32) * np.uint32(0x8DA6B343)) ^ \
            (np.asarray(iz).astype(np.uint32) * np.uint32(0xD8163841))
        h = h ^ np.uint32(seed & 0xFFFFFFFF)
        h = h ^ (h >> np.uint32(16))
        h = h * np.uint32(0x7FEB352D)
        h = h ^ (h >> np.uint32(15))
        h = h * np.uint32(0x846CA68B)
        h = h ^ (h >> np.uint32(16))
    return (h >> np.uint32(8)).astype(np.float64) * (1.0 / (1 << 24))


def fbm2(x, z, seed: int, octaves: int, frequency: float) -> np.ndarray:
    """Value-noise fractal sum in [-1, 1]: octave k has frequency f 2^k and
    weight 2^-k (smoothstep-interpolated lattice values), normalised by the
    total weight."""
    x = np.asarray(x, np.float64)
    z = np.asarray(z, np.float64)
    out = np.zeros(np.broadcast(x, z).shape)
    wsum = 0.0
    for k in range(octaves):
        f, w = frequency * (2.0 ** k), 0.5 ** k
        px, pz = x * f, z * f
        ix, iz = np.floor(px), np.floor(pz)
        fx, fz = px - ix, pz - iz
        sx, sz = fx * fx * (3 - 2 * fx), fz * fz * (3 - 2 * fz)
        ix, iz = ix.astype(np.int64), iz.astype(np.int64)
        sk = seed * 1013 + k * 7919
        # the lattice values over the cells the points touch, then gathered
        x0, z0 = int(ix.min()), int(iz.min())
        gx, gz = np.meshgrid(np.arange(x0, int(ix.max()) + 2), np.arange(z0, int(iz.max()) + 2))
        lat = _hash2(gx, gz, sk)
        jx, jz = ix - x0, iz - z0
        a, b = lat[jz, jx], lat[jz, jx + 1]
        c, d = lat[jz + 1, jx], lat[jz + 1, jx + 1]
        v = (a * (1 - sx) + b * sx) * (1 - sz) + (c * (1 - sx) + d * sx) * sz
        out += w * (2.0 * v - 1.0)
        wsum += w
    return out / wsum


def terrain_patch(n: int, size: float, seed: int, amplitude: float, origin=(0.0, 0.0), uv0=(0.0, 0.0),
                  uv1=(1.0, 1.0), octaves: int = 5, frequency: float = 0.15) -> dict:
    """An n x n-quad heightfield (2 n^2 triangles) over [x0, x0 + size] x
    [z0, z0 + size] in object space (y up), heights amplitude * fbm2 of the
    world position; uvs span [uv0, uv1] linearly in x and z; area-weighted
    vertex normals.  Returns the mesh as write_grid_obj's keyword arguments."""
    x0, z0 = origin
    t = np.linspace(0.0, 1.0, n + 1)
    X, Z = np.meshgrid(x0 + t * size, z0 + t * size)      # rows advance in z
    Y = amplitude * fbm2(X, Z, seed, octaves, frequency)
    verts = np.stack([X, Y, Z], axis=-1).reshape(-1, 3)
    U, V = np.meshgrid(uv0[0] + t * (uv1[0] - uv0[0]), uv0[1] + t * (uv1[1] - uv0[1]))
    uvs = np.stack([U, V], axis=-1).reshape(-1, 2)
    j, i = np.meshgrid(np.arange(n), np.arange(n))
    a = (i * (n + 1) + j).reshape(-1)
    b, c = a + 1, a + (n + 1)
    d = c + 1
    # counter-clockwise seen from +y: (a, c, b) and (b, c, d)
    faces = np.concatenate([np.stack([a, c, b], 1), np.stack([b, c, d], 1)])
    p = verts[faces]
    fn = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    normals = np.zeros_like(verts)
    for k in range(3):
        np.add.at(normals, faces[:, k], fn)
    normals /= np.maximum(np.linalg.norm(normals, axis=1, keepdims=True), 1e-12)
    return {"verts": verts, "faces": faces, "normals": normals, "uvs": uvs}


def albedo_texture(res: int, seed: int, octaves: int = 5, lo: float = 0.1, hi: float = 0.9) -> np.ndarray:
    """A res x res RGB FP32 albedo texture of octave value noise, one field per
    channel, rescaled to [lo, hi]."""
    t = (np.arange(res) + 0.5) / res
    out = np.empty((res, res, 3), np.float32)
    for ch in range(3):
        v = _grid_fbm(t, seed * 3 + ch, octaves, 4.0)
        v = (v - v.min()) / max(v.max() - v.min(), 1e-12)
        out[..., ch] = lo + v * (hi - lo)
    return out


def _grid_fbm(t, seed: int, octaves: int, frequency: float) -> np.ndarray:
    """fbm2 over the grid (x, z) = (t[j], t[i]), interpolating along x once
    per lattice row and then along z: the same operations per point as fbm2,
    in row gathers instead of point gathers."""
    out = np.zeros((t.size, t.size))
    wsum = 0.0
    for k in range(octaves):
        f, w = frequency * (2.0 ** k), 0.5 ** k
        p = t * f
        i = np.floor(p)
        fr = p - i
        sm = fr * fr * (3 - 2 * fr)
        i = i.astype(np.int64)
        i0 = int(i.min())
        gx, gz = np.meshgrid(np.arange(i0, int(i.max()) + 2), np.arange(i0, int(i.max()) + 2))
        lat = _hash2(gx, gz, seed * 1013 + k * 7919)
        j = i - i0
        rows = lat[:, j] * (1 - sm) + lat[:, j + 1] * sm          # along x, per lattice row
        v = rows[j] * (1 - sm)[:, None] + rows[j + 1] * sm[:, None]  # along z
        out += w * (2.0 * v - 1.0)
        wsum += w
    return out / wsum


def quad(p0, p1, p2, p3):
    """Two triangles (p0,p1,p2), (p0,p2,p3)."""
    return [p0, p1, p2, p3], [[0, 1, 2], [0, 2, 3]]


def box_mesh(lo, hi):
    """Axis-aligned box, 12 triangles, outward winding."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    v = [[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
         [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]]
    f = [[0, 3, 2], [0, 2, 1],   # -z
         [4, 5, 6], [4, 6, 7],   # +z
         [0, 4, 7], [0, 7, 3],   # -x
         [1, 2, 6], [1, 6, 5],   # +x
         [0, 1, 5], [0, 5, 4],   # -y
         [3, 7, 6], [3, 6, 2]]   # +y
    return v, f


def icosahedron():
    """Unit-radius icosahedron: 12 vertices, 20 faces, outward winding."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
         [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
         [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
         [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]]
    return v, np.array(f, dtype=np.int64)


# --------------------------------------------------------------------------
# scene assembly
# --------------------------------------------------------------------------

@dataclass
class SceneSpec:
    """In-memory form of one scene file (the PyAurora ``Scene`` contents)."""
    options: dict
    camera: dict
    materials: dict = field(default_factory=dict)
    objects: dict = field(default_factory=dict)
    shaders: dict = field(default_factory=dict)
    meshes: dict = field(default_factory=dict)  # obj file name -> (verts, faces, normals|None[, uvs])
    textures: dict = field(default_factory=dict)  # exr file name -> (H, W, 4) float array, FP16 flag

    def attr_index(self, name: str) -> int:
        """The attribute index of geometry object `name`: geometry objects take
        indices in sorted key order, lights follow (renderer.cpp:118-134)."""
        geo = sorted(k for k, o in self.objects.items() if o.get("type") == "geometry")
        return geo.index(name)

    def write(self, out_dir: str, scene_name: str = "scene.asc") -> str:
        os.makedirs(out_dir, exist_ok=True)
        mesh_dir = os.path.join(out_dir, "meshes")
        os.makedirs(mesh_dir, exist_ok=True)
        grids = []
        for name, mesh in self.meshes.items():
            if isinstance(mesh, dict):
                grids.append((os.path.join(mesh_dir, name), mesh))
            else:
                write_obj(os.path.join(mesh_dir, name), *mesh)
        _write_grids(grids)
        if self.textures:
            from . import exr
            tex_dir = os.path.join(out_dir, "textures")
            os.makedirs(tex_dir, exist_ok=True)
            for name, (rgba, half) in self.textures.items():
                exr.write_exr(os.path.join(tex_dir, name), rgba, compression="zip", half=half)
        cfg = {"basepath": os.path.abspath(out_dir) + "/", "cachepath": "$BASE/cache/",
               "renderpath": "$BASE/render/", "texturepath": "$BASE/textures",
               "hdripath": "$BASE/hdri"}
        with open(os.path.join(out_dir, "config.json"), "w") as fh:
            json.dump(cfg, fh, indent=4)
        # the reference opens OBJ paths verbatim (jsonParser.cpp:227, no template
        # expansion), so geometry paths are written absolute
        objects = {}
        for k, o in self.objects.items():
            o = dict(o)
            if o.get("type") == "geometry":
                o["path"] = os.path.abspath(os.path.join(out_dir, o["path"]))
            if o.get("type") == "envlight" and o.get("envmap"):
                o["envmap"] = os.path.abspath(os.path.join(out_dir, o["envmap"]))
            objects[k] = o
        shaders = {}
        for k, sh in self.shaders.items():
            sh = dict(sh)
            sh["texturename"] = os.path.abspath(os.path.join(out_dir, sh["texturename"]))
            shaders[k] = sh
        root = {"options": self.options, "renderCam": self.camera, "materials": self.materials,
                "objects": objects, "shaders": shaders}
        path = os.path.join(out_dir, scene_name)
        with open(path, "w") as fh:
            json.dump(root, fh, indent=4)
        return path


def _geo(name, obj, material, transforms=None, subdlevels=0, **extra):
    d = {"type": "geometry", "name": name, "path": "meshes/" + obj, "material": material,
         "transforms": transforms or [], "subdlevels": int(subdlevels)}
    d.update(extra)
    return d


def _square_light(name, color, sx, sy, transforms, exposure=0.0):
    return {"type": "light", "name": name, "scaleX": float(sx), "scaleY": float(sy),
            "exposure": float(exposure), "color": [float(c) for c in color], "transforms": transforms}


def _options(w, h, spp, maxdepth, mindepth, filename="render.exr"):
    return {"resolution": [int(w), int(h)], "pixelsamples": int(spp), "lightsamples": 1,
            "mindepth": int(mindepth), "maxdepth": int(maxdepth), "filename": "$RENDER/" + filename}


def box_room(res=256, spp=16, maxdepth=5, mindepth=5, kelemen=False, seed=1) -> SceneSpec:
    """Closed box room (the 256x256, 16 spp, depth-5 CPU-runnable workload).

    Floor, ceiling, back, left and right walls (10 triangles), a matte box and
    a GGX Cook-Torrance box (24 triangles), one ceiling quad light of radiance
    (17, 12, 4); pinhole camera with a 40 degree horizontal field of view
    (``fov`` = tan(20 deg), camera.cpp:36-53).  With ``kelemen`` every
    material is a Kelemen coat (the inverse-rendering variant).
    """
    rng = SplitMix64(seed)
    s = SceneSpec(options=_options(res, res, spp, maxdepth, mindepth),
                  camera={"fov": math.tan(math.radians(20.0)),
                          "transforms": [{"translate": [0.0, 1.0, 3.5]}]})
    x0, x1, y0, y1, z0, z1 = -1.0, 1.0, 0.0, 2.0, -1.0, 1.0
    fv, ff = [], []
    for q in (([x0, y0, z1], [x1, y0, z1], [x1, y0, z0], [x0, y0, z0]),     # floor
              ([x0, y1, z0], [x1, y1, z0], [x1, y1, z1], [x0, y1, z1]),     # ceiling
              ([x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0])):    # back
        v, f = quad(*q)
        ff += [[a + len(fv) for a in t] for t in f]
        fv += v
    s.meshes["walls_white.obj"] = (fv, ff, None)
    v, f = quad([x0, y0, z1], [x0, y0, z0], [x0, y1, z0], [x0, y1, z1])
    s.meshes["wall_left.obj"] = (v, f, None)
    v, f = quad([x1, y0, z0], [x1, y0, z1], [x1, y1, z1], [x1, y1, z0])
    s.meshes["wall_right.obj"] = (v, f, None)
    v, f = box_mesh([-0.65, 0.0, -0.6], [-0.05, 1.2, 0.0])
    s.meshes["box_tall.obj"] = (v, f, None)
    v, f = box_mesh([0.1, 0.0, -0.1], [0.7, 0.6, 0.5])
    s.meshes["box_short.obj"] = (v, f, None)

    if kelemen:
        def kel(name):
            s.materials[name] = {"type": "kelemenMaterial", "diffusecolor": [0.5, 0.5, 0.5],
                                 "speccolor": [0.5, 0.5, 0.5], "reflectance": 0.04,
                                 "roughness": 0.4, "distribution": "ggx"}
        for n in ("white", "red", "green", "boxmatte", "boxspec"):
            kel(n)
    else:
        s.materials["white"] = {"type": "matteMaterial", "color": [0.73, 0.73, 0.73]}
        s.materials["red"] = {"type": "matteMaterial", "color": [0.63, 0.12, 0.05]}
        s.materials["green"] = {"type": "matteMaterial", "color": [0.14, 0.45, 0.09]}
        s.materials["boxmatte"] = {"type": "matteMaterial", "color": [0.73, 0.73, 0.73]}
        s.materials["boxspec"] = {"type": "metalMaterial", "reflectance": [0.04, 0.04, 0.04],
                                  "roughnessA": 0.3, "roughnessB": 0.3, "mix": 0.5,
                                  "gain": [1.0, 1.0, 1.0], "distribution": "ggx"}
    s.objects["walls_white"] = _geo("walls_white", "walls_white.obj", "white")
    s.objects["wall_left"] = _geo("wall_left", "wall_left.obj", "red")
    s.objects["wall_right"] = _geo("wall_right", "wall_right.obj", "green")
    s.objects["box_tall"] = _geo("box_tall", "box_tall.obj", "boxmatte")
    s.objects["box_short"] = _geo("box_short", "box_short.obj", "boxspec")
    # object-space light normal is -z (squareLight.cpp:27-39): rotateX(-90) faces it down
    s.objects["light_ceiling"] = _square_light("light_ceiling", [17.0, 12.0, 4.0], 0.25, 0.25,
                                               [{"translate": [0.0, 1.98, 0.0]}, {"rotateX": -90.0}])
    del rng
    return s


def displaced_spheres(width=1920, height=1080, spp=64, maxdepth=8, mindepth=8, n_spheres=8,
                      levels=7, seed=2, ground=20.0) -> SceneSpec:
    """Large single scene: displaced Loop-subdivided icospheres on a ground quad.

    Eight unit icospheres, Loop-subdivided ``levels`` times (327,680 triangles
    each at 7 levels), displaced radially by 3-octave value noise (amplitude
    0.05 * radius), Kelemen-coated with GGX roughness ~U[0.05, 0.8] and
    albedo ~U[0.1, 0.9]^3, on a 20x20 ground quad, lit by 4 quad lights of
    grey radiance ~U[5, 20] plus a constant environment of 0.1.
    """
    rng = SplitMix64(seed)
    s = SceneSpec(options=_options(width, height, spp, maxdepth, mindepth),
                  camera={"fov": 0.55,
                          "transforms": [{"translate": [0.0, 7.0, 13.0]}, {"rotateX": -34.0}]})
    iv, iff = icosahedron()
    s.meshes["icosahedron.obj"] = (iv, iff, iv)
    h = ground / 2.0
    v, f = quad([-h, 0.0, h], [h, 0.0, h], [h, 0.0, -h], [-h, 0.0, -h])
    s.meshes["ground.obj"] = (v, f, None)
    s.materials["ground"] = {"type": "matteMaterial", "color": [0.5, 0.5, 0.5]}
    s.objects["ground"] = _geo("ground", "ground.obj", "ground")
    placed = []
    for i in range(n_spheres):
        for _ in range(1000):
            x, z = rng.uniform(-h + 2, h - 2), rng.uniform(-h + 2, h - 2)
            if all((x - a) ** 2 + (z - b) ** 2 > (2.3) ** 2 for a, b in placed):
                break
        placed.append((x, z))
        rough = rng.uniform(0.05, 0.8)
        alb = [rng.uniform(0.1, 0.9) for _ in range(3)]
        mname = "sphere%02d_mat" % i
        s.materials[mname] = {"type": "kelemenMaterial", "diffusecolor": alb,
                              "speccolor": [1.0, 1.0, 1.0], "reflectance": 0.04,
                              "roughness": rough, "distribution": "ggx"}
        name = "sphere%02d" % i
        s.objects[name] = _geo(name, "icosahedron.obj", mname, [{"translate": [x, 1.0, z]}],
                               subdlevels=levels,
                               displace={"amplitude": 0.05, "octaves": 3, "frequency": 2.0,
                                         "seed": int(rng.next_u64() & 0x7FFFFFFF)})
    for i in range(4):
        e = rng.uniform(5.0, 20.0)
        x, z = rng.uniform(-6, 6), rng.uniform(-6, 6)
        s.objects["light%d" % i] = _square_light("light%d" % i, [e, e, e], 0.75, 0.75,
                                                 [{"translate": [x, 8.0, z]}, {"rotateX": -90.0}])
    s.objects["env"] = {"type": "envlight", "name": "env", "color": [0.1, 0.1, 0.1],
                        "exposure": 0.0, "envmap": "", "transforms": []}
    return s


def large_scene(width=3840, height=2160, spp=256, maxdepth=16, mindepth=16, n_spheres=32, levels=7,
                patches=(4, 2), patch_quads=708, patch_size=10.0, tex_res=2048, seed=5):
    """The large differentiable scene: displaced icospheres on tessellated
    noise terrain, with an FP32 albedo texture on the terrain as the gradient
    target.

    * terrain: patches[0] x patches[1] heightfield patches of patch_quads^2
      quads (2 * 708^2 = 1,002,528 triangles each), patch_size units square,
      heights 0.6 * 5-octave value noise; one matte material whose diffuse
      colour is a texture2DShader over uvs spanning the whole terrain;
    * n_spheres unit icospheres, Loop-subdivided `levels` times (327,680
      triangles each at 7) and displaced radially by 3-octave value noise
      (amplitude 0.05 radius), resting on the terrain, Kelemen-coated as in
      the displaced-spheres scene;
    * 8 quad lights of grey radiance ~U[5, 20] and a constant environment of 0.1.

    Returns (spec, truth): the scene and the tex_res x tex_res RGB FP32
    ground-truth albedo texture (5-octave value noise in [0.1, 0.9]), which
    the scene's texture file also holds (FP16 rounded).  The terrain's
    attribute is spec.attr_index("terrain00").
    """
    rng = SplitMix64(seed)
    px, pz = patches
    W, D = px * patch_size, pz * patch_size
    s = SceneSpec(options=_options(width, height, spp, maxdepth, mindepth),
                  camera={"fov": 0.62, "transforms": [{"translate": [0.0, 0.55 * D + 3.0, 0.5 * D + 0.45 * W]},
                                                      {"rotateX": -30.0}]})
    truth = albedo_texture(tex_res, seed)
    rgba = np.ones((tex_res, tex_res, 4), np.float32)
    rgba[..., :3] = truth
    s.textures["terrain_albedo.exr"] = (rgba, True)
    s.shaders["terrain_albedo"] = {"type": "texture2DShader", "attributetype": "color",
                                   "texturename": "textures/terrain_albedo.exr"}
    s.materials["terrain"] = {"type": "matteMaterial", "color": "terrain_albedo"}
    amp, tseed = 0.6, int(rng.next_u64() & 0x7FFFFFFF)
    for i in range(px):
        for j in range(pz):
            name = "terrain%d%d" % (i, j)
            # generated when the scene is written
            s.meshes[name + ".obj"] = {"terrain": dict(
                n=patch_quads, size=patch_size, seed=tseed, amplitude=amp,
                origin=(-W / 2 + i * patch_size, -D / 2 + j * patch_size),
                uv0=(i / px, j / pz), uv1=((i + 1) / px, (j + 1) / pz))}
            s.objects[name] = _geo(name, name + ".obj", "terrain")
    iv, iff = icosahedron()
    s.meshes["icosahedron.obj"] = (iv, iff, iv)
    placed = []
    for k in range(n_spheres):
        for _ in range(1000):
            x, z = rng.uniform(-W / 2 + 1.5, W / 2 - 1.5), rng.uniform(-D / 2 + 1.5, D / 2 - 1.5)
            if all((x - a) ** 2 + (z - b) ** 2 > 2.3 ** 2 for a, b in placed):
                break
        placed.append((x, z))
        y = float(amp * fbm2(np.array(x), np.array(z), tseed, 5, 0.15)) + 1.0
        mname = "sphere%02d_mat" % k
        s.materials[mname] = {"type": "kelemenMaterial", "diffusecolor": [rng.uniform(0.1, 0.9) for _ in range(3)],
                              "speccolor": [1.0, 1.0, 1.0], "reflectance": 0.04,
                              "roughness": rng.uniform(0.05, 0.8), "distribution": "ggx"}
        name = "sphere%02d" % k
        s.objects[name] = _geo(name, "icosahedron.obj", mname, [{"translate": [x, y, z]}], subdlevels=levels,
                               displace={"amplitude": 0.05, "octaves": 3, "frequency": 2.0,
                                         "seed": int(rng.next_u64() & 0x7FFFFFFF)})
    for k in range(8):
        e = rng.uniform(5.0, 20.0)
        x, z = rng.uniform(-W / 2 + 2, W / 2 - 2), rng.uniform(-D / 2 + 2, D / 2 - 2)
        s.objects["light%d" % k] = _square_light("light%d" % k, [e, e, e], 1.0, 1.0,
                                                 [{"translate": [x, 10.0, z]}, {"rotateX": -90.0}])
    s.objects["env"] = {"type": "envlight", "name": "env", "color": [0.1, 0.1, 0.1],
                        "exposure": 0.0, "envmap": "", "transforms": []}
    return s, truth


def batch_job(seed: int, k: int) -> SceneSpec:
    """Job k of the batched uneven workload.

    A closed room (the box-room walls, scaled to 4 x 2.5 x 4) holding a random
    mix of displaced Loop-subdivided icospheres and boxes whose triangle count
    is log-uniform in [1e3, 5e5]; resolution uniform over {128, 256, 512,
    1024}^2, spp log-uniform in [4, 256], max depth uniform in [3, 12] (min
    depth = max depth, as the box room).  Materials are matte (albedo
    ~U[0.1, 0.9]^3) or GGX metal (F0 0.04, roughness ~U[0.05, 0.8]); one
    ceiling quad light of radiance (17, 12, 4).
    """
    rng = SplitMix64((seed * 0x9E3779B1 + k * 0x85EBCA77 + 1) & 0xFFFFFFFFFFFF)
    target = rng.log_uniform(1e3, 5e5)
    res = (128, 256, 512, 1024)[rng.randint(0, 3)]
    spp = max(4, min(256, int(round(rng.log_uniform(4, 256)))))
    depth = rng.randint(3, 12)
    s = SceneSpec(options=_options(res, res, spp, depth, depth),
                  camera={"fov": math.tan(math.radians(30.0)),
                          "transforms": [{"translate": [0.0, 1.25, 5.5]}, {"rotateX": -8.0}]})
    x0, x1, y0, y1, z0, z1 = -2.0, 2.0, 0.0, 2.5, -2.0, 2.0
    fv, ff = [], []
    for q in (([x0, y0, z1], [x1, y0, z1], [x1, y0, z0], [x0, y0, z0]),
              ([x0, y1, z0], [x1, y1, z0], [x1, y1, z1], [x0, y1, z1]),
              ([x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0]),
              ([x0, y0, z1], [x0, y0, z0], [x0, y1, z0], [x0, y1, z1]),
              ([x1, y0, z0], [x1, y0, z1], [x1, y1, z1], [x1, y1, z0])):
        v, f = quad(*q)
        ff += [[a + len(fv) for a in t] for t in f]
        fv += v
    s.meshes["room.obj"] = (fv, ff, None)
    s.materials["room"] = {"type": "matteMaterial", "color": [0.73, 0.73, 0.73]}
    s.objects["room"] = _geo("room", "room.obj", "room")
    iv, iff = icosahedron()
    s.meshes["icosahedron.obj"] = (iv, iff, iv)

    def material(name):
        if rng.uniform() < 0.5:
            s.materials[name] = {"type": "matteMaterial", "color": [rng.uniform(0.1, 0.9) for _ in range(3)]}
        else:
            s.materials[name] = {"type": "metalMaterial", "reflectance": [0.04, 0.04, 0.04],
                                 "roughnessA": rng.uniform(0.05, 0.8), "roughnessB": 0.0, "mix": 1.0,
                                 "gain": [1.0, 1.0, 1.0], "distribution": "ggx"}
        return name

    # spheres: n in [1, 4] at the level that brings n * 20 * 4^L closest to the target
    level = max(1, min(7, int(math.floor(math.log(target / 20.0, 4.0)))))
    n_sph = max(1, min(4, int(round(target / (20.0 * 4 ** level)))))
    n_box = rng.randint(1, 3)
    placed = []

    def spot(r):
        for _ in range(200):
            x, z = rng.uniform(x0 + r + 0.1, x1 - r - 0.1), rng.uniform(z0 + r + 0.1, z1 - r - 0.1)
            if all((x - a) ** 2 + (z - b) ** 2 > (r + c + 0.05) ** 2 for a, b, c in placed):
                break
        placed.append((x, z, r))
        return x, z

    for i in range(n_sph):
        r = rng.uniform(0.3, 0.55)
        x, z = spot(r)
        name = "sphere%d" % i
        s.objects[name] = _geo(name, "icosahedron.obj", material(name + "_mat"),
                               [{"translate": [x, r, z]}, {"scale": [r, r, r]}], subdlevels=level,
                               displace={"amplitude": 0.05, "octaves": 3, "frequency": 2.0,
                                         "seed": int(rng.next_u64() & 0x7FFFFFFF)})
    for i in range(n_box):
        hx, hy, hz = rng.uniform(0.15, 0.4), rng.uniform(0.2, 0.8), rng.uniform(0.15, 0.4)
        x, z = spot(max(hx, hz) * 1.42)
        v, f = box_mesh([x - hx, 0.0, z - hz], [x + hx, 2 * hy, z + hz])
        s.meshes["box%d.obj" % i] = (v, f, None)
        name = "box%d" % i
        s.objects[name] = _geo(name, "box%d.obj" % i, material(name + "_mat"))
    s.objects["light_ceiling"] = _square_light("light_ceiling", [17.0, 12.0, 4.0], 0.5, 0.5,
                                               [{"translate": [0.0, 2.48, 0.0]}, {"rotateX": -90.0}])
    return s


def icosphere(level: int):
    """Unit icosphere by midpoint subdivision of the icosahedron (projected to
    the sphere), with the vertex normals of a sphere."""
    v, f = icosahedron()
    verts = [tuple(x) for x in v]
    faces = [tuple(x) for x in f]
    for _ in range(level):
        cache, nf = {}, []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = (np.array(verts[a]) + np.array(verts[b])) / 2.0
                verts.append(tuple(m / np.linalg.norm(m)))
                cache[key] = len(verts) - 1
            return cache[key]
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    v = np.array(verts, dtype=np.float64)
    return v, np.array(faces, dtype=np.int64), v.copy()


def _tex_noise(rng: "SplitMix64", h: int, w: int, lo, hi) -> np.ndarray:
    a = np.array([[rng.uniform() for _ in range(w)] for _ in range(h)])
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    out = np.ones((h, w, 4))
    out[..., :3] = lo + a[..., None] * (hi - lo)
    return out.astype(np.float32)


def textured_stage(res=96, spp=4, maxdepth=4, mindepth=4, sphere_subd=0, seed=5) -> SceneSpec:
    """Open stage exercising the shader and environment-map paths.

    * floor: matte, colour from a uv-mapped checker (texture2DShader), uvs
      tiling 3x and reaching negative values;
    * back wall: Kelemen coat, uv-mapped diffuse colour, uv-mapped float
      roughness, a uv-mapped normal map;
    * sphere: matte with a triplanar colour (``sphere_subd`` > 0 subdivides it
      on the device instead of using the 320-triangle mesh);
    * rotated box without uvs: metal with triplanar reflectance, triplanar
      float roughnessA, a uv-mapped float mix (read at uv = 0) and a constant
      normal map;
    * glass ball (glassMaterial: Fresnel-weighted SpecMirror / SpecTransmit)
      and a mirror panel (glass with reflectance 1), whose rays that miss see
      the environment in the camera ray's direction;
    * lights: a square light, an env light with a lat-long map (stored as FP32,
      rounded to FP16 when loaded) and a constant env light;
    * a rotated camera, so the env map's camera-to-world transform matters.
    """
    rng = SplitMix64(seed)
    s = SceneSpec(options=_options(res, res, spp, maxdepth, mindepth),
                  camera={"fov": math.tan(math.radians(28.0)),
                          "transforms": [{"translate": [0.2, 1.3, 4.6]}, {"rotateY": 4.0}, {"rotateX": -9.0}]})
    # floor with uvs in [-1, 2]
    fv = [[-3.0, 0.0, 3.0], [3.0, 0.0, 3.0], [3.0, 0.0, -3.0], [-3.0, 0.0, -3.0]]
    fuv = [[-1.0, -1.0], [2.0, -1.0], [2.0, 2.0], [-1.0, 2.0]]
    s.meshes["floor.obj"] = (fv, [[0, 1, 2], [0, 2, 3]], None, fuv)
    wv = [[-3.0, 0.0, -2.5], [3.0, 0.0, -2.5], [3.0, 3.0, -2.5], [-3.0, 3.0, -2.5]]
    wuv = [[0.0, -0.2], [1.5, -0.2], [1.5, 1.2], [0.0, 1.2]]
    s.meshes["wall.obj"] = (wv, [[0, 1, 2], [0, 2, 3]], None, wuv)
    if sphere_subd > 0:
        v, f = icosahedron()
        s.meshes["sphere.obj"] = (v, f, v.copy())
    else:
        s.meshes["sphere.obj"] = icosphere(2)
    v, f = box_mesh([-0.5, 0.0, -0.5], [0.5, 1.0, 0.5])
    s.meshes["box.obj"] = (v, f, None)
    s.meshes["glassball.obj"] = icosphere(2)
    mv, mf = quad([-0.6, 0.0, 0.0], [0.6, 0.0, 0.0], [0.6, 1.6, 0.0], [-0.6, 1.6, 0.0])
    s.meshes["mirror.obj"] = (mv, mf, None)

    checker = np.ones((16, 16, 4), np.float32)
    ij = np.add.outer(np.arange(16), np.arange(16)) % 2
    checker[..., :3] = np.where(ij[..., None] == 1, [0.8, 0.78, 0.7], [0.15, 0.2, 0.3])
    checker[0, 0, :3] = [0.9, 0.1, 0.1]  # marks the texel at s = 0, t = 0
    s.textures["checker.exr"] = (checker, True)
    s.textures["wallcolor.exr"] = (_tex_noise(rng, 32, 32, [0.2, 0.3, 0.5], [0.7, 0.6, 0.4]), True)
    s.textures["roughmap.exr"] = (_tex_noise(rng, 32, 32, [0.08, 0.0, 0.0], [0.7, 0.0, 0.0]), True)
    bumps = _tex_noise(rng, 16, 16, [-0.3, -0.3, 1.0], [0.3, 0.3, 1.0])
    s.textures["bumps.exr"] = (bumps, True)
    s.textures["tricol.exr"] = (_tex_noise(rng, 24, 24, [0.1, 0.2, 0.1], [0.9, 0.5, 0.8]), True)
    s.textures["trirefl.exr"] = (_tex_noise(rng, 8, 8, [0.5, 0.4, 0.2], [0.95, 0.8, 0.6]), True)
    s.textures["trirough.exr"] = (_tex_noise(rng, 16, 16, [0.05, 0.0, 0.0], [0.6, 0.0, 0.0]), True)
    s.textures["mixmap.exr"] = (_tex_noise(rng, 4, 4, [0.2, 0.0, 0.0], [0.8, 0.0, 0.0]), True)
    # lat-long sky: vertical gradient, a bright sun, FP32 storage
    H, W = 32, 64
    sky = np.ones((H, W, 4), np.float32)
    t = (np.arange(H) + 0.5) / H
    sky[..., 0] = (0.3 + 0.5 * t)[:, None]
    sky[..., 1] = (0.45 + 0.3 * t)[:, None]
    sky[..., 2] = (0.9 - 0.4 * t)[:, None]
    sky[8:11, 40:44, :3] = [40.0, 36.0, 30.0]
    sky[...] += np.float32(1.0 / 3.0) * 1e-3  # values that FP16 rounds
    s.textures["sky.exr"] = (sky, False)

    s.shaders["checker"] = {"type": "texture2DShader", "attributetype": "color", "texturename": "textures/checker.exr"}
    s.shaders["wallcolor"] = {"type": "texture2DShader", "attributetype": "color",
                              "texturename": "textures/wallcolor.exr"}
    s.shaders["roughmap"] = {"type": "texture2DShader", "attributetype": "float",
                             "texturename": "textures/roughmap.exr"}
    s.shaders["bumps"] = {"type": "texture2DShader", "attributetype": "color", "texturename": "textures/bumps.exr"}
    s.shaders["tricol"] = {"type": "triplanarTextureShader", "attributetype": "color",
                           "texturename": "textures/tricol.exr", "scale": 0.35}
    s.shaders["trirefl"] = {"type": "triplanarTextureShader", "attributetype": "color",
                            "texturename": "textures/trirefl.exr", "scale": 0.5}
    s.shaders["trirough"] = {"type": "triplanarTextureShader", "attributetype": "float",
                             "texturename": "textures/trirough.exr", "scale": 0.4}
    s.shaders["mixmap"] = {"type": "texture2DShader", "attributetype": "float", "texturename": "textures/mixmap.exr"}

    s.materials["floor"] = {"type": "matteMaterial", "color": "checker"}
    s.materials["wall"] = {"type": "kelemenMaterial", "diffusecolor": "wallcolor", "speccolor": [0.9, 0.9, 0.9],
                           "reflectance": 0.1, "roughness": "roughmap", "normalmap": "bumps"}
    s.materials["marble"] = {"type": "matteMaterial", "color": "tricol"}
    s.materials["brushed"] = {"type": "metalMaterial", "reflectance": "trirefl", "roughnessA": "trirough",
                              "roughnessB": 0.3, "mix": "mixmap", "gain": [1.0, 0.95, 0.9],
                              "normalmap": [0.1, 0.05, 1.0], "distribution": "ggx"}
    s.objects["floor"] = _geo("floor", "floor.obj", "floor")
    s.objects["wall"] = _geo("wall", "wall.obj", "wall")
    s.objects["sphere"] = _geo("sphere", "sphere.obj", "marble",
                               [{"translate": [-0.9, 0.7, -0.3]}, {"scale": [0.7, 0.7, 0.7]}],
                               subdlevels=sphere_subd)
    s.objects["box"] = _geo("box", "box.obj", "brushed",
                            [{"translate": [0.9, 0.0, 0.2]}, {"rotateY": 33.0}, {"scale": [0.8, 1.1, 0.8]}])
    s.materials["glass"] = {"type": "glassMaterial", "specColor": [0.95, 0.95, 0.95], "reflectance": 0.08,
                            "ior": 1.5, "transmitColor": [0.9, 0.97, 0.92]}
    s.materials["mirror"] = {"type": "glassMaterial", "specColor": [0.9, 0.9, 0.85], "reflectance": 1.0,
                             "ior": 1.0, "transmitColor": [0.0, 0.0, 0.0]}
    s.objects["glassball"] = _geo("glassball", "glassball.obj", "glass",
                                  [{"translate": [0.05, 0.42, 1.1]}, {"scale": [0.42, 0.42, 0.42]}])
    s.objects["mirror"] = _geo("mirror", "mirror.obj", "mirror",
                               [{"translate": [2.1, 0.0, -1.2]}, {"rotateY": -35.0}])
    s.objects["light_key"] = _square_light("light_key", [10.0, 9.0, 8.0], 0.5, 0.5,
                                           [{"translate": [1.5, 3.2, 1.0]}, {"rotateX": -70.0}])
    s.objects["sky"] = {"type": "envlight", "name": "sky", "exposure": 0.5, "color": [1.0, 1.0, 1.0],
                        "envmap": "textures/sky.exr", "transforms": [{"rotateY": 90.0}]}
    s.objects["sky_fill"] = {"type": "envlight", "name": "sky_fill", "exposure": -2.0, "color": [0.3, 0.3, 0.35],
                             "envmap": "", "transforms": []}
    return s


def batch_jobs(out_dir: str, n_jobs: int = 64, seed: int = 3, max_spp: int = 0) -> list:
    """Writes the batched workload's jobs; returns their scene file paths.
    ``max_spp`` > 0 caps the sample count (for tests)."""
    paths = []
    for k in range(n_jobs):
        spec = batch_job(seed, k)
        if max_spp > 0:
            spec.options["pixelsamples"] = min(spec.options["pixelsamples"], max_spp)
        paths.append(spec.write(os.path.join(out_dir, "job%03d" % k)))
    return paths


def write_config(name: str, out_dir: str, **kw) -> str:
    """Writes a named workload; returns the scene file path."""
    if name in ("box_room", "c1"):
        spec = box_room(**kw)
    elif name in ("box_room_kelemen", "c4"):
        kw.setdefault("res", 512)
        kw.setdefault("spp", 32)
        kw.setdefault("maxdepth", 6)
        kw.setdefault("mindepth", 6)
        spec = box_room(kelemen=True, **kw)
    elif name in ("displaced_spheres", "c2"):
        spec = displaced_spheres(**kw)
    elif name == "textured_stage":
        spec = textured_stage(**kw)
    elif name in ("large_scene", "c5"):
        spec, _ = large_scene(**kw)
    else:
        raise ValueError("unknown workload %r" % name)
    return spec.write(out_dir)
