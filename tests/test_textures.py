"""Textures, shaders and environment maps on the CPU: the oracle's FP16
rounding against numpy's binary16, the scene loader (texels, uvs, transforms,
shader bindings, error cases), and the oracle's texture / triplanar / env-map
routines against closed forms that fix the index conventions (texel centres
map back to their texel, single-axis projections read one texel, the env map's
sampling and eval agree).
"""
import math
import os
import struct

import numpy as np
import pytest

from conftest import host_renderer


# ---------------------------------------------------------------------------
# helpers
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def stage(tmp_path_factory):
    from aurora_amd import scenes
    d = str(tmp_path_factory.mktemp("stage"))
    spec = scenes.textured_stage(res=48, spp=2, maxdepth=3, mindepth=3)
    path = spec.write(d)
    r = host_renderer()
    r.load(path, os.path.join(d, "config.json"))
    e = r.export()
    r.close()
    return spec, d, e


def _tex_index(e, spec, name):
    """Texture index of a file, by matching the exported texel block."""
    want = spec.textures[name][0].astype(np.float16).astype(np.float32)
    for k, t in enumerate(e["textures"]):
        blk = e["texels"][4 * t.offset:4 * (t.offset + t.width * t.height)].reshape(t.height, t.width, 4)
        if blk.shape == want.shape and np.array_equal(blk, want):
            return k
    raise KeyError(name)


def _write_exr_y(path, y):
    """Minimal uncompressed EXR with a single HALF luminance channel Y."""
    h, w = y.shape

    def attr(name, typ, payload):
        return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(payload)) + payload
    hdr = attr("channels", "chlist", b"Y\0" + struct.pack("<iiii", 1, 0, 1, 1) + b"\0")
    hdr += attr("compression", "compression", b"\0")
    hdr += attr("dataWindow", "box2i", struct.pack("<iiii", 0, 0, w - 1, h - 1))
    hdr += attr("displayWindow", "box2i", struct.pack("<iiii", 0, 0, w - 1, h - 1))
    hdr += attr("lineOrder", "lineOrder", b"\0")
    hdr += attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
    hdr += attr("screenWindowCenter", "v2f", struct.pack("<ff", 0.0, 0.0))
    hdr += attr("screenWindowWidth", "float", struct.pack("<f", 1.0))
    head = struct.pack("<ii", 20000630, 2) + hdr + b"\0"
    table_pos = len(head)
    chunks, pos = [], table_pos + 8 * h
    for row in range(h):
        data = y[row].astype(np.float16).tobytes()
        chunks.append(struct.pack("<ii", row, len(data)) + data)
    offs = []
    for c in chunks:
        offs.append(pos)
        pos += len(c)
    with open(path, "wb") as fh:
        fh.write(head + struct.pack("<%dQ" % h, *offs) + b"".join(chunks))


def _oracle_scene(oracle, textures, shaders=(), mats=(), lights=(), camera=None):
    """An oracle scene without geometry that holds textures, shaders, materials
    (for their c2o) and lights."""
    from aurora_amd import _lib as L
    offs, pool = [], []
    off = 0
    tex = []
    for a in textures:
        a = np.asarray(a, np.float32)
        t = L.Texture(a.shape[1], a.shape[0], off)
        tex.append(t)
        pool.append(a.reshape(-1))
        off += a.shape[0] * a.shape[1]
    exp = {"verts": np.zeros((0, 3), np.float32), "tris": np.zeros((0, 3), np.int32),
           "normals": np.zeros((0, 9), np.float32), "attr": np.zeros(0, np.int32), "materials": list(mats),
           "lights": list(lights), "fov": 0.5, "width": 4, "height": 4, "maxdepth": 2, "mindepth": 2,
           "textures": tex, "texels": np.concatenate(pool) if pool else np.zeros(0, np.float32),
           "shaders": list(shaders)}
    if camera is not None:
        exp["camera"] = camera
    return oracle.Scene(exp)


# ---------------------------------------------------------------------------
# FP16 rounding (the texels the reference holds are Imf::Rgba halves)
# ---------------------------------------------------------------------------
def test_oracle_half_rounding_matches_numpy(oracle):
    rng = np.random.default_rng(3)
    v = np.concatenate([rng.normal(0, 1, 4000) * 10.0 ** rng.integers(-9, 6, 4000),
                        [0.0, 65504.0, 65519.0, 65520.0, 7e4, 5.96e-8, 2.98e-8, 2.99e-8, 6.1e-5,
                         1 + 2 ** -11, 1 + 3 * 2 ** -11, 2 ** -14 - 2 ** -25]]).astype(np.float32)
    with np.errstate(over="ignore"):
        want = v.astype(np.float16).astype(np.float32)
    got = np.array([oracle.half_round(float(x)) for x in v], np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


# ---------------------------------------------------------------------------
# scene loading
# ---------------------------------------------------------------------------
def test_loader_texels_are_fp16_of_the_files(stage):
    spec, _, e = stage
    assert len(e["textures"]) == len(spec.textures)
    seen = set()
    for name, (rgba, half) in spec.textures.items():
        k = _tex_index(e, spec, name)  # raises unless some texture equals the FP16 rounding
        seen.add(k)
    assert len(seen) == len(spec.textures)
    # the FP32 sky really needed rounding
    sky = spec.textures["sky.exr"][0]
    assert not np.array_equal(sky, sky.astype(np.float16).astype(np.float32))


def test_loader_uvs_and_transforms(stage):
    from aurora_amd import scenes
    spec, _, e = stage
    attr_of = {}
    names = sorted(k for k, o in spec.objects.items() if o["type"] == "geometry")
    for i, n in enumerate(names):
        attr_of[n] = i
    # floor: per-corner uvs equal the OBJ's vt at the same corners
    fa = attr_of["floor"]
    tris = np.nonzero(e["attr"] == fa)[0]
    fv, ff, _, fuv = spec.meshes["floor.obj"]
    got = e["uvs"][tris].reshape(-1, 2)
    want = np.asarray(fuv, np.float32)[np.asarray(ff).reshape(-1)]
    assert np.array_equal(got, want)
    # objects without vt: zero uvs
    assert not np.any(e["uvs"][e["attr"] == attr_of["box"]])
    # c2o maps the camera-space vertices back to the OBJ's object space
    for n in ("floor", "box", "sphere"):
        a = attr_of[n]
        c2o = np.array(e["materials"][a].c2o, np.float64).reshape(4, 4)
        tv = e["tris"][e["attr"] == a]
        P = e["verts"][tv.reshape(-1)].astype(np.float64)
        obj = (c2o[:3, :3] @ P.T).T + c2o[:3, 3]
        mesh = spec.meshes[n + ".obj"]
        objv = np.asarray(mesh[0], np.float64)[np.asarray(mesh[1]).reshape(-1)]
        assert np.max(np.abs(obj - objv)) < 2e-5, n
    # camera transform: inverse pair, camera origin at its translation
    cam = e["camera"].astype(np.float64)
    M, Mi = cam[:16].reshape(4, 4), cam[16:].reshape(4, 4)
    assert np.max(np.abs(M @ Mi - np.eye(4))) < 1e-6
    assert np.allclose(M[:3, 3], [0.2, 1.3, 4.6], atol=1e-6)
    # shader bindings by slot
    mats = {n: e["materials"][attr_of[n]] for n in names}
    sh = e["shaders"]
    floor_s = mats["floor"].shader[0]
    assert sh[floor_s].type == 0 and sh[floor_s].kind == 0
    wall = mats["wall"]
    assert sh[wall.shader[3]].kind == 1 and sh[wall.shader[6]].type == 0
    box = mats["box"]
    assert sh[box.shader[3]].type == 1 and sh[box.shader[3]].kind == 1 and abs(sh[box.shader[3]].scale - 0.4) < 1e-7
    assert sh[box.shader[6]].type == 2 and list(sh[box.shader[6]].value) == pytest.approx([0.1, 0.05, 1.0])
    assert box.shader[4] == -1 and abs(box.roughnessB - 0.3) < 1e-7
    env = [l for l in e["lights"] if l.type == 1]
    assert sorted(l.envmap >= 0 for l in env) == [False, True]
    del scenes


def test_luminance_only_exr_reads_grey(tmp_path):
    from aurora_amd import exr
    y = np.linspace(0.0, 3.0, 6 * 5, dtype=np.float32).reshape(6, 5)
    p = str(tmp_path / "y.exr")
    _write_exr_y(p, y)
    img = exr.read_exr(p)
    yh = y.astype(np.float16).astype(np.float32)
    for c in range(3):
        assert np.array_equal(img[..., c], yh)
    assert np.all(img[..., 3] == 1.0)


@pytest.mark.parametrize("edit,msg", [
    (lambda s: s.materials["floor"].update(color="nope"), "unknown shader"),
    (lambda s: s.materials["floor"].update(color="roughmap"), "float shader"),
    (lambda s: s.materials["wall"].update(roughness="wallcolor"), "color shader"),
    (lambda s: s.shaders["checker"].update(texturename="textures/missing.exr"), "texture"),
    (lambda s: s.shaders["checker"].update(type="noiseShader"), "unknown shader type"),
    (lambda s: s.shaders["checker"].update(attributetype="vector"), "unknown attribute type"),
])
def test_loader_rejects_bad_shaders(tmp_path, edit, msg):
    from aurora_amd import scenes
    from aurora_amd._lib import AuroraError
    spec = scenes.textured_stage(res=16, spp=1, maxdepth=2, mindepth=2)
    edit(spec)
    path = spec.write(str(tmp_path))
    r = host_renderer()
    with pytest.raises(AuroraError) as ei:
        r.load(path, os.path.join(str(tmp_path), "config.json"))
    r.close()
    assert msg in str(ei.value)


def test_loader_accepts_uv_textures_on_subdivided_meshes(tmp_path):
    """Subdivided meshes carry their texture coordinates (Loop subdivision on
    the GPU, loopSubdivMesh.cpp:174-191), so a uv-mapped texture on one loads."""
    from aurora_amd import scenes
    spec = scenes.textured_stage(res=16, spp=1, maxdepth=2, mindepth=2)
    spec.objects["floor"]["subdlevels"] = 1
    path = spec.write(str(tmp_path))
    r = host_renderer()
    r.load(path, os.path.join(str(tmp_path), "config.json"))
    r.close()


# ---------------------------------------------------------------------------
# oracle routines against closed forms
# ---------------------------------------------------------------------------
def _distinct_texture(h, w):
    a = np.ones((h, w, 4), np.float32)
    i, j = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    a[..., 0] = (i * w + j + 1) / 64.0
    a[..., 1] = i / 8.0
    a[..., 2] = j / 8.0
    return a.astype(np.float16).astype(np.float32)


def test_texture2d_shader_known_texels(oracle):
    from aurora_amd import _lib as L
    tex = _distinct_texture(8, 16)
    sh = L.Shader(0, 0, 0, 0.0)
    sc = _oracle_scene(oracle, [tex], [sh])
    P = Ng = [0.0, 1.0, 0.0]
    # texture2DShader reads (u - floor(u), 1 - (v - floor(v))): row 0 is at v -> 1
    cases = [((0.01, 0.99), (0, 0)), ((0.01, 0.01), (7, 0)), ((0.99, 0.99), (0, 15)),
             ((-0.99, 1.99), (0, 0)), ((3.01, -2.01), (0, 0)), ((0.0, 0.0), (7, 0)),
             ((0.5 + 1 / 32, 0.5 - 1 / 16), (4, 8)), ((1.0, 1.0), (7, 0))]
    for (u, v), (row, col) in cases:
        got = sc.shader_eval(0, P, Ng, u, v, 0)
        assert np.array_equal(got, tex[row, col, :3]), ((u, v), (row, col))


def test_triplanar_projections(oracle):
    from aurora_amd import _lib as L
    tex = _distinct_texture(16, 16)
    const = np.ones((4, 4, 4), np.float32)
    const[..., :3] = [0.3, 0.55, 0.8]
    shaders = [L.Shader(1, 0, 0, 1.0), L.Shader(1, 1, 0, 1.0), L.Shader(1, 0, 1, 0.7), L.Shader(1, 1, 1, 0.7)]
    m = L.Material()
    sc = _oracle_scene(oracle, [tex, const], shaders, [m])
    rng = np.random.default_rng(4)
    # a constant texture: the normalised weights reproduce the constant
    for _ in range(200):
        P = rng.uniform(-3, 3, 3)
        N = rng.normal(size=3)
        N /= np.linalg.norm(N)
        assert np.allclose(sc.shader_eval(2, P, N, 0, 0, 0), const[0, 0, :3], rtol=3e-7)
        assert abs(sc.shader_eval(3, P, N, 0, 0, 0)[0] - const[0, 0, 0]) < 3e-7
    # one projection: Ng along y reads texel (s, t) = (|fmod(z, 1)|, |fmod(x, 1)|)
    for _ in range(200):
        P = rng.uniform(-3, 3, 3).astype(np.float32)
        s, t = abs(math.fmod(float(P[2]), 1.0)), abs(math.fmod(float(P[0]), 1.0))
        want = tex[min(int(t * 16), 15), min(int(s * 16), 15), :3]
        assert np.array_equal(sc.shader_eval(0, P, [0.0, 1.0, 0.0], 0, 0, 0), want)
        # along x: (|fmod(y, 1)|, |fmod(z, 1)|); along z: (|fmod(x, 1)|, |fmod(y, 1)|)
        s, t = abs(math.fmod(float(P[1]), 1.0)), abs(math.fmod(float(P[2]), 1.0))
        want = tex[min(int(t * 16), 15), min(int(s * 16), 15), :3]
        assert np.array_equal(sc.shader_eval(0, P, [-1.0, 0.0, 0.0], 0, 0, 0), want)
        s, t = abs(math.fmod(float(P[0]), 1.0)), abs(math.fmod(float(P[1]), 1.0))
        want = tex[min(int(t * 16), 15), min(int(s * 16), 15), :3]
        assert np.array_equal(sc.shader_eval(0, P, [0.0, 0.0, 1.0], 0, 0, 0), want)
    # thresholds: |N_x| = 0.22 drops the x projection for colour (0.25) but not
    # for float (0.2)
    N = np.array([0.22, math.sqrt(1 - 0.22 ** 2), 0.0], np.float32)
    P = np.array([0.3, 0.6, 0.1], np.float32)
    y_only = tex[int(0.3 * 16), int(0.1 * 16), :3]
    assert np.array_equal(sc.shader_eval(0, P, N, 0, 0, 0), y_only)
    f = sc.shader_eval(1, P, N, 0, 0, 0)[0]
    x_proj = tex[int(0.1 * 16), int(0.6 * 16), 0]
    assert min(y_only[0], x_proj) < f < max(y_only[0], x_proj)


def _sky(h=16, w=32):
    a = _distinct_texture(h, w)
    a[..., :3] += 0.05
    return a.astype(np.float16).astype(np.float32)


def _env_light(envmap=0, exposure=0.5):
    from aurora_amd import _lib as L
    l = L.Light()
    l.type, l.envmap, l.exposure = 1, envmap, exposure
    l.color[0] = l.color[1] = l.color[2] = 1.0
    return l


def _camera(rx=-9.0, ry=4.0):
    """A rotation camera (row-major camera-to-world and its transpose)."""
    a, b = math.radians(rx), math.radians(ry)
    Ry = np.array([[math.cos(b), 0, math.sin(b)], [0, 1, 0], [-math.sin(b), 0, math.cos(b)]])
    Rx = np.array([[1, 0, 0], [0, math.cos(a), -math.sin(a)], [0, math.sin(a), math.cos(a)]])
    R = (Ry @ Rx).astype(np.float32)
    M = np.eye(4, dtype=np.float32)
    M[:3, :3] = R
    Mi = np.eye(4, dtype=np.float32)
    Mi[:3, :3] = R.T
    return np.concatenate([M.reshape(-1), Mi.reshape(-1)])


def test_env_map_eval_texel_centres(oracle):
    sky = _sky()
    H, W = sky.shape[:2]
    cam = _camera()
    sc = _oracle_scene(oracle, [sky], lights=[_env_light()], camera=cam)
    M = cam[:16].reshape(4, 4).astype(np.float64)
    scale = 2.0 ** 0.5
    for i in range(H):
        for j in range(W):
            u, v = (j + 0.5) / W, (i + 0.5) / H
            y = 1 - 2 * v
            r = math.sqrt(1 - y * y)
            world = np.array([r * math.cos(2 * math.pi * u), y, r * math.sin(2 * math.pi * u)])
            d = M[:3, :3].T @ world  # world -> camera
            got = sc.env_eval(0, d.astype(np.float32))
            want = (sky[i, j, :3].astype(np.float64) * scale).astype(np.float32)
            assert np.array_equal(got, want), (i, j)


def test_env_map_sampling_consistency(oracle):
    sky = _sky()
    H, W = sky.shape[:2]
    cam = _camera()
    l = _env_light()
    sc = _oracle_scene(oracle, [sky], lights=[l], camera=cam)
    light = sc.lights[0]
    # initPdf: alpha = half((r+g+b)/3), normalised by the mean
    alpha = np.array([oracle.half_round(float((p[0] + p[1] + p[2]) / np.float32(3))) for p in sky.reshape(-1, 4)])
    rng = np.random.default_rng(6)
    P, N = [0.0, 0.0, 0.0], [0.0, 1.0, 0.0]
    pdfs = []
    for _ in range(300):
        i, j = rng.integers(0, H), rng.integers(0, W)
        x = (j + rng.uniform(0.2, 0.8)) / W
        y = (i + rng.uniform(0.2, 0.8)) / H
        d, tmax, pdf, col = oracle.light_sample(light, P, N, float(x), float(y), scene=sc)
        assert tmax == 1e6 and abs(np.linalg.norm(d) - 1) < 1e-6
        # the sampled direction looks up the texel the uniforms point at
        assert np.array_equal(col, sc.env_eval(0, d))
        assert np.array_equal(col, (sky[i, j, :3].astype(np.float64) * 2 ** 0.5).astype(np.float32))
        ratio = pdf * 4 * math.pi / alpha[i * W + j]
        pdfs.append(ratio)
        # MIS pdf of the same direction, hemisphere domain: twice the sphere one
        lp, _ = oracle.light_pdf(light, P, d, scene=sc)
        assert abs(lp / pdf - 2.0) < 1e-6
    # pdf = alpha / norm / (4 pi), with norm the mean of the unrounded means
    norm = np.float32(np.mean(sky[..., :3].astype(np.float64).sum(-1) / 3))
    assert np.allclose(pdfs, 1.0 / norm, rtol=1e-5)


def test_oracle_renders_textured_stage(stage, oracle):
    spec, _, e = stage
    s = oracle.Scene(e)
    acc, st = s.render(3, 0, 1)
    assert np.all(np.isfinite(acc))
    assert st["paths"] == e["width"] * e["height"]
    assert np.all(acc[..., 3] == 1.0)  # every camera ray sees geometry or the environment


# ---------------------------------------------------------------------------
# the FP32 albedo texture (aur_set_albedo_texture's restatement in the oracle)
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def terrain(tmp_path_factory):
    """Two terrain patches of the large differentiable scene with its uv
    texture shader (no spheres: a host-only context flattens it)."""
    from aurora_amd import scenes
    d = str(tmp_path_factory.mktemp("terrain"))
    spec, truth = scenes.large_scene(width=48, height=27, spp=2, maxdepth=3, mindepth=3, n_spheres=0, levels=0,
                                     patches=(2, 1), patch_quads=12, patch_size=6.0, tex_res=16, seed=7)
    path = spec.write(d)
    r = host_renderer()
    r.load(path, os.path.join(d, "config.json"))
    e = r.export()
    r.close()
    attrs = [spec.attr_index("terrain00"), spec.attr_index("terrain10")]
    return e, truth.astype(np.float16).astype(np.float32), attrs


def test_albedo_texture_reads_what_the_texture_shader_reads(oracle, terrain):
    """With texel values FP16 represents exactly, the FP32 albedo texture and
    the scene's own texture2DShader (texture2DShader.h:28-31 over Texture2D's
    FP16 texels) give the same image bit for bit; a scaled texture darkens it."""
    e, truth, attrs = terrain
    ref, rst = oracle.Scene(e).render(11, 0, 2, threads=8)
    got, gst = oracle.Scene(e, ptex=truth, ptex_attrs=attrs).render(11, 0, 2, threads=8)
    assert ref[..., :3].max() > 0 and np.array_equal(got, ref) and rst == gst
    half, _ = oracle.Scene(e, ptex=truth * 0.5, ptex_attrs=attrs).render(11, 0, 2, threads=8)
    assert np.mean(half[..., :3] < ref[..., :3]) > 0.3


def test_albedo_texture_gradient_matches_central_differences(oracle, terrain):
    """The image is a polynomial in the texels with fixed random numbers and
    the largest texel gradients of L = sum w . I match central differences
    (h = 1e-3) to 1e-2 relative."""
    e, truth, attrs = terrain
    rng = np.random.default_rng(2)
    H, W = e["height"], e["width"]
    w = rng.uniform(-1, 1, size=(H, W, 3))
    dl = np.zeros((H, W, 4), np.float32)
    dl[..., :3] = w
    s = oracle.Scene(e, ptex=truth, ptex_attrs=attrs)
    out = s.adjoint(13, 0, 4, dl, threads=8)
    gt = out[3]
    assert gt.shape == truth.shape and np.count_nonzero(gt) > 20

    def loss(tex):
        acc, _ = oracle.Scene(e, ptex=tex, ptex_attrs=attrs).render(13, 0, 4, threads=8)
        return float(np.sum((acc[..., :3] / 4).astype(np.float64) * w))

    h = 1e-3
    for idx in np.argsort(-np.abs(gt.ravel()))[:6]:
        y, x, c = np.unravel_index(int(idx), gt.shape)
        tp, tm = truth.copy(), truth.copy()
        tp[y, x, c] += h
        tm[y, x, c] -= h
        fd = (loss(tp) - loss(tm)) / (2 * h)
        assert gt[y, x, c] == pytest.approx(fd, rel=1e-2, abs=1e-3 * np.abs(gt).max()), (y, x, c)
