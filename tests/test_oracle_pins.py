"""Pins the CPU oracle against closed forms, invariants and brute force.

Each test is chosen so that a plausible slip in the restatement -- a dropped
term, a wrong sign or index, a transposed operand -- makes it fail.  The
expected values come from the mathematics (solid angles, normalisation
integrals, smooth-surface limits, hand-derived stencils), not from re-typing
the oracle's own formulas.
"""
import math

import numpy as np
import pytest


# ---------------------------------------------------------------------------
# Moller-Trumbore (renderableTriangle.cpp:33-66)
# ---------------------------------------------------------------------------
def test_triangle_hit_barycentrics(oracle):
    p, e1, e2 = [0, 0, 0], [1, 0, 0], [0, 1, 0]
    hit, t, b1, b2 = oracle.intersect_tri(p, e1, e2, [0.2, 0.3, 1.0], [0, 0, -1], 0.0, 10.0)
    assert hit and t == pytest.approx(1.0) and b1 == pytest.approx(0.2) and b2 == pytest.approx(0.3)
    # b1 weights the e1 vertex, b2 the e2 vertex: a transposed operand swaps them
    hit, t, b1, b2 = oracle.intersect_tri(p, e1, e2, [0.7, 0.1, 2.0], [0, 0, -1], 0.0, 10.0)
    assert hit and t == pytest.approx(2.0) and (b1, b2) == (pytest.approx(0.7), pytest.approx(0.1))


def test_triangle_misses_and_inclusive_bounds(oracle):
    p, e1, e2 = [0, 0, 0], [1, 0, 0], [0, 1, 0]
    assert not oracle.intersect_tri(p, e1, e2, [0.8, 0.8, 1.0], [0, 0, -1], 0.0, 10.0)[0]   # outside
    assert not oracle.intersect_tri(p, e1, e2, [0.2, 0.2, 1.0], [1, 0, 0], 0.0, 10.0)[0]    # parallel
    assert not oracle.intersect_tri(p, e1, e2, [0.2, 0.2, 1.0], [0, 0, -1], 0.0, 0.5)[0]    # beyond tmax
    assert not oracle.intersect_tri(p, e1, e2, [0.2, 0.2, 1.0], [0, 0, -1], 1.5, 9.0)[0]    # before tmin
    assert oracle.intersect_tri(p, e1, e2, [0.5, 0.5, 1.0], [0, 0, -1], 0.0, 10.0)[0]       # edge b1+b2 == 1
    assert oracle.intersect_tri(p, e1, e2, [0.2, 0.2, 1.0], [0, 0, -1], 0.0, 1.0)[0]        # t == tmax
    hit, t, _, _ = oracle.intersect_tri(p, e1, e2, [0.2, 0.2, -1.0], [0, 0, 1], 0.0, 10.0)   # back side
    assert hit and t == pytest.approx(1.0)


# ---------------------------------------------------------------------------
# BSDFs (lambert.h:44-80, cookTorrance.cpp:40-144)
# ---------------------------------------------------------------------------
N = np.array([0, 0, 1], np.float32)


def _grid(n):
    g = (np.arange(n) + 0.5) / n
    return np.meshgrid(g, g, indexing="ij")


def test_lambert_white_furnace(oracle):
    albedo = [0.3, 0.6, 0.9]
    u1, u2 = _grid(64)
    acc = np.zeros(3)
    for a, b in zip(u1.ravel(), u2.ravel()):
        d, pdf, col = oracle.bsdf_sample(0, albedo, 0.0, 0, [0, 0, 1], N, float(a), float(b))
        assert d[2] >= 0.0
        acc += col * d[2] / pdf
    np.testing.assert_allclose(acc / u1.size, albedo, rtol=2e-3)
    f, pdf = oracle.bsdf_eval(0, albedo, 0.0, 0, [0, 0, 1], N, [0.6, 0, 0.8])
    np.testing.assert_allclose(f, np.array(albedo) / math.pi, rtol=1e-6)
    assert pdf == pytest.approx(1.0 / (2 * math.pi), rel=1e-6)
    f, _ = oracle.bsdf_eval(0, albedo, 0.0, 0, [0, 0, 1], N, [0.6, 0, -0.8])
    assert np.all(f == 0)


def _sphere_quadrature(n_theta=720, n_phi=360):
    th = (np.arange(n_theta) + 0.5) / n_theta * math.pi
    ph = (np.arange(n_phi) + 0.5) / n_phi * 2 * math.pi
    w = (math.pi / n_theta) * (2 * math.pi / n_phi) * np.sin(th)
    return th, ph, w


@pytest.mark.parametrize("dist", [0, 1])
@pytest.mark.parametrize("rough", [0.25, 0.5])
def test_microfacet_pdf_normalisation(oracle, dist, rough):
    """With V = N every half-vector has v.h > 0, so pdf_L integrates to 1 over the sphere."""
    th, ph, w = _sphere_quadrature(360, 90)
    total = 0.0
    for i, t in enumerate(th):
        for p in ph:
            L = [math.sin(t) * math.cos(p), math.sin(t) * math.sin(p), math.cos(t)]
            _, pdf = oracle.bsdf_eval(1, [1, 1, 1], rough, dist, [0, 0, 1], N, L)
            if pdf > 0 and math.isfinite(pdf):
                total += pdf * w[i]
    assert total == pytest.approx(1.0, abs=1.5e-2)


@pytest.mark.parametrize("dist", [0, 1])
def test_microfacet_sample_consistent_with_pdf_and_eval(oracle, dist):
    """getSample's pdf/colour agree with pdf()/evalSampleWorld() for the sampled L."""
    V = np.array([0.3, -0.2, 0.9327379], np.float32)
    V /= np.linalg.norm(V)
    u1, u2 = _grid(24)
    checked = 0
    for a, b in zip(u1.ravel(), u2.ravel()):
        d, spdf, scol = oracle.bsdf_sample(1, [0.8, 0.5, 0.2], 0.35, dist, V, N, float(a), float(b))
        if not (1e-3 < spdf < 1e4) or d[2] <= 0.01:
            continue
        f, pdf = oracle.bsdf_eval(1, [0.8, 0.5, 0.2], 0.35, dist, V, N, d)
        h = (d + V) / np.linalg.norm(d + V)
        if h[2] > 0.999 or np.dot(V, h) <= 0.01:
            continue
        assert pdf == pytest.approx(spdf, rel=2e-4)
        np.testing.assert_allclose(f, scol, rtol=2e-4)
        checked += 1
    assert checked > 200


@pytest.mark.parametrize("dist", [0, 1])
def test_microfacet_sampling_moments(oracle, dist):
    """E[cos theta_h] of getSample's half-vectors equals the D(h)cos(h) quadrature."""
    rough = 0.4
    th, ph, w = _sphere_quadrature(2000, 1)
    num = den = 0.0
    for i, t in enumerate(th):
        if t >= math.pi / 2:
            continue
        L = [math.sin(2 * t), 0.0, math.cos(2 * t)]  # V = N: h at angle t reflects to L at 2t
        _, pdf = oracle.bsdf_eval(1, [1, 1, 1], rough, dist, [0, 0, 1], N, L)
        dens = pdf * 4 * math.cos(t)  # D(h)cos(h) from pdf_L = D cos / (4 v.h)
        num += math.cos(t) * dens * w[i] * 2 * math.pi
        den += dens * w[i] * 2 * math.pi
    expect = num / den
    u1, u2 = _grid(200)
    vals = []
    for a, b in zip(u1[:, 0], u2[:, 0]):
        d, _, _ = oracle.bsdf_sample(1, [1, 1, 1], rough, dist, [0, 0, 1], N, float(a), 0.25)
        h = (d + np.array([0, 0, 1])) / np.linalg.norm(d + np.array([0, 0, 1]))
        vals.append(h[2])
    assert np.mean(vals) == pytest.approx(expect, abs=3e-3)


# ---------------------------------------------------------------------------
# square light (squareLight.h:57-122)
# ---------------------------------------------------------------------------
def _facing_light(oracle, height=2.0, sx=0.5, sy=0.3):
    # object normal -z; rotateX(-90) makes it face -y; placed at y = height
    m = np.array([[1, 0, 0, 0], [0, 0, 1, height], [0, -1, 0, 0], [0, 0, 0, 1]], np.float32)
    return oracle.make_light(m.ravel(), sx, sy, [10.0, 5.0, 1.0], exposure=1.0)


def test_light_sample_solid_angle(oracle):
    """E[1/pdf] over uniform quad samples is the quad's solid angle."""
    light = _facing_light(oracle)
    P, Nn = [0.0, 0.0, 0.0], [0.0, 1.0, 0.0]
    u1, u2 = _grid(128)
    inv = []
    for a, b in zip(u1.ravel(), u2.ravel()):
        d, tmax, pdf, col = oracle.light_sample(light, P, Nn, float(a), float(b))
        inv.append(1.0 / pdf)
    a, b, h = 0.5, 0.3, 2.0
    omega = 4 * math.atan(a * b / (h * math.sqrt(h * h + a * a + b * b)))
    assert np.mean(inv) == pytest.approx(omega, rel=1e-3)
    # Le = color * 2^exposure
    np.testing.assert_allclose(col, [20.0, 10.0, 2.0], rtol=1e-6)
    # the shadow ray stops 10 * RAY_BIAS short of the light
    d, tmax, pdf, _ = oracle.light_sample(light, P, Nn, 0.5, 0.5)
    assert tmax == pytest.approx(2.0 - 0.005, abs=1e-6)
    np.testing.assert_allclose(d, [0, 1, 0], atol=1e-6)


def test_light_pdf_two_paths_agree(oracle):
    light = _facing_light(oracle)
    P = [0.1, 0.0, -0.05]
    d, tmax, spdf, _ = oracle.light_sample(light, P, [0, 1, 0], 0.3, 0.8)
    lpdf, hit_tmax = oracle.light_pdf(light, P, d)
    assert lpdf == pytest.approx(spdf, rel=1e-5)
    assert hit_tmax == pytest.approx(tmax, rel=1e-5)
    # below the light and facing away: invisible
    assert oracle.lib().orc_light_visible(light, oracle.fp(np.array([0, 0, 0], np.float32)),
                                          oracle.fp(np.array([0, -1, 0], np.float32))) == 0


# ---------------------------------------------------------------------------
# Kelemen albedo table (kelemenMaterial.cpp:113-223)
# ---------------------------------------------------------------------------
def test_kelemen_table_smooth_limit(oracle):
    """At roughness -> 0 the microfacet normal is N, so A(cos) = schlick(cos, F0)."""
    tab = oracle.kelemen_table(0.05, 0, n_samples=512).reshape(64, 32)
    cos = np.maximum(1e-4, np.arange(64) / 63.0)
    expect = 0.05 + (1 - cos) ** 5 * 0.95
    np.testing.assert_allclose(tab[1:, 0], expect[1:], rtol=1e-4)
    assert np.all(tab <= 1.0 + 1e-6) and np.all(tab >= 0)


def test_kelemen_lookup_golden(oracle):
    """Bilinear lookup with the reference's rB = max(rA + 1, 31) row."""
    i, j = np.meshgrid(np.arange(64), np.arange(32), indexing="ij")
    tab = (i + 100.0 * j).astype(np.float32).ravel()
    # cos 0.5 -> i = 31, mix 0.5; r 0.25 -> rA = 7, rB = 31, rmix = 0.75
    v = oracle.lib().orc_kelemen_albedo(oracle.fp(tab), 0.5, 0.25)
    assert v == pytest.approx(0.25 * (31.5 + 700) + 0.75 * (31.5 + 3100), rel=1e-6)


# ---------------------------------------------------------------------------
# camera (camera.cpp:34-62)
# ---------------------------------------------------------------------------
def test_camera_ray_closed_form(oracle, box_room_export):
    s = oracle.Scene(box_room_export)
    f = box_room_export["fov"]
    ray = oracle.ORay()
    oracle.lib().orc_camera_ray(s.s, 0, 0, 0.0, 0.0, ray)
    d = np.array(ray.d[:])
    np.testing.assert_allclose(d, np.array([-f, f, -1]) / math.sqrt(2 * f * f + 1), rtol=1e-6)
    oracle.lib().orc_camera_ray(s.s, 128, 128, 0.0, 0.0, ray)
    np.testing.assert_allclose(np.array(ray.d[:]), [0, 0, -1], atol=1e-7)
    assert ray.tmin == pytest.approx(0.1) and ray.o[:] == [0, 0, 0]


# ---------------------------------------------------------------------------
# integrator (integrator.cpp:33-54, 152-300) on minimal scenes
# ---------------------------------------------------------------------------
def _floor_light_scene(albedo=0.5, emission=(10.0, 10.0, 10.0), maxdepth=1, res=16):
    """Floor quad (camera space y = -1), one light 2 units above it, camera at origin looking down -z."""
    from aurora_amd import _lib as L
    verts = np.array([[-50, -1, 50], [50, -1, 50], [50, -1, -50], [-50, -1, -50]], np.float32)
    tris = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    normals = np.tile(np.array([0, 1, 0], np.float32), (2, 3))
    mats = [L.Material(), L.Material()]
    mats[0].type = 0
    mats[0].diffuse[:] = [albedo] * 3
    mats[1].type = 3
    mats[1].emission[:] = list(emission)
    light = L.Light()
    light.type = 0
    m = np.array([[1, 0, 0, 0], [0, 0, 1, 1.0], [0, -1, 0, -3.0], [0, 0, 0, 1]], np.float32)
    light.m[:] = list(m.ravel())
    light.sx, light.sy = 0.5, 0.5
    light.color[:] = list(emission)
    return {"verts": verts, "tris": tris, "normals": normals, "attr": np.array([0, 0], np.int32),
            "materials": mats, "lights": [light], "fov": 0.01, "width": res, "height": res,
            "maxdepth": maxdepth, "mindepth": maxdepth}


def test_camera_miss_without_environment_is_black(oracle):
    e = _floor_light_scene()
    s = oracle.Scene(e)
    n = 64
    out, cnt = s.render_paths(0, np.full(n, 8 * 16 + 8, np.int32), np.arange(n, dtype=np.int32))
    hit = s.trace(np.array([[0, 0, 0, 0.1, 0, 0, -1, 1e7]], np.float32))
    assert hit["prim"][0] < 0  # the centre ray runs parallel to the floor
    assert np.allclose(out, 0.0)
    assert cnt["ext_rays"] == n and cnt["shadow_rays"] == 0


def test_direct_lighting_quadrature_tilted_camera(oracle):
    """Camera ray hits the floor at P; Lo = rho/pi * Le * sum(cos_s cos_l / d^2 dA)."""
    e = _floor_light_scene(res=1)
    e["fov"] = 1e-4
    # tilt: rotate the scene geometry instead of the camera (camera space is fixed)
    ang = math.radians(-45.0)
    R = np.array([[1, 0, 0], [0, math.cos(ang), -math.sin(ang)], [0, math.sin(ang), math.cos(ang)]], np.float32)
    Rinv = R.T
    e["verts"] = (e["verts"] @ Rinv.T).astype(np.float32)
    e["normals"] = (e["normals"].reshape(-1, 3) @ Rinv.T).reshape(-1, 9).astype(np.float32)
    M = np.eye(4, dtype=np.float32)
    M[:3, :3] = Rinv
    lm = np.array(e["lights"][0].m[:], np.float32).reshape(4, 4)
    e["lights"][0].m[:] = list((M @ lm).ravel())
    s = oracle.Scene(e)
    hit = s.trace(np.array([[0, 0, 0, 0.1, 0, 0, -1, 1e7]], np.float32))
    assert hit["prim"][0] >= 0
    t = float(hit["t"][0])
    P = Rinv.T @ np.array([0, 0, -t])  # hit point back in the unrotated frame (y == -1)
    assert P[1] == pytest.approx(-1.0, abs=1e-5)
    n = 250 * 250
    g = (np.arange(250) + 0.5) / 250
    gx, gy = np.meshgrid(g, g, indexing="ij")
    slots = np.zeros((n, 11), np.float32)
    slots[:, 0] = slots[:, 1] = 0.5
    # vertex 0: slot 2 lobe, 3 light pick, 4-5 light sample (integrator.cpp:116, squareLight.h:63-64)
    slots[:, 4], slots[:, 5] = gx.ravel(), gy.ravel()
    out, cnt = s.render_paths(0, np.zeros(n, np.int32), np.zeros(n, np.int32), slots=slots)
    k = 400
    q = (np.arange(k) + 0.5) / k - 0.5
    X, Z = np.meshgrid(q, q - 3.0, indexing="ij")
    dx, dy, dz = X - P[0], 1.0 - P[1], Z - P[2]
    d2 = dx * dx + dy * dy + dz * dz
    cos_s = dy / np.sqrt(d2)
    cos_l = dy / np.sqrt(d2)
    E = np.sum(cos_s * cos_l / d2) / (k * k)  # area 1
    expect = 0.5 / math.pi * 10.0 * E
    assert out[:, 0].mean() == pytest.approx(expect, rel=2e-3)
    assert cnt["shadow_rays"] == n


@pytest.mark.parametrize("emission,expect", [(1000.0, 200.0), (1001.0, 1001.0 * 0.04), (150.0, 150.0)])
def test_firefly_clamp_on_emitter(oracle, emission, expect):
    """A camera ray that hits an emitter sees its emission, scaled by 0.2 while any channel > 200."""
    e = _floor_light_scene(emission=(emission,) * 3, res=1)
    e["fov"] = 1e-4
    # put the light quad in front of the camera, facing it
    e["lights"][0].m[:] = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, -2.0, 0, 0, 0, 1]
    e["verts"] = np.array([[-0.5, -0.5, -2.0], [-0.5, 0.5, -2.0], [0.5, 0.5, -2.0], [0.5, -0.5, -2.0]], np.float32)
    e["tris"] = np.array([[0, 1, 3], [2, 1, 3]], np.int32)
    e["attr"] = np.array([1, 1], np.int32)
    s = oracle.Scene(e)
    out, cnt = s.render_paths(0, np.zeros(4, np.int32), np.arange(4, dtype=np.int32))
    np.testing.assert_allclose(out[:, 0], expect, rtol=1e-6)
    assert cnt["shadow_rays"] == 0  # the constant BRDF of a light adds no direct lighting


def _emitter_in_view(emission):
    e = _floor_light_scene(emission=emission, res=1)
    e["fov"] = 1e-4
    e["lights"][0].m[:] = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, -2.0, 0, 0, 0, 1]
    e["verts"] = np.array([[-0.5, -0.5, -2.0], [-0.5, 0.5, -2.0], [0.5, 0.5, -2.0], [0.5, -0.5, -2.0]], np.float32)
    e["tris"] = np.array([[0, 1, 3], [2, 1, 3]], np.int32)
    e["attr"] = np.array([1, 1], np.int32)
    return e


def test_firefly_clamp_bounded_steps(oracle):
    """The clamp takes at most 64 steps of 0.2: the largest finite radiances
    end exactly where the unbounded loop of integrator.cpp:277-281 ends (the
    float32 loop below), and an infinite one ends the loop as +inf instead
    of spinning."""
    for v in (3.0e38, 1.0e30, 201.0):
        s = oracle.Scene(_emitter_in_view((v,) * 3))
        out, _ = s.render_paths(0, np.zeros(2, np.int32), np.arange(2, dtype=np.int32))
        x, k = np.float32(v), 0
        while x > np.float32(200):
            x = np.float32(x * np.float32(0.2))
            k += 1
        assert k <= 53
        assert np.all(out[:, 0] == x), (v, out[:, 0], x)
    s = oracle.Scene(_emitter_in_view((math.inf, 1.0, 1.0)))
    out, _ = s.render_paths(0, np.zeros(2, np.int32), np.arange(2, dtype=np.int32))
    assert np.all(np.isinf(out[:, 0]))


# ---------------------------------------------------------------------------
# oracle BVH against its brute force
# ---------------------------------------------------------------------------
def test_oracle_bvh_matches_brute_force(oracle, box_room_export):
    from conftest import random_rays
    s = oracle.Scene(box_room_export, bvh=True, leaf=2)
    rng = np.random.default_rng(5)
    rays = random_rays(rng, 4000, [-0.95, -0.95, -4.45], [0.95, 0.95, -2.6])
    a, b = s.trace(rays), s.trace(rays, use_bvh=True)
    assert np.array_equal(a["prim"], b["prim"]) and np.array_equal(a["t"], b["t"])
    a, b = s.trace(rays, any_hit=True), s.trace(rays, any_hit=True, use_bvh=True)
    assert np.array_equal(a["prim"] >= 0, b["prim"] >= 0)


# ---------------------------------------------------------------------------
# Loop subdivision restatement (loopSubdivMesh.cpp)
# ---------------------------------------------------------------------------
def _ico():
    from aurora_amd.scenes import icosahedron
    v, f = icosahedron()
    return v.astype(np.float32), f.astype(np.int32)


@pytest.mark.parametrize("level", [0, 1, 2, 3])
def test_loop_counts_and_euler(oracle, level):
    v, f = _ico()
    P, F, Nn = oracle.loop_subdivide(v, f, level)
    assert F.shape[0] == 20 * 4 ** level
    assert P.shape[0] == 10 * 4 ** level + 2
    edges = {tuple(sorted((int(a[i]), int(a[(i + 1) % 3])))) for a in F for i in range(3)}
    assert len(edges) == 30 * 4 ** level
    assert P.shape[0] - len(edges) + F.shape[0] == 2


def test_loop_icosahedron_symmetry_and_normals(oracle):
    v, f = _ico()
    P, F, Nn = oracle.loop_subdivide(v, f, 3)
    r = np.linalg.norm(P[:12], axis=1)  # the 12 original vertices keep their indices
    assert np.ptp(r) < 1e-5
    # the limit surface is close to, but not exactly, a sphere
    radial = P / np.linalg.norm(P, axis=1, keepdims=True)
    cosang = np.abs(np.sum(radial * Nn, axis=1))
    assert cosang.min() > math.cos(math.radians(5.0))
    # limit normals agree with the face normals of the limit mesh around each vertex
    fn = np.cross(P[F[:, 1]] - P[F[:, 0]], P[F[:, 2]] - P[F[:, 0]])
    acc = np.zeros_like(P)
    for k in range(3):
        np.add.at(acc, F[:, k], fn)
    acc /= np.linalg.norm(acc, axis=1, keepdims=True)
    assert np.abs(np.sum(acc * Nn, axis=1)).min() > math.cos(math.radians(3.0))


def test_loop_affine_invariance(oracle):
    """Every stencil's weights sum to one, so subdivision commutes with affine maps."""
    v, f = _ico()
    A = np.array([[1.3, 0.2, 0.0], [0.1, 0.7, 0.3], [0.0, -0.2, 1.1]], np.float32)
    b = np.array([0.5, -1.0, 2.0], np.float32)
    P1, F1, _ = oracle.loop_subdivide(v @ A.T + b, f, 2)
    P2, F2, _ = oracle.loop_subdivide(v, f, 2)
    assert np.array_equal(F1, F2)
    np.testing.assert_allclose(P1, P2 @ A.T + b, atol=2e-5)


def test_loop_single_triangle_boundary_stencils(oracle):
    """One triangle: all vertices are boundary.  Level 1 by hand: corners 3/4 v + 1/8 (n1 + n2),
    edge points are midpoints; then the 1/5 limit rule."""
    v = np.array([[0, 0, 0], [4, 0, 0], [0, 8, 0]], np.float32)
    f = np.array([[0, 1, 2]], np.int32)
    P, F, Nn = oracle.loop_subdivide(v, f, 1)
    a, b, c = v.astype(np.float64)
    even = [0.75 * a + 0.125 * (b + c), 0.75 * b + 0.125 * (a + c), 0.75 * c + 0.125 * (a + b)]
    odd = {(0, 1): 0.5 * (a + b), (1, 2): 0.5 * (b + c), (2, 0): 0.5 * (c + a)}
    # level-1 mesh: corners keep indices 0..2, edge e (face order) -> 3 + e
    lvl1 = np.array(even + [odd[(0, 1)], odd[(1, 2)], odd[(2, 0)]])
    # limit: boundary rule 3/5 v + 1/5 (two boundary neighbours)
    nb = {0: (3, 5), 1: (3, 4), 2: (4, 5), 3: (0, 1), 4: (1, 2), 5: (2, 0)}
    expect = np.array([0.6 * lvl1[i] + 0.2 * (lvl1[j] + lvl1[k]) for i, (j, k) in nb.items()])
    np.testing.assert_allclose(P, expect, atol=1e-5)
    assert F.tolist() == [[0, 3, 5], [3, 1, 4], [5, 4, 2], [3, 4, 5]]
    np.testing.assert_allclose(np.abs(Nn[:, 2]), 1.0, atol=1e-6)


# ---------------------------------------------------------------------------
# integrateDirectMIS (integrator.cpp:56-98) on a metal floor
# (metalMaterial.cpp:32-62): quadrature over the light quad
# ---------------------------------------------------------------------------
def _metal_floor_scene(tilt_deg, dist, rA, rB, mix, F0, gain, light_z, half=0.5, light_y=1.0):
    """A metal floor (camera space y = -1 before the tilt) under a square quad
    light of half-size `half` at y = light_y centred at z = light_z, emitting
    10; the scene is
    rotated by tilt_deg about x so the single camera ray (straight down -z)
    meets the floor at tilt_deg from its normal.  Returns the export and the
    rotation back to the untilted frame."""
    from aurora_amd import _lib as L
    verts = np.array([[-50, -1, 50], [50, -1, 50], [50, -1, -50], [-50, -1, -50]], np.float32)
    tris = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    normals = np.tile(np.array([0, 1, 0], np.float32), (2, 3))
    m0 = L.Material()
    m0.type, m0.distribution = 1, dist
    m0.spec[:], m0.gain[:] = list(F0), list(gain)
    m0.roughness, m0.roughnessB, m0.mix = rA, rB, mix
    m1 = L.Material()
    m1.type = 3
    m1.emission[:] = [10.0] * 3
    light = L.Light()
    light.type = 0
    light.sx = light.sy = half
    light.color[:] = [10.0] * 3
    m = np.array([[1, 0, 0, 0], [0, 0, 1, light_y], [0, -1, 0, light_z], [0, 0, 0, 1]], np.float32)
    ang = math.radians(-tilt_deg)
    R = np.array([[1, 0, 0], [0, math.cos(ang), -math.sin(ang)], [0, math.sin(ang), math.cos(ang)]], np.float32)
    Rinv = R.T
    M = np.eye(4, dtype=np.float32)
    M[:3, :3] = Rinv
    light.m[:] = list((M @ m).ravel())
    e = {"verts": (verts @ Rinv.T).astype(np.float32),
         "normals": (normals.reshape(-1, 3) @ Rinv.T).reshape(-1, 9).astype(np.float32),
         "tris": tris, "attr": np.array([0, 0], np.int32), "materials": [m0, m1], "lights": [light],
         "fov": 1e-4, "width": 1, "height": 1, "maxdepth": 1, "mindepth": 1}
    return e, Rinv


def _metal_nee_quadrature(P, V, dist, rough, F0, gain, light_z, half=0.5, light_y=1.0, k=1000):
    """Expected NEE radiance at P seen from direction V (untilted frame) by
    midpoint quadrature over the quad (side 2 half, Le 10, normal -y):
        f = schlick(n.v, F0) gain D(h) cos(theta_h) / (4 (v.h)^2)
    (cookTorrance.cpp:125-144: no Fresnel or G inside the lobe; the material
    applies Schlick at n.v), with D Beckmann exp(-tan^2/m^2)/(pi m^2 cos^4)
    or GGX m^2/(pi cos^4 (m^2 + tan^2)^2).  Returns the plain integral
    int f Le cos_s cos_l / d^2 dA, the light-strategy part with power-
    heuristic weight pL^2/(pL^2 + pB^2), and the MIS sum whose BSDF weight uses
    the pdf getSample reports (cookTorrance.cpp:62-68 take tan = 0 inside
    cos(theta_h) > 0.999)."""
    q = ((np.arange(k) + 0.5) / k - 0.5) * 2 * half
    X, Z = np.meshgrid(q, q + light_z, indexing="ij")
    dv = np.stack([X - P[0], np.full_like(X, light_y) - P[1], Z - P[2]], -1)
    d2 = (dv ** 2).sum(-1)
    l = dv / np.sqrt(d2)[..., None]
    cs = cl = l[..., 1]
    h = l + V
    h /= np.linalg.norm(h, axis=-1, keepdims=True)
    vh, nh, vn = h @ V, h[..., 1], V[1]
    m2 = rough * rough
    t2 = (1 - nh * nh) / (nh * nh)
    if dist == 0:
        Dc = np.exp(-t2 / m2) / (math.pi * m2 * nh ** 3)
    else:
        Dc = m2 / (math.pi * nh ** 3 * (m2 + t2) ** 2)
    F = np.asarray(F0) + (1 - vn) ** 5 * (1 - np.asarray(F0))
    f = (F * np.asarray(gain))[None, None, :] * (Dc / (4 * vh * vh))[..., None]
    f *= ((vh > 1e-3) & (cs > 1e-3) & (vn > 1e-3))[..., None]
    area = 4 * half * half
    pL = d2 / (cl * area)
    pB = Dc / (4 * vh)
    pBs = np.where(nh > 0.999, 1.0 / (math.pi * max(1e-3, m2)), Dc) / (4 * vh)
    wL = pL ** 2 / (pL ** 2 + pB ** 2)
    wB = pBs ** 2 / (pBs ** 2 + pL ** 2)
    g = 10.0 * cs * cl / d2
    tot = lambda w: (f * (g * w)[..., None]).sum((0, 1)) * area / (k * k)
    return tot(1.0), tot(wL), tot(wL + wB), tot(pL / (pL + pB))


def _metal_floor_run(oracle, tilt, dist, rA, rB, mix, F0, gain, lobe, light_z, bsdf_miss=False, N=512, half=0.5,
                     light_y=1.0):
    e, Rinv = _metal_floor_scene(tilt, dist, rA, rB, mix, F0, gain, light_z, half, light_y)
    s = oracle.Scene(e)
    hit = s.trace(np.array([[0, 0, 0, 0.1, 0, 0, -1, 1e7]], np.float32))
    assert hit["prim"][0] >= 0
    P = Rinv.T @ np.array([0, 0, -float(hit["t"][0])], np.float64)
    V = Rinv.T @ np.array([0, 0, 1.0])
    n = N * N
    g = (np.arange(N) + 0.5) / N
    gx, gy = np.meshgrid(g, g, indexing="ij")
    perm = np.random.default_rng(1).permutation(n)
    # vertex 0 (integrator.cpp:100-130, metalMaterial.cpp:34): slot 2 lobe, 3
    # light pick, 4-5 light sample, 6-7 the MIS BSDF sample, 8-9 continuation
    slots = np.full((n, 11), 0.5, np.float32)
    slots[:, 2] = lobe
    slots[:, 4], slots[:, 5] = gx.ravel(), gy.ravel()
    if bsdf_miss:
        # theta_h near 90 degrees: the BSDF sample leaves below the horizon
        slots[:, 6], slots[:, 7] = 0.999999, gy.ravel()[perm]
    else:
        slots[:, 6], slots[:, 7] = gx.ravel()[perm], gy.ravel()[perm]
    out, cnt = s.render_paths(0, np.zeros(n, np.int32), np.zeros(n, np.int32), slots=slots)
    r = rA if lobe > mix else rB
    return (out[:, :3].astype(np.float64).mean(0), cnt,
            _metal_nee_quadrature(P, V, dist, r, F0, gain, light_z, half, light_y), n)


@pytest.mark.parametrize("dist,tilt,light_z", [(0, 45, -3.0), (1, 45, -3.0), (0, 70, -2.0), (1, 70, -2.0)])
@pytest.mark.parametrize("lobe", [0.7, 0.2])
def test_mis_nee_on_metal_floor_matches_quadrature(oracle, dist, tilt, light_z, lobe):
    """Light sample + BSDF sample with the power heuristic (integrator.cpp:56-98)
    on a Beckmann / GGX metal floor: the injected-slot mean over a stratified
    512^2 grid equals the quadrature of the MIS weights times f Le cos cos / d^2
    to 3e-5.  The lobe slot picks roughnessA above mix and roughnessB below it
    (metalMaterial.cpp:34-62), and the lobe colour is schlick(n.v, F0) * gain
    with F0 and gain differing per channel.  At light_z = -3 the light covers
    the specular peak, where the plain integral differs from the MIS sum by
    the getSample pdf quirk (up to 1.2e-4), so that case checks the weights
    getSample's pdf enters."""
    F0, gain = (0.04, 0.5, 0.9), (1.0, 0.8, 0.6)
    mc, cnt, (E, EL, Emis, _), n = _metal_floor_run(oracle, tilt, dist, 0.3, 0.6, 0.4, F0, gain, lobe, light_z)
    np.testing.assert_allclose(mc, Emis, rtol=3e-5)
    # both strategies traced shadow rays: one light ray per path, BSDF rays where the sample met the quad
    assert n < cnt["shadow_rays"] < 1.2 * n
    # the other roughness or the colour without the Schlick angle term is far off
    r_other = 0.6 if lobe > 0.4 else 0.3
    _, _, (E2, _, Emis2, _), _ = _metal_floor_run(oracle, tilt, dist, r_other, r_other, 0.4, F0, gain, lobe, light_z,
                                                  N=16)
    assert np.all(np.abs(Emis2 / Emis - 1) > 0.05)


@pytest.mark.parametrize("dist", [0, 1])
def test_mis_light_strategy_weight_is_the_power_heuristic(oracle, dist):
    """With the BSDF samples forced below the horizon only the light strategy
    contributes: its mean equals the quadrature with the power-heuristic weight
    pL^2 / (pL^2 + pB^2) (integrator.cpp:28-31, 74), not the balance heuristic
    or weight 1, and exactly one shadow ray per path is traced."""
    F0, gain = (0.04, 0.5, 0.9), (1.0, 0.8, 0.6)
    # the light covers the lobe's peak: the balance heuristic pL / (pL + pB)
    # or weight 1 would move the mean by far more than the tolerance
    mc, cnt, (E, EL, Emis, EL_balance), n = _metal_floor_run(oracle, 45, dist, 0.3, 0.3, 0.4, F0, gain, 0.7, -3.0,
                                                             bsdf_miss=True)
    assert cnt["shadow_rays"] == n
    np.testing.assert_allclose(mc, EL, rtol=3e-5)
    assert np.all(np.abs(EL_balance / EL - 1) > 1e-2) and np.all(np.abs(E / EL - 1) > 5e-3)


# ---------------------------------------------------------------------------
# Russian roulette (integrator.cpp:219-248): cumulative continue probability,
# kill when the uniform exceeds it, survivors divided by it
# ---------------------------------------------------------------------------
def _grey_room(box_room_export, mindepth, maxdepth=4):
    """The box room with every surface matte grey 0.5 (lum(albedo/pi) ndl /
    (1/2pi) = ndl <= 1, so the continue probability stays below one)."""
    from aurora_amd import _lib as L
    e = dict(box_room_export)
    mats = []
    for m in e["materials"]:
        if m.type == 3:
            mats.append(m)
            continue
        g = L.Material()
        g.type = 0
        g.diffuse[:] = [0.5, 0.5, 0.5]
        mats.append(g)
    e["materials"] = mats
    e["mindepth"], e["maxdepth"] = mindepth, maxdepth
    return e


def test_roulette_divides_by_its_survival_probability(oracle, box_room_export):
    """Paths of depth 4 with every continuation cosine at 0.5 have continue
    probability 0.5, 0.25, 0.125 at their vertices 1-3 (mindepth 0; vertex 0 has
    bounces == mindepth).  Over an 8^3 grid of the three roulette uniforms the
    mean radiance of each path equals its radiance without roulette (one
    division by cp per survival, kill when u > cp), and the extension rays of a
    path that stays in the room are 2 + 1/2 + 1/8 + 1/64 per grid point on
    average: the probability is the cumulative product (a per-vertex one gives
    2 + 1/2 + 1/4 + 1/8)."""
    norr = oracle.Scene(_grey_room(box_room_export, 4))
    rr = oracle.Scene(_grey_room(box_room_export, 0))
    W = box_room_export["width"]
    rng = np.random.default_rng(12)
    nslots = 2 + 9 * 4
    K = 8
    g = (np.arange(K) + 0.5) / K
    u1, u2, u3 = [a.ravel() for a in np.meshgrid(g, g, g, indexing="ij")]
    full, checked = 0, 0
    for path in range(24):
        base = rng.uniform(0.02, 0.98, nslots).astype(np.float32)
        base[0:2] = rng.uniform(0.3, 0.7, 2)
        pix = np.array([(W // 2) * W + W // 2 + int(rng.integers(-20, 20))], np.int32)
        for v in range(4):
            base[2 + 9 * v + 6] = 0.5   # continuation: z = cos = 0.5 (lambert.h:44-59)
        ref, rc = norr.render_paths(0, pix, np.zeros(1, np.int32), slots=base[None])
        slots = np.repeat(base[None], K ** 3, 0)
        slots[:, 2 + 9 * 1 + 8], slots[:, 2 + 9 * 2 + 8], slots[:, 2 + 9 * 3 + 8] = u1, u2, u3
        out, oc = rr.render_paths(0, np.repeat(pix, K ** 3), np.zeros(K ** 3, np.int32), slots=slots)
        # float32 paths: cp and 1/cp round within an ulp or two of 2^-k
        np.testing.assert_allclose(out[:, :3].astype(np.float64).mean(0), ref[0, :3], rtol=2e-5, atol=1e-9)
        if rc["ext_rays"] == 5 and rc["hit_rays"] == 5:  # every vertex shaded, none on the light
            assert oc["ext_rays"] == K ** 3 * 2 + K ** 3 // 2 + K ** 3 // 8 + K ** 3 // 64
            full += 1
        checked += ref[0, :3].sum() > 0
    assert full >= 4 and checked >= 12


def test_roulette_is_unbiased_in_full_paths(oracle, box_room_export):
    """Same paths (same seed, so the same directions) with mindepth 0 and with
    mindepth = maxdepth = 5 in the grey room: roulette kills paths (fewer
    extension rays) and leaves the means unchanged.  The continue probability
    is proportional to the sampled cosine, so E[1/cp] -- the variance of the
    roulette estimator -- diverges logarithmically and the sample deviation
    understates the spread: per pixel the paired difference stays within
    4 sigma (at 256..4096 spp about 1.5 % of pixels pass 3 sigma), pooled over
    300 pixels within 1 %.  The exact check of the roulette step is the
    injected-grid test above."""
    a = oracle.Scene(_grey_room(box_room_export, 0, 5))
    b = oracle.Scene(_grey_room(box_room_export, 5, 5))
    W, H = box_room_export["width"], box_room_export["height"]
    rng = np.random.default_rng(3)
    pix = rng.choice(W * H, 300, replace=False).astype(np.int32)
    spp = 256
    P = np.repeat(pix, spp)
    S = np.tile(np.arange(spp, dtype=np.int32), pix.size)
    la, ca = a.render_paths(9, P, S)
    lb, cb = b.render_paths(9, P, S)
    assert ca["ext_rays"] < 0.9 * cb["ext_rays"]
    d = (la[:, 0] - lb[:, 0]).astype(np.float64).reshape(pix.size, spp)  # grey: r = g = b
    z = np.abs(d.mean(1)) / (d.std(1) / math.sqrt(spp) + 1e-12)
    assert np.mean(z > 4.0) <= 0.01, np.sort(z)[-5:]
    assert abs(la[:, :3].astype(np.float64).mean() / lb[:, :3].astype(np.float64).mean() - 1) < 0.01


@pytest.mark.parametrize("dist", [0, 1])
def test_scene_kelemen_table_is_the_fp16_250k_estimate(oracle, dist):
    """A scene's Kelemen table is the 250,000-sample estimate per entry
    (SPECALBEDO_SAMPLES, constants.h:26) rounded to FP16 as the reference's
    EXR round trip stores it (kelemenMaterial.cpp:53-76): every entry is an
    FP16 value within half an FP16 ulp of the float estimate, and the estimate
    is within its Monte-Carlo error of a 4x smaller independent one."""
    assert oracle.KELEMEN_SAMPLES == 250000
    t = oracle.scene_kelemen_table(0.05, dist)
    est = oracle.kelemen_table(0.05, dist)
    assert np.array_equal(t, t.astype(np.float16).astype(np.float32))
    ulp = np.spacing(est.astype(np.float16)).astype(np.float32)
    assert np.all(np.abs(t - est) <= 0.5 * ulp + 1e-12)
    small = oracle.kelemen_table(0.05, dist, n_samples=62500, seed=7)
    assert np.max(np.abs(small - est)) < 0.02 and np.mean(np.abs(small - est)) < 2e-3


# ---------------------------------------------------------------------------
# texture coordinates through Loop subdivision (loopSubdivMesh.cpp:174-191,
# 360-413, 565-613, 639-692, 477-482)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("levels", [1, 2, 3])
def test_loop_uv_uses_the_position_stencils(oracle, levels):
    """Seamless uvs equal to each vertex's (x, y) on the icosahedron (valence
    5 everywhere, every edge interior) come out as the (x, y) of the vertices
    before the limit step: the even rule with Warren's beta and the odd rule
    3/8, 3/8, 1/8, 1/8 act on uvs exactly as on positions, and the uvs are not
    pushed to the limit (with the limit step the positions move away)."""
    from aurora_amd.scenes import icosahedron
    v, f = icosahedron()
    P, F = v.astype(np.float32), f.astype(np.int32)
    uv = P[F][..., :2].copy()
    Pp, Fp, _, UVo = oracle.loop_subdivide(P, F, levels, uv=uv, limit=False)
    np.testing.assert_allclose(UVo, Pp[Fp][..., :2], atol=2e-6)
    Pl, Fl, _, UVl = oracle.loop_subdivide(P, F, levels, uv=uv)
    np.testing.assert_array_equal(UVl, UVo)
    assert np.array_equal(Fl, Fp) and np.abs(Pl - Pp).max() > 1e-3


def test_loop_uv_seam_one_level(oracle):
    """Two triangles (a quad) sharing the edge v0-v2.  Without a seam the
    shared edge's odd uv is 3/8 (u0 + u2) + 1/8 (u1 + u3) and the corners take
    the boundary rule 3/4 u + 1/8 (neighbours); with the right triangle's uvs
    shifted by (10, 0) the shared edge is a seam: each face gets the midpoint
    of its own end uvs there, and the seam's end vertices keep their uvs."""
    P = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32)
    F = np.array([[0, 1, 2], [3, 4, 5]], np.int32)
    u = {0: np.array([0.0, 0.0]), 1: np.array([1.0, 0.0]), 2: np.array([1.0, 1.0]), 3: np.array([0.0, 1.0])}
    # face 0 = (v0, v1, v2), face 1 = (v0, v2, v3) in position
    uv = np.array([[u[0], u[1], u[2]], [u[0], u[2], u[3]]], np.float32)
    _, Fo, _, UVo = oracle.loop_subdivide(P, F, 1, uv=uv)
    # child 0 of face 0: (even v0, odd e01, odd e20); child 3: (e01, e12, e20)
    even0 = 0.75 * u[0] + 0.125 * (u[1] + u[3])
    odd02 = 3 / 8 * (u[0] + u[2]) + 1 / 8 * (u[1] + u[3])
    np.testing.assert_allclose(UVo[0, 0], even0, atol=1e-6)
    np.testing.assert_allclose(UVo[0, 2], odd02, atol=1e-6)
    np.testing.assert_allclose(UVo[3, 0], 0.5 * (u[0] + u[1]), atol=1e-6)
    # the seam: face 1's uvs shifted
    uv2 = uv.copy()
    uv2[1] += np.array([10.0, 0.0], np.float32)
    _, _, _, UVs = oracle.loop_subdivide(P, F, 1, uv=uv2)
    np.testing.assert_allclose(UVs[0, 0], u[0], atol=1e-6)                    # v0 keeps its uv (face 0)
    np.testing.assert_allclose(UVs[4, 0], u[0] + [10, 0], atol=1e-6)          # and in face 1
    np.testing.assert_allclose(UVs[0, 2], 0.5 * (u[0] + u[2]), atol=1e-6)      # face 0's seam midpoint
    np.testing.assert_allclose(UVs[4, 1], 0.5 * (u[0] + u[2]) + [10, 0], atol=1e-6)  # face 1's
    np.testing.assert_allclose(UVs[0, 1], 0.5 * (u[0] + u[1]), atol=1e-6)      # a boundary edge
    # v1, off the seam: boundary rule on its own face's uvs
    np.testing.assert_allclose(UVs[1, 1], 0.75 * u[1] + 0.125 * (u[0] + u[2]), atol=1e-6)
