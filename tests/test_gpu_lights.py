"""GPU: per-query light calls (aur_light_queries) against the oracle's
restatement: SquareLight generateSample / eval / pdf / visible
(squareLight.h:57-122) in the box room, and the environment lights of the
textured stage -- a lat-long map with its sampling pdf and a constant one
(infiniteAreaLight.cpp:35-106)."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _queries(rng, n, n_lights, lo, hi):
    from aurora_amd import _lib as L
    qs = []
    for i in range(n):
        q = L.LightQuery()
        q.light = i % n_lights
        P = rng.uniform(lo, hi)
        N = rng.normal(size=3)
        N /= np.linalg.norm(N)
        D = rng.normal(size=3)
        D /= np.linalg.norm(D)
        q.P[:], q.N[:], q.D[:] = list(P), list(N), list(D)
        q.u1, q.u2 = float(rng.uniform()), float(rng.uniform())
        q.tmax = 1e7
        qs.append(q)
    return qs


def _check(r, s, oracle, qs):
    from oracle import oracle as O
    res = r.light(qs)
    e = r.export()
    n_area = n_env = vis = 0
    for q, g in zip(qs, res):
        li = q.light
        light = s.lights[li]
        P, N, D = np.array(q.P[:], np.float32), np.array(q.N[:], np.float32), np.array(q.D[:], np.float32)
        d, tmax, pdf, col = oracle.light_sample(light, P, N, q.u1, q.u2, scene=s)
        np.testing.assert_allclose(g.sample_dir[:], d, atol=2e-6)
        assert g.sample_tmax == pytest.approx(tmax, rel=1e-6, abs=1e-6)
        assert g.sample_pdf == pytest.approx(pdf, rel=1e-5, abs=1e-12)
        np.testing.assert_allclose(g.sample_color[:], col, rtol=1e-5)
        lp, lt = oracle.light_pdf(light, P, D, tmin=0.0005, tmax=q.tmax, scene=s)
        assert g.pdf == pytest.approx(lp, rel=1e-5, abs=1e-12)
        assert g.pdf_tmax == pytest.approx(lt, rel=1e-6)
        if e["lights"][li].type == 1:
            n_env += 1
            np.testing.assert_allclose(g.Le[:], s.env_eval(li, D), rtol=1e-5)
        else:
            n_area += 1
            np.testing.assert_allclose(g.Le[:], col, rtol=1e-6)  # a quad's eval is its radiance
        v = O.lib().orc_light_visible(C.byref(light), O.fp(P), O.fp(N))
        assert g.visible == v
        vis += v
    return n_area, n_env, vis


def test_square_light_queries(gpu_box_room, oracle):
    s = oracle.Scene(gpu_box_room.export())
    rng = np.random.default_rng(3)
    qs = _queries(rng, 3000, 1, [-0.9, -0.9, -4.4], [0.9, 0.9, -2.6])
    n_area, _, vis = _check(gpu_box_room, s, oracle, qs)
    res = gpu_box_room.light(qs)
    # some rays from inside the room hit the quad (pdf > 0, maxt shortened), most miss it
    hits = sum(1 for g in res if g.pdf > 0)
    assert n_area == 3000 and 0 < hits < 3000 and 0 < vis < 3000
    assert all(g.pdf_tmax < 1e7 for g in res if g.pdf > 0)


def test_environment_light_queries(tmp_path, oracle):
    from aurora_amd import scenes
    from aurora_amd.renderer import Renderer
    path = scenes.textured_stage(res=32, spp=1, maxdepth=2, mindepth=2).write(str(tmp_path))
    r = Renderer(path, os.path.join(str(tmp_path), "config.json"))
    r.build()
    e = r.export()
    s = oracle.Scene(e)
    rng = np.random.default_rng(4)
    qs = _queries(rng, 3000, len(e["lights"]), [-1, 0, -4], [1, 1, -2])
    n_area, n_env, _ = _check(r, s, oracle, qs)
    assert n_env > 0
    from aurora_amd import _lib as L
    with pytest.raises(L.AuroraError):
        bad = L.LightQuery()
        bad.light = len(e["lights"])
        r.light([bad])
    r.close()
