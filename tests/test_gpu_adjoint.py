"""GPU path-replay adjoint: against the CPU reference gradient, against central
finite differences (h = 1e-3) of the GPU forward renderer, and one short
inverse-rendering run through the torch.autograd.Function."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def kroom(tmp_path_factory):
    from aurora_amd import scenes
    from aurora_amd.renderer import Renderer
    d = str(tmp_path_factory.mktemp("kroom"))
    scenes.box_room(res=48, spp=8, maxdepth=4, mindepth=4, kelemen=True).write(d)
    r = Renderer(os.path.join(d, "scene.asc"))
    r.build()
    info = r.info()
    rng = np.random.default_rng(4)
    albedo = rng.uniform(0.2, 0.8, size=(info["n_tris"], 3)).astype(np.float32)
    spec = rng.uniform(0.2, 0.8, size=info["n_attrs"]).astype(np.float32)
    rough = rng.uniform(0.1, 0.7, size=info["n_attrs"]).astype(np.float32)
    r.set_tri_albedo(albedo)
    r.set_material_params(rough, spec)
    w = rng.uniform(-1, 1, size=(info["height"], info["width"], 3)).astype(np.float32)
    yield r, albedo, rough, spec, w
    r.close()


def _loss(r, albedo, rough, spec, w, passes, seed):
    r.set_tri_albedo(albedo)
    r.set_material_params(rough, spec)
    r.set_seed(seed)
    acc, _ = r.render_passes(0, passes)
    return float(np.sum((acc[..., :3] / passes).astype(np.float64) * w))


def test_adjoint_matches_cpu_reference(kroom, oracle):
    r, albedo, rough, spec, w = kroom
    r.set_seed(13)
    ga, gr, gs, st = r.adjoint(0, 8, w)
    s = oracle.Scene(r.export(), tri_albedo=albedo)
    s.set_material_params(rough, spec)
    dl = np.zeros(w.shape[:2] + (4,), np.float32)
    dl[..., :3] = w
    oa, orr, osp = s.adjoint(13, 0, 8, dl, threads=8)
    scale = np.abs(oa).max()
    np.testing.assert_allclose(ga, oa, rtol=2e-3, atol=2e-4 * scale)
    np.testing.assert_allclose(gs, osp, rtol=2e-3, atol=2e-4 * np.abs(osp).max())
    np.testing.assert_allclose(gr, orr, rtol=5e-3, atol=1e-3 * np.abs(orr).max())
    assert st["paths"] == 8 * w.shape[0] * w.shape[1]


def test_adjoint_matches_finite_differences(kroom):
    """Albedo and spec weight: exact for fixed random numbers.  Roughness: the
    reference's lookup makes the image piecewise in roughness (jumps every 1/31),
    so h = 1e-3 stays inside one cell; agreement is statistical."""
    r, albedo, rough, spec, w = kroom
    passes, seed, h = 16, 21, 1e-3
    r.set_tri_albedo(albedo)
    r.set_material_params(rough, spec)
    r.set_seed(seed)
    ga, gr, gs, _ = r.adjoint(0, passes, w)
    worst = 0.0
    for idx in np.argsort(-np.abs(ga.ravel()))[:24]:
        t, c = divmod(int(idx), 3)
        ap, am = albedo.copy(), albedo.copy()
        ap[t, c] += h
        am[t, c] -= h
        fd = (_loss(r, ap, rough, spec, w, passes, seed) - _loss(r, am, rough, spec, w, passes, seed)) / (2 * h)
        worst = max(worst, abs(ga[t, c] - fd) / max(abs(fd), 1e-3 * np.abs(ga).max()))
    for m in range(len(spec)):
        if abs(gs[m]) < 1e-2 * np.abs(gs).max():
            continue
        sp, sm = spec.copy(), spec.copy()
        sp[m] += h
        sm[m] -= h
        fd = (_loss(r, albedo, rough, sp, w, passes, seed) - _loss(r, albedo, rough, sm, w, passes, seed)) / (2 * h)
        worst = max(worst, abs(gs[m] - fd) / abs(fd))
    assert worst < 1e-2, worst
    r.set_tri_albedo(albedo)
    r.set_material_params(rough, spec)


def test_roughness_gradient_statistically(kroom):
    r, albedo, rough, spec, w = kroom
    passes, h = 256, 1e-3
    m = int(np.argmax([abs(x) for x in rough]))
    fds, adj = [], []
    for seed in range(4):
        r.set_tri_albedo(albedo)
        r.set_material_params(rough, spec)
        r.set_seed(100 + seed)
        _, gr, _, _ = r.adjoint(0, passes, w)
        adj.append(gr)
        fd = []
        for k in range(len(rough)):
            rp, rm = rough.copy(), rough.copy()
            rp[k] += h
            rm[k] -= h
            fd.append((_loss(r, albedo, rp, spec, w, passes, 100 + seed) -
                       _loss(r, albedo, rm, spec, w, passes, 100 + seed)) / (2 * h))
        fds.append(fd)
    fds, adj = np.array(fds), np.array(adj)
    big = np.abs(adj.mean(0)) > 0.2 * np.abs(adj.mean(0)).max()
    se = np.sqrt(fds.var(0) / len(fds) + adj.var(0) / len(adj)) + 1e-6
    z = np.abs(fds.mean(0) - adj.mean(0)) / se
    assert np.all(z[big] < 4.0), (fds.mean(0), adj.mean(0), se)
    r.set_material_params(rough, spec)
    del m


def test_autograd_function_reduces_loss(tmp_path):
    torch = pytest.importorskip("torch")
    from aurora_amd import inverse
    p = inverse.make_problem(str(tmp_path), res=64, spp=32, depth=4)
    hist, params = inverse.optimize(p, steps=30, lr=0.05)
    assert np.isfinite(hist).all()

    def loss(prm):  # common random numbers for a noise-free comparison
        p.renderer.set_seed(77)
        img = inverse.render_image(p.renderer, prm["albedo"], prm["rough"], prm["spec"], 0, 64)
        return float(np.mean((img - p.target) ** 2))

    l0, l1 = loss(p.init), loss(params)
    assert l1 < 0.9 * l0, (l0, l1)
    seen = np.abs(params["albedo"] - p.init["albedo"]).sum(1) > 0  # triangles the camera sees
    err0 = np.abs(p.init["albedo"] - p.truth["albedo"])[seen].mean()
    err1 = np.abs(params["albedo"] - p.truth["albedo"])[seen].mean()
    assert err1 < err0, (err0, err1)
    p.renderer.close()
    del torch


@pytest.fixture(scope="module")
def metal_room(tmp_path_factory):
    """The config-1 box room (GGX metal box, attribute 0) with per-triangle
    albedo and a roughness parameter per material."""
    from aurora_amd import scenes
    from aurora_amd.renderer import Renderer
    d = str(tmp_path_factory.mktemp("mroom"))
    scenes.box_room(res=48, spp=8, maxdepth=4, mindepth=4).write(d)
    r = Renderer(os.path.join(d, "scene.asc"))
    r.build()
    info = r.info()
    assert r.export()["materials"][0].type == 1
    rng = np.random.default_rng(6)
    albedo = rng.uniform(0.2, 0.8, size=(info["n_tris"], 3)).astype(np.float32)
    rough = np.full(info["n_attrs"], 0.35, np.float32)
    spec = np.full(info["n_attrs"], 0.5, np.float32)
    r.set_tri_albedo(albedo)
    r.set_material_params(rough, spec)
    w = rng.uniform(-1, 1, size=(info["height"], info["width"], 3)).astype(np.float32)
    yield r, albedo, rough, spec, w
    r.close()


def test_metal_roughness_adjoint_matches_cpu_reference(metal_room, oracle):
    """metalMaterial.cpp:32-62: the GGX lobe's roughness (roughnessA =
    roughnessB) gets a gradient; GPU and oracle replay the same paths."""
    r, albedo, rough, spec, w = metal_room
    r.set_seed(19)
    ga, gr, gs, _ = r.adjoint(0, 8, w)
    s = oracle.Scene(r.export(), tri_albedo=albedo)
    s.set_material_params(rough, spec)
    dl = np.zeros(w.shape[:2] + (4,), np.float32)
    dl[..., :3] = w
    oa, orr, _ = s.adjoint(19, 0, 8, dl, threads=8)
    assert abs(orr[0]) > 0 and np.count_nonzero(orr) == 1  # only the metal box has a rough lobe
    np.testing.assert_allclose(ga, oa, rtol=2e-3, atol=2e-4 * np.abs(oa).max())
    np.testing.assert_allclose(gr, orr, rtol=5e-3, atol=1e-3 * np.abs(orr).max())


def test_metal_roughness_gradient_statistically(metal_room):
    """The adjoint's metal-roughness gradient against central differences
    (common random numbers) over eight seeds: the two estimate the same
    derivative with different noise, so their means agree within 4 sigma.
    Roughness moves the sampled directions, and a path that changes what it
    hits makes a step of its own, so the difference uses h = 0.02 (bias of
    order h^2) rather than 1e-3 (noise of order 1/h)."""
    r, albedo, rough, spec, w = metal_room
    passes, h = 128, 0.02
    fds, adj = [], []
    for seed in range(8):
        r.set_tri_albedo(albedo)
        r.set_material_params(rough, spec)
        r.set_seed(300 + seed)
        _, gr, _, _ = r.adjoint(0, passes, w)
        adj.append(gr[0])
        rp, rm = rough.copy(), rough.copy()
        rp[0] += h
        rm[0] -= h
        fds.append((_loss(r, albedo, rp, spec, w, passes, 300 + seed) -
                    _loss(r, albedo, rm, spec, w, passes, 300 + seed)) / (2 * h))
    fds, adj = np.array(fds), np.array(adj)
    se = np.sqrt(fds.var() / len(fds) + adj.var() / len(adj)) + 1e-9
    assert abs(fds.mean() - adj.mean()) < 4 * se, (fds, adj)
    assert se < 0.5 * abs(adj.mean()), (fds, adj)  # the comparison resolves the gradient
    r.set_material_params(rough, spec)


def test_adjoint_at_inverse_rendering_size(tmp_path, oracle):
    """The inverse-rendering workload's size: 512x512, depth 6, 32 passes.  GPU
    and oracle gradients agree on a 32x32 subset of pixels (dL/dI zero
    elsewhere: those pixels add nothing on either side)."""
    from aurora_amd import scenes
    from aurora_amd.renderer import Renderer
    p = scenes.write_config("box_room_kelemen", str(tmp_path))
    r = Renderer(p)
    r.build()
    info = r.info()
    assert (info["width"], info["height"], info["maxdepth"]) == (512, 512, 6)
    rng = np.random.default_rng(8)
    albedo = rng.uniform(0.1, 0.9, size=(info["n_tris"], 3)).astype(np.float32)
    rough = rng.uniform(0.05, 0.8, size=info["n_attrs"]).astype(np.float32)
    spec = rng.uniform(0.0, 1.0, size=info["n_attrs"]).astype(np.float32)
    r.set_tri_albedo(albedo)
    r.set_material_params(rough, spec)
    w = np.zeros((512, 512, 3), np.float32)
    w[8::16, 8::16] = rng.uniform(-1, 1, size=(32, 32, 3))
    r.set_seed(23)
    ga, gr, gs, st = r.adjoint(0, 32, w)
    assert st["paths"] == 32 * 512 * 512
    s = oracle.Scene(r.export(), tri_albedo=albedo)
    s.set_material_params(rough, spec)
    dl = np.zeros((512, 512, 4), np.float32)
    dl[..., :3] = w
    oa, orr, osp = s.adjoint(23, 0, 32, dl, threads=16)
    np.testing.assert_allclose(ga, oa, rtol=2e-3, atol=2e-4 * np.abs(oa).max())
    np.testing.assert_allclose(gs, osp, rtol=2e-3, atol=2e-4 * np.abs(osp).max())
    np.testing.assert_allclose(gr, orr, rtol=5e-3, atol=1e-3 * np.abs(orr).max())
    r.close()


def test_kept_records_replace_the_first_replay(kroom):
    """A forward render that kept its per-path records lets the adjoint of the
    same passes skip its first replay: the gradients equal those of the full
    two-replay adjoint (up to the order of float atomic adds), the call traces
    only the gradient replay's rays, and any change of inputs, or a pass range
    the records do not cover, falls back to the full adjoint."""
    r, albedo, rough, spec, w = kroom
    r.set_tri_albedo(albedo)
    r.set_material_params(rough, spec)
    r.set_seed(29)
    r.keep_path_records(0, 0)
    full = r.adjoint(0, 8, w)
    r.keep_path_records(0, 8)
    r.render_passes(0, 8, fetch=False)
    kept = r.adjoint(0, 8, w)
    assert kept[3]["paths"] == full[3]["paths"]
    assert kept[3]["ext_rays"] < full[3]["ext_rays"] and kept[3]["shadow_rays"] < full[3]["shadow_rays"]
    for a, b in zip(kept[:3], full[:3]):
        np.testing.assert_allclose(a, b, rtol=1e-4, atol=1e-6 * np.abs(b).max())
    # a sub-range of the kept passes is served from the records too
    sub = r.adjoint(2, 4, w)
    r.keep_path_records(0, 0)
    sub_full = r.adjoint(2, 4, w)
    np.testing.assert_allclose(sub[1], sub_full[1], rtol=1e-4, atol=1e-6 * np.abs(sub_full[1]).max())
    assert sub[3]["ext_rays"] < sub_full[3]["ext_rays"]
    # a new seed invalidates the records: the full adjoint runs
    r.keep_path_records(0, 8)
    r.render_passes(0, 8, fetch=False)
    r.set_seed(31)
    stale = r.adjoint(0, 8, w)
    r.keep_path_records(0, 0)
    fresh = r.adjoint(0, 8, w)
    assert stale[3]["ext_rays"] == fresh[3]["ext_rays"]
    np.testing.assert_allclose(stale[1], fresh[1], rtol=1e-4, atol=1e-6 * np.abs(fresh[1]).max())
    # records of only part of the range: the full adjoint runs
    r.keep_path_records(0, 8)
    r.render_passes(0, 4, fetch=False)
    part = r.adjoint(0, 8, w)
    assert part[3]["ext_rays"] == fresh[3]["ext_rays"]
    r.keep_path_records(0, 0)
