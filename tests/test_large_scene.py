"""The large differentiable scene's generator and the host side of its
loading: terrain patches (heightfield grid, uvs, normals), the noise texture,
the grid OBJ writer read back through the library's OBJ parser, and the
attribute index of the terrain."""
import os

import numpy as np

from conftest import host_renderer


def test_terrain_patch_grid():
    from aurora_amd import scenes
    m = scenes.terrain_patch(12, 5.0, 3, 0.6, origin=(-2.0, 1.0), uv0=(0.25, 0.5), uv1=(0.5, 1.0))
    v, f, n, uv = m["verts"], m["faces"], m["normals"], m["uvs"]
    assert v.shape == (13 * 13, 3) and f.shape == (2 * 12 * 12, 3)
    assert np.allclose(v[:, 0].min(), -2.0) and np.allclose(v[:, 0].max(), 3.0)
    assert np.allclose(v[:, 2].min(), 1.0) and np.allclose(v[:, 2].max(), 6.0)
    assert np.abs(v[:, 1]).max() <= 0.6 + 1e-12
    # heights are the noise of the world position
    np.testing.assert_allclose(v[:, 1], 0.6 * scenes.fbm2(v[:, 0], v[:, 2], 3, 5, 0.15), rtol=0, atol=1e-12)
    # uvs linear in x and z over [uv0, uv1]
    np.testing.assert_allclose(uv[:, 0], 0.25 + (v[:, 0] + 2.0) / 5.0 * 0.25, atol=1e-12)
    np.testing.assert_allclose(uv[:, 1], 0.5 + (v[:, 2] - 1.0) / 5.0 * 0.5, atol=1e-12)
    # every face is counter-clockwise seen from above, normals point up
    p = v[f]
    fn = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    assert np.all(fn[:, 1] > 0)
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0) and np.all(n[:, 1] > 0.5)
    # every vertex is used and every quad is split once
    assert np.unique(f).size == v.shape[0]


def test_noise_and_texture():
    from aurora_amd import scenes
    t = (np.arange(48) + 0.5) / 48
    X, Z = np.meshgrid(t, t)
    # the separable grid evaluation is the point evaluation, bit for bit
    assert np.array_equal(scenes._grid_fbm(t, 9, 5, 4.0), scenes.fbm2(X, Z, 9, 5, 4.0))
    tex = scenes.albedo_texture(64, 5)
    assert tex.shape == (64, 64, 3) and tex.dtype == np.float32
    for c in range(3):
        assert abs(tex[..., c].min() - 0.1) < 1e-6 and abs(tex[..., c].max() - 0.9) < 1e-6
    assert not np.array_equal(tex[..., 0], tex[..., 1])
    assert np.array_equal(tex, scenes.albedo_texture(64, 5))          # deterministic
    # lattice hash: uniform-looking values in [0, 1)
    h = scenes._hash2(np.arange(-500, 500), np.zeros(1000, np.int64), 77)
    assert h.min() >= 0 and h.max() < 1 and abs(h.mean() - 0.5) < 0.05


def test_large_scene_loads_with_uvs(tmp_path):
    """A small instance of the large scene (no subdivided spheres, so the host
    can flatten it): the terrain triangles carry the generator's uvs, and the
    terrain attribute is attr_index('terrain00')."""
    from aurora_amd import scenes
    spec, truth = scenes.large_scene(width=32, height=18, spp=1, maxdepth=2, mindepth=2, n_spheres=0,
                                     patches=(2, 2), patch_quads=6, patch_size=4.0, tex_res=16)
    assert truth.shape == (16, 16, 3)
    path = spec.write(str(tmp_path))
    r = host_renderer()
    r.load(path)
    e = r.export()
    r.close()
    n_terrain = 4 * 2 * 6 * 6
    attr = e["attr"]
    t_attr = spec.attr_index("terrain00")
    assert [spec.attr_index("terrain%d%d" % (i, j)) for i in range(2) for j in range(2)] == [0, 1, 2, 3]
    assert (attr < 4).sum() == n_terrain and t_attr == 0
    # per-corner uvs of patch (0, 0) equal the generator's
    m = scenes.terrain_patch(**spec.meshes["terrain00.obj"]["terrain"])
    uv = e["uvs"].reshape(-1, 3, 2)[attr == t_attr]
    np.testing.assert_allclose(uv, m["uvs"][m["faces"]].astype(np.float32), atol=1e-6)
    # the grid writer's text round-trips the positions to float precision
    cam = e["verts"][e["tris"][attr == t_attr]]
    assert cam.shape == m["verts"][m["faces"]].shape


def test_obj_parser_line_forms(tmp_path):
    """CRLF line ends, tabs, comments and a missing trailing newline read as
    before; v/t/n corners take their uv from the middle index."""
    from aurora_amd import scenes
    p = tmp_path / "m.obj"
    p.write_bytes(b"# c\r\nv 0 0 0\r\nv\t1 0 0\r\nv 0 1 0\r\nv 1 1 0\r\nvt 0.5 0.25\r\nvt 1 1\r\nvn 0 0 1\r\n"
                  b"f 1/1/1 2/2/1 3/1/1\r\nf 2/2/1\t4/1/1 3/2/1")
    spec = scenes.box_room()
    d = tmp_path / "scene"
    spec.write(str(d))
    os.replace(str(p), str(d / "meshes" / "box_short.obj"))
    r = host_renderer()
    r.load(str(d / "scene.asc"))
    e = r.export()
    r.close()
    first = e["attr"] == 0
    assert first.sum() == 2
    uv = e["uvs"].reshape(-1, 3, 2)[first]
    np.testing.assert_allclose(uv[0], [[0.5, 0.25], [1, 1], [0.5, 0.25]])
    np.testing.assert_allclose(uv[1], [[1, 1], [0.5, 0.25], [1, 1]])
    tri = e["tris"][first]
    np.testing.assert_allclose(e["verts"][tri[1]] - e["verts"][tri[0][0]], [[1, 0, 0], [1, 1, 0], [0, 1, 0]],
                               atol=1e-6)
