"""Light samples (renderer.cpp:233, 257; integrator.cpp:152-300) in the oracle:
pass 0 traces one path per pixel, later passes `lightsamples` paths that share
the pass's camera sample (camera ray traced once), each with its own vertex
stream, averaged as sum(Lo_j * (1/L)); a camera miss returns the environment
once.  The property below rebuilds a pass from single injected-slot paths."""
import numpy as np


def _pcg(v):
    v = np.uint32(v)
    s = np.uint32(v * np.uint32(747796405) + np.uint32(2891336453))
    w = np.uint32(((s >> np.uint32((s >> np.uint32(28)) + np.uint32(4))) ^ s) * np.uint32(277803737))
    return np.uint32((w >> np.uint32(22)) ^ w)


def test_pass_is_mean_of_its_light_sample_paths(oracle, box_room_export):
    e = dict(box_room_export)
    W, H = e["width"], e["height"]
    L = 3
    e["lightsamples"] = L
    s = oracle.Scene(e)
    seed, p = 9, 2
    acc, st = s.render(seed, p, 1)
    rng = np.random.default_rng(0)
    nslots = 2 + 9 * e["maxdepth"]
    pix = rng.choice(W * H, 40, replace=False)
    lib = oracle.lib()
    with np.errstate(over="ignore"):
        for px in pix:
            key_cam = np.uint32(lib.orc_path_key(seed, int(px), p))
            outs = []
            for j in range(L):
                key = key_cam if j == 0 else _pcg(np.uint32(key_cam ^ np.uint32(j * 0x68E31DA4 & 0xFFFFFFFF)))
                sl = np.array([lib.orc_slot_uniform(int(key_cam if k < 2 else key), k) for k in range(nslots)],
                              np.float32)
                o, _ = s.render_paths(0, [px], [0], slots=sl[None, :])
                outs.append(o[0])
            inv = np.float32(1) / np.float32(L)
            want = np.zeros(3, np.float32)
            for o in outs:
                want = (want + o[:3] * inv).astype(np.float32)
            got = acc[px // W, px % W]
            assert np.array_equal(got[:3], want), px
            assert got[3] == 1.0
    # L paths per pixel, the camera ray counted once per pixel
    assert st["paths"] == W * H * L
    one, st1 = oracle.Scene(dict(box_room_export)).render(seed, p, 1)
    assert st["ext_rays"] > st1["ext_rays"]


def test_first_pass_has_one_path(oracle, box_room_export):
    e = dict(box_room_export)
    e["lightsamples"] = 5
    a, st = oracle.Scene(e).render(4, 0, 1)
    b, st1 = oracle.Scene(dict(box_room_export)).render(4, 0, 1)
    assert np.array_equal(a, b) and st == st1
