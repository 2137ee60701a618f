"""GPU: aur_trace_noalloc (device buffers in and out, no device allocation,
the traversal stack's overflow in the library's fixed per-GPU scratch)
against aur_trace on the same rays: hit ids, t, u, v bit for bit, closest and
any hit, on two calls in a row with the second one smaller, on the box room
(4- and 8-wide BVHs) and on subdivided spheres (a deeper BVH)."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import random_rays

pytestmark = pytest.mark.gpu

HIT = [("t", np.float32), ("u", np.float32), ("v", np.float32), ("prim", np.int32)]


def _hip():
    """The HIP runtime the library itself is bound to (device buffers for the
    test).  Opening it by its soname returns the copy already loaded for the
    library -- torch ships a runtime of the same soname, and a second runtime
    in the process would not share the library's device state."""
    from aurora_amd import _lib as L
    L.load_library()
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    return h


def _trace_noalloc(r, rays, mode):
    from aurora_amd import _lib as L
    hip = _hip()
    rays = np.ascontiguousarray(rays, np.float32)
    n = rays.shape[0]
    d_rays, d_hits = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(d_rays), rays.nbytes) == 0
    assert hip.hipMalloc(C.byref(d_hits), 16 * n) == 0
    try:
        stale = np.full((n, 4), -7.0, np.float32)
        assert hip.hipMemcpy(d_rays, rays.ctypes.data, rays.nbytes, 1) == 0
        assert hip.hipMemcpy(d_hits, stale.ctypes.data, stale.nbytes, 1) == 0
        assert hip.hipDeviceSynchronize() == 0
        L.check(r.lib.aur_trace_noalloc(r._ctx, d_rays, n, d_hits, mode))
        assert hip.hipDeviceSynchronize() == 0  # the call only enqueues on the context's stream
        out = np.zeros((n, 4), np.float32)
        assert hip.hipMemcpy(out.ctypes.data, d_hits, out.nbytes, 2) == 0
    finally:
        hip.hipFree(d_rays)
        hip.hipFree(d_hits)
    return out.view(HIT).reshape(n)


def _same(a, b):
    assert np.array_equal(a["prim"], b["prim"])
    hit = b["prim"] >= 0
    for k in ("t", "u", "v"):
        assert np.array_equal(a[k][hit], b[k][hit]), k


def _check_two_calls(r, rays_a, rays_b):
    assert rays_b.shape[0] < rays_a.shape[0]
    for rays in (rays_a, rays_b):
        print("aur_trace_noalloc: %d rays" % rays.shape[0], flush=True)
        _same(_trace_noalloc(r, rays, 0), r.trace(rays))
        g, o = _trace_noalloc(r, rays, 1), r.trace(rays, any_hit=True)
        assert np.array_equal(g["prim"] >= 0, o["prim"] >= 0)


def test_noalloc_trace_matches_aur_trace_box_room(gpu_box_room):
    rng = np.random.default_rng(41)
    a = random_rays(rng, 30000, [-0.99, -0.99, -4.49], [0.99, 0.99, -2.51])
    b = random_rays(rng, 4000, [-0.99, -0.99, -4.49], [0.99, 0.99, -2.51])
    for width in (4, 8):
        print("box room, BVH width %d" % width, flush=True)
        gpu_box_room.build(width=width, max_leaf=2)
        _check_two_calls(gpu_box_room, a, b)
    gpu_box_room.build()


def test_noalloc_trace_matches_aur_trace_spheres(tmp_path):
    from aurora_amd import scenes
    from aurora_amd.renderer import Renderer
    spec = scenes.displaced_spheres(width=64, height=36, n_spheres=3, levels=5)
    r = Renderer(spec.write(str(tmp_path)))
    r.build()
    rng = np.random.default_rng(42)
    a = random_rays(rng, 50000, [-6, -1, -12], [6, 2, -2])
    b = random_rays(rng, 7000, [-6, -1, -12], [6, 2, -2])
    assert np.mean(r.trace(a)["prim"] >= 0) > 0.1
    _check_two_calls(r, a, b)
    r.close()


def test_noalloc_trace_refuses_what_it_cannot_do(gpu_box_room):
    from aurora_amd import _lib as L
    rays = random_rays(np.random.default_rng(1), 8, [-0.5, -0.5, -4], [0.5, 0.5, -3])
    with pytest.raises(L.AuroraError) as ei:
        _trace_noalloc(gpu_box_room, rays, 2)  # the extension kernel needs the path queues
    assert ei.value.code == -1
    assert gpu_box_room.lib.aur_trace_noalloc(gpu_box_room._ctx, None, -1, None, 0) == -1
