"""GPU: a render does not depend on how it is cut into chunks or launched.
Paths draw their random numbers from (seed, pixel, pass, slot) and passes are
resolved in pass order, so the chunk size (pool_paths), the traversal
workgroup size and the persistent grid change the time, not the image."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_chunk_size_and_launch_shape_do_not_change_the_image(gpu_box_room):
    r = gpu_box_room
    info = r.info()
    wh = info["width"] * info["height"]
    before = r.get_launch()
    r.set_seed(41)
    ref, st = r.render_passes(0, 6)
    try:
        for pool, block, waves in ((wh, 64, 8), (2 * wh + 100, 256, 16), (64 << 20, 128, 32)):
            r.set_launch(trace_block=block, trace_waves_per_cu=waves, pool_paths=pool)
            r.set_seed(41)
            img, st2 = r.render_passes(0, 6)
            assert np.array_equal(img, ref), (pool, block, waves)
            assert st2["paths"] == st["paths"] and st2["ext_rays"] == st["ext_rays"]
            assert st2["shadow_rays"] == st["shadow_rays"] and st2["hit_rays"] == st["hit_rays"]
    finally:
        r.set_launch(**before)


def test_pass_ranges_compose(gpu_box_room):
    """Passes [0, 2) then [2, 5) accumulate to the same framebuffer as [0, 5)
    in one call."""
    r = gpu_box_room
    r.set_seed(8)
    one, _ = r.render_passes(0, 5)
    r.set_seed(8)
    r.render_passes(0, 2)
    two, _ = r.render_passes(2, 3)
    assert np.array_equal(one, two)


def test_two_pipelines_match_one_pipeline(tmp_path):
    """A render of at least 2 M paths traces its chunks on two path pipelines at
    once, each chunk's resolve waiting for the previous chunk's; with kernel
    timing on it runs on one pipeline and one stream.  Framebuffer, FP16
    display and ray counts agree bit for bit, with one chunk per pipeline, an
    even and an odd number of chunks, and one pass per chunk."""
    import os

    from aurora_amd import scenes
    from aurora_amd.renderer import Renderer
    path = scenes.write_config("box_room", str(tmp_path), res=256, spp=40)
    r = Renderer(path, os.path.join(str(tmp_path), "config.json"))
    try:
        r.build()
        wh = 256 * 256
        print("two pipelines: box room 256x256, 40 passes, one pipeline first", flush=True)
        r.set_profiling(True)
        r.set_seed(5)
        ref, st = r.render_passes(0, 40)
        dref = r.display()
        r.set_profiling(False)
        assert st["paths"] == 40 * wh >= 1 << 21
        # chunks per render: 2, 10, 7 (six passes each, the last four) and 40
        for pool in (64 << 20, 8 * wh, 2 * (6 * wh + 7), wh):
            print("two pipelines: pool", pool, flush=True)
            r.set_launch(pool_paths=pool)
            r.set_seed(5)
            img, st2 = r.render_passes(0, 40)
            assert np.array_equal(img, ref), pool
            assert np.array_equal(r.display(), dref), pool
            for key in ("paths", "ext_rays", "hit_rays", "shadow_rays"):
                assert st2[key] == st[key], (pool, key)
    finally:
        r.close()
