"""GPU: the reference display's FP16 running mean (OpenexrDisplay::
appendValue, openexrDisplay.cpp:61-72), kept beside the FP32 sum.

Each pass's per-pixel value is taken from a fresh context that renders that
pass alone (its framebuffer starts at zero, so it holds the value exactly);
the running mean restated in numpy (float32 arithmetic, FP16 storage rounded
to nearest even) must equal the device's bit for bit, for per-scene renders
in one call or several, for the batched launch, and in the written EXR."""
import os

import numpy as np
import pytest


def _running_mean(values):
    d = np.zeros(values[0].shape, np.float16)
    for k, v in enumerate(values):
        w = np.float32(1.0) / np.float32(k + 1)
        d = (d.astype(np.float32) * (np.float32(1.0) - w) + v.astype(np.float32) * w).astype(np.float16)
    return d


def test_running_mean_closed_forms():
    """Pins the restatement above (CPU): a constant stays exact after every
    update (a 1/(n+1) weight halves the first value), and the values 1..n
    give the arithmetic mean (n+1)/2 within FP16 rounding (swapped weights
    track the newest value instead)."""
    for c in (0.375, 1.0, 3.5, 1024.0):
        vals = [np.full(4, c, np.float32)] * 9
        for n in range(1, 10):
            assert np.all(_running_mean(vals[:n]) == np.float16(c))
    for n in range(1, 9):
        vals = [np.full(4, k, np.float32) for k in range(1, n + 1)]
        got = _running_mean(vals).astype(np.float64)
        assert np.allclose(got, (n + 1) / 2.0, rtol=2e-3, atol=0), (n, got[0])
    # one update from zero is the value itself, rounded to FP16
    v = np.array([0.1, 1e-5, 70000.0, -2.5e-8], np.float32)
    with np.errstate(over="ignore"):  # 70000 overflows FP16 to inf, as Imf::Rgba's half does
        assert np.array_equal(_running_mean([v]), v.astype(np.float16))


@pytest.fixture(scope="module")
def room(tmp_path_factory):
    from aurora_amd import scenes
    d = str(tmp_path_factory.mktemp("disp"))
    scenes.write_config("box_room", d, res=48, spp=5, maxdepth=4, mindepth=4)
    return os.path.join(d, "scene.asc"), os.path.join(d, "config.json")


@pytest.mark.gpu
def test_display_is_the_reference_running_mean(room, tmp_path):
    from aurora_amd.exr import read_exr
    from aurora_amd.renderer import Renderer, batch_submit
    P = 5
    values = []
    for k in range(P):
        rk = Renderer(*room)
        rk.build()
        rk.set_seed(13)
        v, _ = rk.render_passes(k, 1)
        values.append(v)
        rk.close()
    want = _running_mean(values)
    r = Renderer(*room)
    r.build()
    r.set_seed(13)
    r.render_passes(0, P)
    assert np.array_equal(r.display().view(np.uint16), want.view(np.uint16))
    # several calls in pass order
    r.set_seed(13)
    r.render_passes(0, 2)
    r.render_passes(2, 3)
    assert np.array_equal(r.display().view(np.uint16), want.view(np.uint16))
    # the batched launch keeps the same display
    ex = Renderer()
    r.set_seed(13)
    batch_submit(ex, [(r, 0, P)])
    assert np.array_equal(r.display().view(np.uint16), want.view(np.uint16))
    # the written file holds it
    path = str(tmp_path / "disp.exr")
    r.write_image(path, 0)
    img = read_exr(path)
    assert np.array_equal(np.asarray(img, np.float32).astype(np.float16).view(np.uint16), want.view(np.uint16))
    ex.close()
    r.close()


@pytest.mark.gpu
def test_display_counts_passes_since_cleared(room):
    """A fresh framebuffer that starts at pass 3 holds the running mean of
    passes 3, 4, 5 with weights 1, 1/2, 1/3 (appendValue counts the values
    appended since the display was cleared), in one call, in two calls, and
    through the batched launch."""
    from aurora_amd.renderer import Renderer, batch_submit
    values = []
    for k in range(3, 6):
        rk = Renderer(*room)
        rk.build()
        rk.set_seed(17)
        v, _ = rk.render_passes(k, 1)
        values.append(v)
        rk.close()
    want = _running_mean(values)
    r = Renderer(*room)
    r.build()
    r.set_seed(17)
    r.render_passes(3, 3)
    assert np.array_equal(r.display().view(np.uint16), want.view(np.uint16))
    r2 = Renderer(*room)
    r2.build()
    r2.set_seed(17)
    r2.render_passes(3, 1)
    r2.render_passes(4, 2)
    assert np.array_equal(r2.display().view(np.uint16), want.view(np.uint16))
    r3 = Renderer(*room)
    r3.build()
    r3.set_seed(17)
    batch_submit(r, [(r3, 3, 3)])
    assert np.array_equal(r3.display().view(np.uint16), want.view(np.uint16))
    for x in (r, r2, r3):
        x.close()
