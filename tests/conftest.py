import ctypes as C
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def lib():
    from aurora_amd import _lib as L
    return L.load_library()


def host_renderer():
    """A host-only context (device -1): parses and flattens, runs no GPU work."""
    from aurora_amd import _lib as L
    from aurora_amd.renderer import Renderer
    r = Renderer.__new__(Renderer)
    r.lib = L.load_library()
    r._ctx = C.c_void_p()
    L.check(r.lib.aur_create(C.byref(r._ctx), -1))
    r.passes_done = 0
    return r


@pytest.fixture(scope="session")
def box_room_dir(tmp_path_factory):
    from aurora_amd import scenes
    d = str(tmp_path_factory.mktemp("box_room"))
    scenes.write_config("box_room", d)
    return d


@pytest.fixture(scope="session")
def box_room_export(box_room_dir):
    r = host_renderer()
    r.load(os.path.join(box_room_dir, "scene.asc"), os.path.join(box_room_dir, "config.json"))
    e = r.export()
    r.close()
    return e


@pytest.fixture(scope="session")
def gpu_box_room(box_room_dir):
    from aurora_amd.renderer import Renderer
    r = Renderer(os.path.join(box_room_dir, "scene.asc"), os.path.join(box_room_dir, "config.json"))
    r.build()
    yield r
    r.close()


def random_rays(rng, n, lo, hi, tmin=1e-4, tmax=1e7):
    o = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, np.full((n, 1), tmin, np.float32), d.astype(np.float32),
                           np.full((n, 1), tmax, np.float32)], 1)
