"""Specular (MirrorBrdf) lobes of the oracle against closed forms:
SpecMirror and SpecTransmit (specMirror.cpp:30-35, specTransmit.cpp:30-58).

Both go to tangent space through worldToTangent, whose coordinateSystem frame
is not normalised (its two tangents have length s = |N x Y|, or 1 for N = +Y
exactly), and come back through tangentToWorld, whose frame is.  So for
N = +Y the mirror direction is the textbook reflection and the transmitted one
obeys Snell's law; for other normals the tangential part is scaled by s:

    mirror:    R ~ (V.N) N - s (V - (V.N) N)
    transmit:  T ~ -cos_t N - eta s (V - (V.N) N),  eta = 1 / ior,
               cos_t = sqrt(1 - eta^2 (1 - (V.N)^2))   (pdf 0 when that is < 0)

A normalised worldToTangent, a swapped tangent or a wrong sign fails these.
"""
import math

import numpy as np
import pytest

MIRROR, TRANSMIT = 2, 3


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def _frame_scale(n):
    if n[0] == 0 and n[1] != 0 and n[2] == 0:
        return 1.0
    return math.sqrt(float(n[0]) ** 2 + float(n[2]) ** 2)


def _cases(rng, k):
    out = []
    for i in range(k):
        n = _unit(rng.normal(size=3)) if i % 4 else np.array([0.0, 1.0, 0.0])
        v = _unit(rng.normal(size=3))
        if v @ n < 0:
            v = -v
        out.append((n.astype(np.float32), v.astype(np.float32)))
    return out


def test_mirror_direction(oracle):
    rng = np.random.default_rng(5)
    for n, v in _cases(rng, 400):
        d, pdf, col = oracle.bsdf_sample(MIRROR, [0.9, 0.8, 0.7], 1.0, 0, v, n, 0.3, 0.6)
        n64, v64 = n.astype(np.float64), v.astype(np.float64)
        vn = v64 @ n64
        want = _unit(vn * n64 - _frame_scale(n) * (v64 - vn * n64))
        assert np.max(np.abs(d - want)) < 2e-6, (n, v)
        assert pdf == 1.0 and np.array_equal(col, np.float32([0.9, 0.8, 0.7]))
        if n[1] == 1.0:  # the textbook reflection
            assert np.max(np.abs(d - (2 * vn * n64 - v64))) < 2e-6
        f, p = oracle.bsdf_eval(MIRROR, [0.9, 0.8, 0.7], 1.0, 0, v, n, d)
        assert p == 0.0 and not f.any()


@pytest.mark.parametrize("ior", [1.5, 1.33, 0.5])
def test_transmit_direction_and_total_internal_reflection(oracle, ior):
    rng = np.random.default_rng(6)
    eta = np.float32(1.0) / np.float32(ior)
    n_tir = 0
    for n, v in _cases(rng, 400):
        d, pdf, col = oracle.bsdf_sample(TRANSMIT, [0.5, 0.6, 0.7], ior, 0, v, n, 0.1, 0.2)
        n64, v64 = n.astype(np.float64), v.astype(np.float64)
        vn = v64 @ n64
        s2 = float(eta) ** 2 * (1 - vn * vn)
        if s2 >= 1 - 1e-6:
            n_tir += s2 >= 1 + 1e-6 and pdf == 0.0
            if s2 >= 1 + 1e-6:
                assert pdf == 0.0
            continue
        cost = math.sqrt(1 - s2)
        want = _unit(-cost * n64 - float(eta) * _frame_scale(n) * (v64 - vn * n64))
        assert pdf == 1.0 and np.array_equal(col, np.float32([0.5, 0.6, 0.7]))
        assert np.max(np.abs(d - want)) < 3e-6, (n, v)
        if n[1] == 1.0:  # Snell's law, into the surface
            sin_i = math.sqrt(max(0.0, 1 - vn * vn))
            sin_t = math.sqrt(max(0.0, 1 - (d @ n64) ** 2))
            assert abs(sin_t - sin_i / ior) < 3e-6 and d @ n64 < 0
    if ior < 1:
        assert n_tir > 50  # grazing views beyond the critical angle terminate
