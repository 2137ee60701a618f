"""Halton QMC restatement (oracle/oracle_halton.c) against the reference's own
library: the fixture tests/golden/halton_golden.npz (written by
tests/golden/make_halton_golden.py from Aurora/external/halton compiled as is)
bit for bit, and oracle/_ref/halton_ref when it has been built; plus the
properties that define the sequence (radical inverses, one point per pixel
per enumerated index)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

GOLDEN = os.path.join(ROOT, "tests", "golden", "halton_golden.npz")


def test_samples_match_reference_library(oracle):
    g = np.load(GOLDEN)
    got = np.array([oracle.halton_sample(int(d), int(i)) for d, i in zip(g["sample_dim"], g["sample_index"])],
                   np.float32)
    assert np.array_equal(got.view(np.uint32), g["sample_bits"])


def test_enumeration_matches_reference_library(oracle):
    g = np.load(GOLDEN)
    for w, h, inc, sx, sy in g["res"]:
        e = oracle.halton_enum(int(w), int(h))
        assert e["inc"] == inc
        assert np.float32(2 ** e["p2"]).view(np.uint32) == sx and np.float32(3 ** e["p3"]).view(np.uint32) == sy
    for w, h, i, x, y, idx in g["enum"]:
        assert oracle.halton_index(int(w), int(h), int(i), int(x), int(y)) == idx


def test_against_built_reference_when_present(oracle):
    exe = os.path.join(ROOT, "oracle", "_ref", "halton_ref")
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref/halton_ref not built (needs the reference sources)")
    out = subprocess.run([exe], stdout=subprocess.PIPE, text=True, check=True).stdout.split("\n")
    n = 0
    for line in out:
        f = line.split()
        if f and f[0] == "S":
            d, i, bits = int(f[1]), int(f[2]), int(f[3])
            assert np.float32(oracle.halton_sample(d, i)).view(np.uint32) == bits, (d, i)
            n += 1
    assert n == 256 * 40


def test_radical_inverse_properties(oracle):
    # base 2 and base 3 (identity permutations): the textbook radical inverse
    for i in range(2000):
        r2, r3, k, b = 0.0, 0.0, i, 0.5
        while k:
            r2 += (k & 1) * b
            k >>= 1
            b *= 0.5
        k, b = i, 1.0 / 3
        while k:
            r3 += (k % 3) * b
            k //= 3
            b /= 3
        assert oracle.halton_sample(0, i) == np.float32(r2)
        assert abs(oracle.halton_sample(1, i) - r3) < 3e-7
    # every dimension: values in [0, 1), the first b^2 points hit distinct
    # cells of width 1/b^2 (a permutation of the digits keeps the
    # stratification); the points of index < b^2 sit on the cell corners
    # (up to the 1 - 2^-22 scale and float rounding)
    for d in (2, 5, 17, 100, 255):
        pts = np.array([oracle.halton_sample(d, i) for i in range(4096)])
        assert pts.min() >= 0 and pts.max() < 1
    for d, b in ((2, 5), (3, 7), (4, 11)):
        v = np.array([oracle.halton_sample(d, i) for i in range(b * b)], np.float64) * b * b
        cells = np.round(v)
        assert np.max(np.abs(v - cells)) < 1e-4
        assert sorted(cells.astype(int).tolist()) == list(range(b * b))


def test_enumerated_points_fall_in_their_pixel(oracle):
    for w, h in ((7, 5), (64, 36), (48, 48)):
        e = oracle.halton_enum(w, h)
        for i in range(3):
            for x in range(0, w, max(1, w // 7)):
                for y in range(0, h, max(1, h // 5)):
                    u = oracle.halton_slot(w, h, i, x, y, 0)
                    v = oracle.halton_slot(w, h, i, x, y, 1)
                    # inside the pixel, up to the 1 - 2^-22 scale at a cell corner
                    assert -1e-5 < u < 1 and -1e-5 < v < 1, (w, h, i, x, y, u, v)
        assert e["inc"] == 2 ** e["p2"] * 3 ** e["p3"]


def test_oracle_halton_paths_read_the_halton_slots(oracle, box_room_export):
    """With the Halton source the oracle's paths consume exactly the slot
    values orc_halton_slot gives (camera offsets, then dimensions 2..)."""
    e = dict(box_room_export)
    W, H = e["width"], e["height"]
    e["sampler"] = "halton"
    s = oracle.Scene(e)
    rng = np.random.default_rng(3)
    pix = rng.integers(0, W * H, 300).astype(np.int32)
    smp = rng.integers(0, 40, 300).astype(np.int32)
    nslots = 2 + 9 * e["maxdepth"]
    slots = np.array([[oracle.halton_slot(W, H, int(i), int(p) % W, int(p) // W, k) for k in range(nslots)]
                      for p, i in zip(pix, smp)], np.float32)
    a, _ = s.render_paths(0, pix, smp)
    b, _ = s.render_paths(0, pix, smp, slots=slots)
    assert np.array_equal(a, b)
    # the seed plays no part
    c, _ = s.render_paths(12345, pix, smp)
    assert np.array_equal(a, c)
