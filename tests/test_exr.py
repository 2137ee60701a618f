"""OpenEXR codec: FP16 rounding against numpy's IEEE binary16, round trips for
every supported compression and pixel type (including partial 16-line ZIP
chunks), and an independent decoder written from the published file layout
(header attribute list, chunk offset table, ZIP = zlib over a byte-delta
predictor over even/odd-split bytes) that must read what the writer wrote."""
import os
import struct
import zlib

import numpy as np
import pytest

from conftest import ROOT  # noqa: F401


def _parse(path):
    """Independent reader: returns (header dict, {channel: (H, W) array})."""
    data = open(path, "rb").read()
    magic, version = struct.unpack_from("<ii", data, 0)
    assert magic == 20000630 and (version & 0xFF) == 2 and not (version & 0x200)
    pos, hdr = 8, {}
    while True:
        end = data.index(b"\0", pos)
        name = data[pos:end].decode()
        pos = end + 1
        if not name:
            break
        end = data.index(b"\0", pos)
        typ = data[pos:end].decode()
        pos = end + 1
        (size,) = struct.unpack_from("<i", data, pos)
        pos += 4
        hdr[name] = (typ, data[pos:pos + size])
        pos += size
    chans, p = [], 0
    raw = hdr["channels"][1]
    while raw[p] != 0:
        e = raw.index(b"\0", p)
        cname = raw[p:e].decode()
        ptype, _, xs, ys = struct.unpack_from("<iiii", raw, e + 1)
        assert xs == 1 and ys == 1
        chans.append((cname, ptype))
        p = e + 1 + 16
    assert [c for c, _ in chans] == sorted(c for c, _ in chans)
    comp = hdr["compression"][1][0]
    x0, y0, x1, y1 = struct.unpack("<iiii", hdr["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    lpc = 16 if comp == 3 else 1
    n_chunks = (h + lpc - 1) // lpc
    offs = struct.unpack_from("<%dQ" % n_chunks, data, pos)
    out = {c: np.zeros((h, w), np.float32) for c, _ in chans}
    bpp = sum(2 if t == 1 else 4 for _, t in chans)
    for off in offs:
        y, n = struct.unpack_from("<ii", data, off)
        blob = data[off + 8: off + 8 + n]
        lines = min(lpc, h - (y - y0))
        size = lines * w * bpp
        if n < size:
            t = bytearray(zlib.decompress(blob))
            for i in range(1, len(t)):
                t[i] = (t[i - 1] + t[i] - 128) & 0xFF
            half = (len(t) + 1) // 2
            blob = bytes(b for pair in zip(t[:half], t[half:] + b"\0") for b in pair)[:len(t)]
        q = 0
        for ly in range(lines):
            for c, t in chans:
                dt = np.float16 if t == 1 else np.float32
                k = 2 if t == 1 else 4
                out[c][y - y0 + ly] = np.frombuffer(blob[q:q + k * w], dt).astype(np.float32)
                q += k * w
    return hdr, out


def test_half_rounding_matches_ieee_binary16(tmp_path):
    from aurora_amd import exr
    rng = np.random.default_rng(1)
    v = np.concatenate([rng.normal(0, 1, 5000) * 10.0 ** rng.integers(-9, 6, 5000),
                        [0.0, -0.0, 65504.0, 65520.0, 1e6, -1e6, 5.96e-8, 2.98e-8, 2.99e-8, 6.1e-5, 1.0 + 2 ** -11,
                         1.0 + 3 * 2 ** -11, np.inf, -np.inf]]).astype(np.float32)
    v = v[: (v.size // 4) * 4]
    img = v.reshape(1, -1, 4)
    p = str(tmp_path / "h.exr")
    exr.write_exr(p, img, compression="none", half=True)
    _, ch = _parse(p)
    with np.errstate(over="ignore"):  # values past 65504 round to inf on purpose
        want = img.astype(np.float16).astype(np.float32)
    for k, c in enumerate("RGBA"):
        got = ch[c][0]
        assert np.array_equal(got.view(np.uint32), want[0, :, k].view(np.uint32)), c


@pytest.mark.parametrize("comp", ["none", "zips", "zip"])
@pytest.mark.parametrize("half", [True, False])
def test_round_trip_and_independent_decoder(tmp_path, comp, half):
    from aurora_amd import exr
    rng = np.random.default_rng(7)
    img = rng.uniform(0, 4, (37, 19, 4)).astype(np.float32)  # 37 rows: a partial 16-line ZIP chunk
    img[5:9, 3:7] = 0.25                                       # compressible runs
    p = str(tmp_path / ("i_%s_%d.exr" % (comp, half)))
    exr.write_exr(p, img, compression=comp, half=half, meta={"renderer": "test", "RaySpeed": "12"})
    want = img.astype(np.float16).astype(np.float32) if half else img
    back = exr.read_exr(p)
    assert back.shape == img.shape and np.array_equal(back, want)
    hdr, ch = _parse(p)
    assert hdr["compression"][1][0] == {"none": 0, "zips": 2, "zip": 3}[comp]
    assert hdr["renderer"] == ("string", b"test") and hdr["RaySpeed"] == ("string", b"12")
    for k, c in enumerate("RGBA"):
        assert np.array_equal(ch[c], want[..., k])
    assert not os.path.exists(p + "_tmp")


def test_errors(tmp_path):
    from aurora_amd import exr, _lib as L
    with pytest.raises(L.AuroraError) as ei:
        exr.read_exr(str(tmp_path / "missing.exr"))
    assert ei.value.code == -2
    p = str(tmp_path / "t.exr")
    exr.write_exr(p, np.ones((4, 4, 4), np.float32), compression="zip")
    open(p, "r+b").truncate(60)
    with pytest.raises(L.AuroraError) as ei:
        exr.read_exr(p)
    assert ei.value.code == -2
    with pytest.raises(L.AuroraError) as ei:
        L.check(L.load_library().aur_exr_write(p.encode(), 2, 2, L.fptr(np.ones(16, np.float32)), 1, 1, None, None, 0))
    assert ei.value.code == -1
