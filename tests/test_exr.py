"""OpenEXR output of SH lightmaps (shb_save_lightmap; reference
src/saver.cpp:103-422): the three layouts, HALF storage, ZIP blocks.

Pinned by: float->half against numpy's float16 conversion (round to nearest
even) on every half, every midpoint between neighbouring halves and the
special values; the ZIP block bytes against the layout the file format
specifies (even/odd byte split, byte deltas + 128) on a block built by hand;
the header attributes against the format's fixed encodings. Runs on the CPU
(the saver is host code; no device context is needed)."""
import os
import struct
import zlib

import numpy as np
import pytest

import shb
from shb import exr

NAMES = ["L0", "L1m1", "L10", "L11", "L2m2", "L2m1", "L20", "L21", "L22"]


def half_bits(x):
    return np.asarray(x, np.float32).astype(np.float16).view(np.uint16)


def test_half_conversion_matches_numpy(tmp_path):
    # every finite half, the midpoints between neighbours (ties to even),
    # values just either side of the midpoints, subnormals, overflow, inf, nan
    h = np.arange(0, 0x7c00, dtype=np.uint16).view(np.float16).astype(np.float32)
    mids = ((h[:-1].astype(np.float64) + h[1:]) / 2).astype(np.float32)
    near = np.concatenate([np.nextafter(mids, np.inf), np.nextafter(mids, -np.inf)])
    extra = np.array([65504, 65519.99, 65520, 1e9, np.inf, 2.0 ** -25, 2.0 ** -26,
                      np.nextafter(np.float32(2.0 ** -25), np.float32(1)), 1e-30, 0.0],
                     np.float32)
    vals = np.concatenate([h, mids, near, extra])
    vals = np.concatenate([vals, -vals]).astype(np.float32)
    w = 4096
    n = ((vals.size + w - 1) // w) * w
    plane = np.zeros(n, np.float32)
    plane[:vals.size] = vals
    sh = np.zeros((n, 27), np.float32)
    sh[:, 0] = plane
    path = str(tmp_path / "h.exr")
    shb.save_lightmap(path, sh, None, w, n // w)
    _, ch = exr.read_exr(path)
    got = ch["L0.R"].ravel().astype(np.float16).view(np.uint16)
    with np.errstate(over="ignore"):
        exp = half_bits(plane)
    assert np.array_equal(got, exp)
    nanpath = str(tmp_path / "nan.exr")
    sh[:4, 0] = [np.nan, -np.nan, np.inf, -np.inf]
    shb.save_lightmap(nanpath, sh[:w], None, w, 1)
    r = exr.read_exr(nanpath)[1]["L0.R"].ravel()
    assert np.isnan(r[:2]).all() and r[2] == np.inf and r[3] == -np.inf


def test_zip_block_layout_from_the_format(tmp_path):
    # one channel-row pattern written by hand: halves 0x3c00 (1.0) and 0xc000
    # (-2.0) alternating, 8 pixels x 2 rows of an RGB file (split mode, L1m1)
    w, hgt = 8, 2
    sh = np.zeros((w * hgt, 27), np.float32)
    sh[:, 3] = np.tile([1.0, -2.0], w * hgt // 2)
    base = str(tmp_path / "b.exr")
    shb.save_lightmap(base, sh, None, w, hgt, shb.SAVE_SPLIT)
    buf = open(str(tmp_path / "b_L1m1.exr"), "rb").read()
    hdr, _ = exr.read_exr(str(tmp_path / "b_L1m1.exr"))
    # single block: offset table of one entry right after the header
    end = buf.index(b"screenWindowWidth\0float\0") + len(b"screenWindowWidth\0float\0") + 8 + 1
    (off,) = struct.unpack_from("<Q", buf, end)
    y, size = struct.unpack_from("<iI", buf, off)
    assert y == 0
    # uncompressed block: per scanline, channels in name order B, G, R, each
    # a row of little-endian halves
    rows = []
    for r in range(hgt):
        for c in ("B", "G", "R"):
            v = sh[r * w:(r + 1) * w, 3 + "RGB".index(c)]
            rows.append(half_bits(v).astype("<u2").tobytes())
    raw = np.frombuffer(b"".join(rows), np.uint8)
    split = np.concatenate([raw[0::2], raw[1::2]])
    pred = split.astype(np.int64)
    pred[1:] = (split[1:].astype(np.int64) - split[:-1] + 128) % 256
    block = buf[off + 8: off + 8 + size]
    if size < raw.size:
        assert np.array_equal(np.frombuffer(zlib.decompress(block), np.uint8), pred.astype(np.uint8))
    else:
        assert block == raw.tobytes()
    assert hdr["compression"] == 3


def test_header_attributes(tmp_path):
    path = str(tmp_path / "a.exr")
    sh = np.ones((5 * 3, 27), np.float32)
    shb.save_lightmap(path, sh, np.ones(15, np.float32), 5, 3)
    buf = open(path, "rb").read()
    assert struct.unpack_from("<II", buf, 0) == (20000630, 2)
    hdr, ch = exr.read_exr(path)
    names = [c[0] for c in hdr["channels"]]
    assert names == sorted([f"{n}.{s}" for n in NAMES for s in "RGB"] + ["EnvVisibility"])
    assert all(c[1:] == (1, 1, 1) for c in hdr["channels"])  # HALF, 1x1 sampling
    assert hdr["dataWindow"] == (0, 0, 4, 2) and hdr["displayWindow"] == (0, 0, 4, 2)
    assert hdr["lineOrder"] == 0
    assert hdr["pixelAspectRatio"] == ("float", struct.pack("<f", 1.0))
    assert hdr["screenWindowCenter"] == ("v2f", struct.pack("<ff", 0.0, 0.0))
    assert hdr["screenWindowWidth"] == ("float", struct.pack("<f", 1.0))


@pytest.fixture
def lightmap():
    rng = np.random.default_rng(11)
    w, h = 45, 37  # 3 ZIP blocks, the last one partial
    sh = (rng.normal(size=(w * h, 27)) * rng.uniform(0.01, 40, (w * h, 1))).astype(np.float32)
    sh[::7] = 0.0  # compressible stretches next to noise
    vis = rng.random(w * h).astype(np.float32)
    return w, h, sh, vis


def test_combined_round_trip(tmp_path, lightmap):
    w, h, sh, vis = lightmap
    path = str(tmp_path / "lm.exr")
    shb.save_lightmap(path, sh, vis, w, h, shb.SAVE_COMBINED)
    assert os.listdir(tmp_path) == ["lm.exr"]
    _, ch = exr.read_exr(path)
    assert len(ch) == 28
    for i, n in enumerate(NAMES):
        for c, s in enumerate("RGB"):
            assert np.array_equal(half_bits(ch[f"{n}.{s}"].ravel()), half_bits(sh[:, 3 * i + c])), n + s
    assert np.array_equal(half_bits(ch["EnvVisibility"].ravel()), half_bits(vis))
    # visibility absent -> zeros (saver.cpp:166-172)
    shb.save_lightmap(path, sh, None, w, h)
    assert not exr.read_exr(path)[1]["EnvVisibility"].any()


def test_split_round_trip(tmp_path, lightmap):
    w, h, sh, vis = lightmap
    shb.save_lightmap(str(tmp_path / "lm.exr"), sh, vis, w, h, shb.SAVE_SPLIT)
    assert sorted(os.listdir(tmp_path)) == sorted(f"lm_{n}.exr" for n in NAMES)
    for i, n in enumerate(NAMES):
        hdr, ch = exr.read_exr(str(tmp_path / f"lm_{n}.exr"))
        assert sorted(ch) == (["A", "B", "G", "R"] if i == 0 else ["B", "G", "R"])
        for c, s in enumerate("RGB"):
            assert np.array_equal(half_bits(ch[s].ravel()), half_bits(sh[:, 3 * i + c]))
        if i == 0:
            assert np.array_equal(half_bits(ch["A"].ravel()), half_bits(vis))


def test_packed_luminance(tmp_path, lightmap):
    w, h, sh, vis = lightmap
    shb.save_lightmap(str(tmp_path / "out.exr"), sh, vis, w, h, shb.SAVE_LUMINANCE_PACKED)
    assert sorted(os.listdir(tmp_path)) == [f"out_packed_{i}.exr" for i in range(3)]
    files = [exr.read_exr(str(tmp_path / f"out_packed_{i}.exr"))[1] for i in range(3)]
    for f in files:
        assert sorted(f) == ["A", "B", "G", "R"]
    for c, s in enumerate("RGB"):
        assert np.array_equal(half_bits(files[0][s].ravel()), half_bits(sh[:, c]))
    assert np.array_equal(half_bits(files[0]["A"].ravel()), half_bits(vis))
    # Rec. 709 luma of coefficients 1..8, four per file in R, G, B, A; against
    # a float64 luma: within one half rounding of the float32 sum
    c = sh.reshape(-1, 9, 3).astype(np.float64)
    y = c[:, :, 0] * 0.2126 + c[:, :, 1] * 0.7152 + c[:, :, 2] * 0.0722
    for k in range(1, 9):
        got = files[1 + (k - 1) // 4]["RGBA"[(k - 1) % 4]].ravel().astype(np.float64)
        tol = np.abs(y[:, k]) * (2.0 ** -11 + 1e-6) + 2.0 ** -24
        assert (np.abs(got - y[:, k]) <= tol).all(), k
    # a luma weight typed wrong would miss: pure-red, pure-green and pure-blue
    # coefficients give the three weights themselves
    one = np.zeros((3, 27), np.float32)
    for k in range(3):
        one[k, 3 + k] = 1.0
    shb.save_lightmap(str(tmp_path / "w.exr"), one, None, 3, 1, shb.SAVE_LUMINANCE_PACKED)
    r = exr.read_exr(str(tmp_path / "w_packed_1.exr"))[1]["R"].ravel()
    assert np.array_equal(half_bits(r), half_bits([0.2126, 0.7152, 0.0722]))


def test_save_errors(tmp_path):
    sh = np.zeros((4, 27), np.float32)
    with pytest.raises(shb.ShbError):
        shb.save_lightmap(str(tmp_path / "x.exr"), sh, None, 2, 2, mode=7)
    with pytest.raises(shb.ShbError):
        shb.save_lightmap(str(tmp_path / "missing_dir" / "x.exr"), sh, None, 2, 2)
    with pytest.raises(ValueError):
        shb.save_lightmap(str(tmp_path / "x.exr"), sh, None, 3, 2)
