"""Reading OpenEXR scanline files (NONE / ZIPS / ZIP compression, HALF /
FLOAT / UINT channels) into numpy arrays, to load lightmaps written by
shb_save_lightmap back (and to check them)."""
from __future__ import annotations

import struct
import zlib

import numpy as np

_LINES = {0: 1, 2: 1, 3: 16}  # compression -> scanlines per block
_DTYPE = {0: np.dtype("<u4"), 1: np.dtype("<f2"), 2: np.dtype("<f4")}


def _undo_zip(data: bytes, n: int) -> bytes:
    d = np.frombuffer(zlib.decompress(data), np.uint8)
    if d.size != n:
        raise ValueError(f"block inflates to {d.size} bytes, expected {n}")
    # undo the byte-delta predictor (t[i] = t[i-1] + d[i] - 128, t[0] = d[0]),
    # then re-interleave the two halves
    t = (np.cumsum(d.astype(np.int64) - 128) + 128).astype(np.uint8)
    out = np.empty(n, np.uint8)
    h = (n + 1) // 2
    out[0::2] = t[:h]
    out[1::2] = t[h:]
    return out.tobytes()


def read_exr(path):
    """Returns (header dict, {channel name: (height, width) array})."""
    buf = open(path, "rb").read()
    magic, version = struct.unpack_from("<II", buf, 0)
    if magic != 20000630:
        raise ValueError("not an OpenEXR file")
    if version & 0x200:
        raise ValueError("tiled files are not supported")
    pos = 8
    header = {}
    while buf[pos] != 0:
        end = buf.index(0, pos)
        name = buf[pos:end].decode()
        pos = end + 1
        end = buf.index(0, pos)
        typ = buf[pos:end].decode()
        pos = end + 1
        (size,) = struct.unpack_from("<I", buf, pos)
        pos += 4
        val = buf[pos:pos + size]
        pos += size
        if typ == "chlist":
            chans, q = [], 0
            while val[q] != 0:
                e = val.index(0, q)
                cname = val[q:e].decode()
                ptype, _, xs, ys = struct.unpack_from("<iB3xii", val, e + 1)
                chans.append((cname, ptype, xs, ys))
                q = e + 1 + 16
            header[name] = chans
        elif typ == "compression" or typ == "lineOrder":
            header[name] = val[0]
        elif typ == "box2i":
            header[name] = struct.unpack("<4i", val)
        else:
            header[name] = (typ, val)
    pos += 1
    comp = header["compression"]
    if comp not in _LINES:
        raise ValueError(f"compression {comp} is not supported")
    x0, y0, x1, y1 = header["dataWindow"]
    w, h = x1 - x0 + 1, y1 - y0 + 1
    chans = header["channels"]
    nl = _LINES[comp]
    nblocks = (h + nl - 1) // nl
    offsets = struct.unpack_from(f"<{nblocks}Q", buf, pos)
    out = {c[0]: np.zeros((h, w), np.float32) for c in chans}
    for off in offsets:
        y, size = struct.unpack_from("<iI", buf, off)
        data = buf[off + 8: off + 8 + size]
        rows = min(nl, y1 + 1 - y)
        n = sum(rows * w * _DTYPE[c[1]].itemsize for c in chans)
        raw = data if size == n else _undo_zip(data, n)
        q = 0
        for r in range(rows):
            for cname, ptype, _, _ in chans:
                dt = _DTYPE[ptype]
                row = np.frombuffer(raw, dt, w, q)
                q += w * dt.itemsize
                out[cname][y - y0 + r] = row.astype(np.float32)
    return header, out
