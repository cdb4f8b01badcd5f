"""shb_env_sh_latlong_noalloc: the environment SH projection with its
partial sums in the library's fixed device scratch (no allocation in the
call). Checked through the C ABI against shb_env_sh_latlong on the same map
(same kernels and summation order: bit for bit), on two calls in a row, the
second with a smaller map, so the second reads a scratch the first left
filled. Device buffers come from the HIP runtime the library itself uses."""
import ctypes as C

import numpy as np
import pytest

import shb

pytestmark = pytest.mark.gpu

_hip = C.CDLL("libamdhip64.so.7")
_hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
_hip.hipFree.argtypes = [C.c_void_p]
_hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
_hip.hipDeviceSynchronize.argtypes = []
H2D, D2H = 1, 2


class DevBuf:
    def __init__(self, nbytes):
        self.p = C.c_void_p()
        assert _hip.hipMalloc(C.byref(self.p), nbytes) == 0
        self.n = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.n
        assert _hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, n, dtype):
        out = np.zeros(n, dtype)
        assert _hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        _hip.hipFree(self.p)


def sky(h, w, ch, seed):
    rng = np.random.default_rng(seed)
    th = ((np.arange(h) + 0.5) * np.pi / h).astype(np.float32)[:, None]
    ph = ((np.arange(w) + 0.5) * 2 * np.pi / w).astype(np.float32)[None, :]
    base = (1.0 + np.cos(th) * (1.5 + 0.5 * np.sin(3 * ph))).astype(np.float32)
    px = np.empty((h, w, ch), np.float32)
    for c in range(ch):
        px[..., c] = (0.5 + 0.2 * c) * base
    return px * rng.uniform(0.8, 1.2, (h, w, 1)).astype(np.float32)


def test_noalloc_matches_allocating_entry_point_twice(baker):
    print("no-alloc env SH: 2048x1024x3 then 512x256x4", flush=True)
    big, small = sky(1024, 2048, 3, 1), sky(256, 512, 4, 2)
    d_px = DevBuf(big.nbytes)
    d_out = DevBuf(27 * 4)
    try:
        for px in (big, small):
            h, w, ch = px.shape
            ref, _ = baker.env_sh_latlong(px)
            d_px.put(px)
            d_out.put(np.full(27, np.nan, np.float32))
            baker.env_sh_latlong_noalloc(d_px.p.value, w, h, ch, d_out.p.value)
            got = d_out.get(27, np.float32)
            assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), (got, ref)
    finally:
        d_px.free()
        d_out.free()


def test_noalloc_rejects_bad_buffers_and_oversized_maps(baker):
    d_out = DevBuf(27 * 4)
    host = np.zeros(64 * 32 * 3, np.float32)
    try:
        # a host pointer is refused before any launch
        with pytest.raises(shb.ShbError) as e:
            baker.env_sh_latlong_noalloc(host.ctypes.data, 64, 32, 3, d_out.p.value)
        assert e.value.code == -1
        # a map whose partial sums exceed the fixed scratch: refused by size
        # alone (the pointer is never read)
        with pytest.raises(shb.ShbError) as e:
            baker.env_sh_latlong_noalloc(d_out.p.value, 32768, 16384, 3, d_out.p.value)
        assert e.value.code == -5
    finally:
        d_out.free()
