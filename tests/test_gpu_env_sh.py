"""GPU LatLong environment projection (shb_env_sh_latlong, replacing
ProjectLatLongMap, src/scene.cpp:31-71) against the oracle's FP64 sum of the
reference's own per-texel float terms, the uniform-white KAT of
scene_test.cpp:96-126 and a basis-function map. The GPU sums in FP64 in a
fixed order, so it agrees with the FP64 sum to float rounding of the terms
(the reference's serial float sum is further off; see tests/test_env_sh.py)."""
import math

import numpy as np
import pytest

import shb
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def check(g, s64, rtol=1e-5):
    assert np.abs(g.astype(np.float64) - s64).max() <= rtol * np.abs(s64).max(), (g, s64)


def test_uniform_white_kat(baker):
    g, _ = baker.env_sh_latlong(np.ones((32, 64, 3), np.float32))
    exp = 0.282095 * 4 * math.pi
    assert np.abs(g[:3] - exp).max() < 0.1 and np.abs(g[3:]).max() < 0.1
    check(g, orc.env_latlong(np.ones((32, 64, 3), np.float32))[1])


@pytest.mark.parametrize("h,w,ch", [(200, 400, 3), (37, 300, 4), (4096, 64, 3), (1, 5, 3),
                                    (513, 1031, 3)])
def test_random_maps_against_fp64(baker, h, w, ch):
    rng = np.random.default_rng(h * 7 + w)
    px = rng.uniform(0, 4, (h, w, ch)).astype(np.float32)
    px[..., :3] *= np.exp(rng.normal(0, 1, (h, w, 1))).astype(np.float32)  # HDR spread
    g, sec = baker.env_sh_latlong(px)
    check(g, orc.env_latlong(px)[1])
    assert sec > 0
    # deterministic, and the resident map is reused with pixels=None
    g2, _ = baker.env_sh_latlong(None, shape=(h, w, ch))
    assert np.array_equal(g, g2)


def test_basis_map_and_alpha_ignored(baker):
    w, h = 512, 256
    th = (np.arange(h) + 0.5) * math.pi / h
    ph = (np.arange(w) + 0.5) * 2 * math.pi / w
    T, P = np.meshgrid(th, ph, indexing="ij")
    x, y, z = np.sin(T) * np.sin(P), np.cos(T), np.sin(T) * np.cos(P)
    px = np.zeros((h, w, 4), np.float32)
    px[..., 0] = 1.092548 * x * z          # Y_7
    px[..., 1] = 0.546274 * (x * x - y * y)  # Y_8
    px[..., 3] = -1e6
    g, _ = baker.env_sh_latlong(px)
    c = g.reshape(9, 3).astype(np.float64)
    e7, e8 = np.eye(9)[7], np.eye(9)[8]
    np.testing.assert_allclose(c[:, 0], e7, atol=2e-4)
    np.testing.assert_allclose(c[:, 1], e8, atol=2e-4)
    assert not c[:, 2].any()


def test_large_map_against_fp64(baker):
    rng = np.random.default_rng(5)
    px = rng.uniform(0, 10, (2048, 4096, 3)).astype(np.float32)
    g, _ = baker.env_sh_latlong(px)
    ref, s64 = orc.env_latlong(px)
    check(g, s64)
    # the GPU's FP64 sum is closer to the exact sum than the serial float one
    assert np.abs(g - s64).max() <= np.abs(ref - s64).max()


def test_errors(baker):
    with pytest.raises(shb.ShbError):
        baker.env_sh_latlong(np.ones((4, 8, 2), np.float32))
    with pytest.raises(shb.ShbError):
        baker.env_sh_latlong(None, shape=(3, 3, 3))
