"""The HIP path against the committed golden fixtures (tests/golden/): the
box-room scene loaded from the fixture (not regenerated), its 4096 seeded
rays traced bit-exact against the oracle's FP32 hits, the 256^2 raster
checked by coverage, SHA-256 of every covered texel and 512 full records,
dilation and downsample bit-exact, and BakeSimpleQuad (baker_test.cpp:
10-65) against its closed form."""
import numpy as np
import pytest

import shb
from golden import fixtures as fx
from shb import SURFACE_DTYPE, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1(baker):
    z = fx.load("box_room_256")
    baker.upload_scene(fx.scene(z))
    return z


def test_trace_fixture_bit_exact(baker, c1):
    z = c1
    h = baker.trace_rays(z["rays"])
    exp = z["hits_fp32"]
    assert np.array_equal(h["prim_id"], exp["prim_id"])
    hit = exp["prim_id"] >= 0
    assert fx.same_bits(h[hit], exp[hit])
    a = baker.trace_rays(z["rays"], any_hit=True)
    assert np.array_equal((a["prim_id"] >= 0).astype(np.uint8), z["anyhit_fp32"])


def test_trace_fixture_every_bvh_width(baker, c1):
    z = c1
    saved = baker.tuning()
    try:
        for w in (2, 4, 8):
            baker.set_tuning(bvh_width=w)
            h = baker.trace_rays(z["rays"])
            hit = z["hits_fp32"]["prim_id"] >= 0
            assert np.array_equal(h["prim_id"], z["hits_fp32"]["prim_id"]), w
            assert fx.same_bits(h[hit], z["hits_fp32"][hit]), w
    finally:
        baker.set_tuning(**saved)


def test_raster_fixture(baker, c1):
    z = c1
    pts = baker.rasterize(256, 256, 1)
    assert pts.dtype == SURFACE_DTYPE
    assert np.array_equal(np.packbits(pts["material_id"] >= 0), z["raster_coverage"])
    assert np.array_equal(fx.surface_digest(pts), z["raster_sha256"])
    assert fx.same_bits(pts[z["raster_sample_index"]], z["raster_sample"])


def test_image_ops_fixture(baker):
    z = fx.load("image_ops")
    for key in ("dilate_27ch_2p", "dilate_1ch_5p"):
        w, h, passes = (int(x) for x in z[key + "_shape"])
        d, m = baker.dilate(z[key + "_in"], z[key + "_mask_in"], w, h, passes)
        assert fx.same_bits(d.reshape(z[key + "_out"].shape), z[key + "_out"]), key
        assert fx.same_bits(m, z[key + "_mask_out"]), key
    w, h, sc = (int(x) for x in z["downsample_shape"])
    d = baker.downsample(z["downsample_in"], w // sc, h // sc, sc)
    assert fx.same_bits(d.reshape(z["downsample_out"].shape), z["downsample_out"])
    m = baker.downsample_mask(z["downsample_mask_in"], w // sc, h // sc, sc)
    assert fx.same_bits(m, z["downsample_mask_out"])


def test_simple_quad_fixture(baker):
    z = fx.load("bake_simple_quad")
    baker.upload_scene(scenes.quad_scene())
    pts = np.zeros(16, SURFACE_DTYPE)
    pts["material_id"] = -1
    t = int(z["texel"][0])
    pts[t]["material_id"] = 0
    pts[t]["position"] = (-0.25, -0.25, 0)
    pts[t]["normal"] = (0, 0, 1)
    pts[t]["tangent"] = (1, 0, 0, 1)
    baker.set_surface(pts, 4, 4)
    sh, vis, _ = baker.bake(int(z["samples"][0]), 0)
    np.testing.assert_allclose(sh[t], z["sh"], rtol=3e-6, atol=1e-7)
    assert vis[t] == z["vis"][0]
    others = np.arange(16) != t
    assert not sh[others].any() and not vis[others].any()
