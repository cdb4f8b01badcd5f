"""The oracle against the committed golden fixtures (tests/golden/), and the
fixtures against independent checks: the FP32 watertight hits against the
FP64 brute force, the SH table against the reference's basis constants
(sh_coeffs.cpp:27-31) and the BakeSimpleQuad closed form (baker_test.cpp:
10-65). CPU only; tests/test_gpu_golden.py runs the HIP path against the
same fixtures."""
import numpy as np
import pytest

from golden import fixtures as fx
from oracle import oracle as orc
from shb import SURFACE_DTYPE, scenes


@pytest.fixture(scope="module")
def c1():
    z = fx.load("box_room_256")
    return z, fx.scene(z)


def test_fixture_scene_is_the_box_room(c1):
    z, s = c1
    ref = scenes.box_room_scene(256, 256)
    for k in ("positions", "normals", "texture_uvs", "lightmap_uvs", "tangents", "indices",
              "texels"):
        assert fx.same_bits(getattr(s, k), getattr(ref, k)), k
    assert s.num_triangles == 24 and len(s.lights) == 1


def test_oracle_reproduces_trace_fixture(c1):
    z, s = c1
    o = orc.Oracle(s)
    assert fx.same_bits(o.trace(z["rays"], mode=0), z["hits_fp32"])
    assert np.array_equal((o.trace(z["rays"], mode=2)["prim_id"] >= 0).astype(np.uint8),
                          z["anyhit_fp32"])
    h64 = o.trace(z["rays"], mode=1)
    assert np.array_equal(h64["prim_id"], z["hits_fp64_prim"])
    assert fx.same_bits(h64["t"], z["hits_fp64_t"])


def test_fp32_hits_against_fp64_bruteforce(c1):
    z, _ = c1
    h, p64, t64 = z["hits_fp32"], z["hits_fp64_prim"], z["hits_fp64_t"]
    assert (p64 >= 0).mean() > 0.3
    same = h["prim_id"] == p64
    assert same.mean() >= 0.995
    both = same & (p64 >= 0)
    dt = np.abs(h["t"][both].astype(np.float64) - t64[both])
    assert np.quantile(dt / np.abs(t64[both]), 0.999) < 1e-5
    # every hit point within 1e-5 of the coordinate scale (origins a few
    # 1e-4 above a wall give t ~ 1e-4, where relative t is ill-conditioned)
    r = z["rays"][both]
    scale = np.maximum(1.0, np.abs(r["origin"]).max(axis=1))
    assert (dt * np.linalg.norm(r["direction"], axis=1) / scale).max() < 1e-5
    # any hit within [tnear, tfar] exactly when a closest hit exists
    assert np.array_equal(z["anyhit_fp32"] == 1, h["prim_id"] >= 0)


def test_oracle_reproduces_raster_fixture(c1):
    z, s = c1
    pts = orc.Oracle(s).rasterize(256, 256, 1)
    assert pts.dtype == SURFACE_DTYPE
    assert np.array_equal(np.packbits(pts["material_id"] >= 0), z["raster_coverage"])
    assert np.array_equal(fx.surface_digest(pts), z["raster_sha256"])
    assert fx.same_bits(pts[z["raster_sample_index"]], z["raster_sample"])


def test_oracle_reproduces_image_fixtures():
    z = fx.load("image_ops")
    for key in ("dilate_27ch_2p", "dilate_1ch_5p"):
        w, h, passes = (int(x) for x in z[key + "_shape"])
        d, m = orc.dilate(z[key + "_in"], z[key + "_mask_in"], w, h, passes)
        assert fx.same_bits(d, z[key + "_out"]) and fx.same_bits(m, z[key + "_mask_out"]), key
    w, h, sc = (int(x) for x in z["downsample_shape"])
    assert fx.same_bits(orc.downsample(z["downsample_in"], w, h, sc), z["downsample_out"])
    assert fx.same_bits(orc.downsample_mask(z["downsample_mask_in"], w // sc, h // sc, sc),
                        z["downsample_mask_out"])


def test_sh_table_and_oracle_basis():
    z = fx.load("sh9_fibonacci")
    d, Y = z["directions"], z["sh"]
    # the table itself: orthonormality on the sphere (quadrature over the
    # equal-area Fibonacci points) and Y_00 = 1 / (2 sqrt(pi))
    G = Y.T @ Y * (4 * np.pi / len(d))
    assert np.abs(G - np.eye(9)).max() < 5e-3
    assert np.allclose(Y[:, 0], 0.5 / np.sqrt(np.pi))
    worst = 0.0
    for i in range(0, len(d), 4):
        sh = orc.sh_accumulate([1, 0, 0], d[i]).reshape(9, 3)[:, 0]
        worst = max(worst, np.abs(sh - Y[i]).max())
    assert worst < 2e-6, worst  # the reference's 6-digit constants


def test_simple_quad_fixture_and_oracle():
    z = fx.load("bake_simple_quad")
    exp = z["sh"]
    assert np.allclose(exp.reshape(9, 3)[:, 0], [0.282095, 0, 0.488603, 0, 0, 0, 0.630784, 0, 0],
                       atol=1e-6)  # baker_test.cpp:58-64 expectations
    s = scenes.quad_scene()
    pts = np.zeros(16, SURFACE_DTYPE)
    pts["material_id"] = -1
    t = int(z["texel"][0])
    pts[t]["material_id"] = 0
    pts[t]["position"] = (-0.25, -0.25, 0)
    pts[t]["normal"] = (0, 0, 1)
    pts[t]["tangent"] = (1, 0, 0, 1)
    sh, vis, _, _ = orc.Oracle(s).bake(pts, 4, 4, int(z["samples"][0]), 0)
    np.testing.assert_allclose(sh[t], exp, rtol=3e-6, atol=1e-7)
    assert vis[t] == z["vis"][0]
