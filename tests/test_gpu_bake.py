"""GPU bake parity (BakeSHLightMap, src/baker.cpp:186-304):
  * same stream: the GPU and the oracle in its GPU-stream mode consume the
    same counter-based random numbers, so every texel matches up to FP32
    rounding of transcendentals (rare path flips at triangle edges allowed);
  * statistics: against the oracle's reference stream (std::mt19937 +
    libstdc++ distributions) per texel within the Monte Carlo error;
  * the closed form of BakerTest.BakeSimpleQuad (baker_test.cpp:10-65);
  * the post-bake pipeline (main.cpp:138-163) bit exact against the
    oracle's downsample / mask / dilate applied to the GPU's raw bake."""
import numpy as np
import pytest

import shb
from shb import SURFACE_DTYPE, scenes
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

Y_PLUS_Z = np.array([0.282095, 0, 0.488603, 0, 0, 0, 0.630784, 0, 0], np.float32)


@pytest.fixture(scope="module")
def room():
    s = scenes.box_room_scene(48, 48)
    return s, orc.Oracle(s)


def test_bake_simple_quad_closed_form(baker):
    s = scenes.quad_scene()
    baker.upload_scene(s)
    pts = np.zeros(16, SURFACE_DTYPE)
    pts["material_id"] = -1
    pts[5]["material_id"] = 0
    pts[5]["position"] = (-0.25, -0.25, 0)
    pts[5]["normal"] = (0, 0, 1)
    pts[5]["tangent"] = (1, 0, 0, 1)
    baker.set_surface(pts, 4, 4, 1)
    sh, vis, st = baker.bake(64, 0)
    exp = np.repeat(Y_PLUS_Z, 3)
    np.testing.assert_allclose(sh[5], exp, rtol=1e-6, atol=1e-7)
    assert sh[5][0] > 0.1  # baker_test.cpp:64
    assert vis[5] == 1.0
    others = np.delete(np.arange(16), 5)
    assert (sh[others] == 0).all() and (vis[others] == 0).all()
    assert st["valid_texels"] == 1 and st["rays_extension"] == 64 and st["rays_shadow"] == 64


def test_bake_same_stream_as_oracle(baker, room):
    s, o = room
    baker.upload_scene(s)
    pts = baker.rasterize(48, 48, 1)
    spp, B = 16, 3
    g_sh, g_vis, g_st = baker.bake(spp, B, seed=12345)
    o_sh, o_vis, _, o_st = o.bake(pts, 48, 48, spp, B, mode=orc.MODE_GPU, seed=12345)
    v = pts["material_id"] >= 0
    scale = np.abs(o_sh[v]).max(axis=1, keepdims=True) + 1e-6
    rel = np.abs(g_sh[v] - o_sh[v]) / scale
    per_texel = rel.max(axis=1)
    assert (per_texel < 1e-3).mean() > 0.97, np.sort(per_texel)[-20:]
    assert np.abs(g_vis[v] - o_vis[v]).max() <= 2.0 / spp
    n_ext, n_sh = o_st["rays_extension"], o_st["rays_shadow"]
    assert abs(g_st["rays_extension"] - n_ext) <= 1e-3 * n_ext
    assert abs(g_st["rays_shadow"] - n_sh) <= 2e-3 * n_sh


def test_bake_within_3_sigma_of_reference_stream(baker, room):
    s, o = room
    baker.upload_scene(s)
    pts = baker.rasterize(48, 48, 1)
    spp, B = 64, 3
    g_sh, g_vis, _ = baker.bake(spp, B, seed=777)
    r_sh, r_vis, r_var, _ = o.bake(pts, 48, 48, spp, B, mode=orc.MODE_REF)
    # the GPU's per-texel sample variance: the oracle's GPU-stream mode draws
    # the same samples, so its variance estimate is the GPU's
    _, _, g_var, _ = o.bake(pts, 48, 48, spp, B, mode=orc.MODE_GPU, seed=777)
    v = pts["material_id"] >= 0
    live = (r_var[v] + g_var[v]) > 0
    z = (g_sh[v] - r_sh[v])[live] / np.sqrt(r_var[v] + g_var[v])[live]
    assert 0.6 < np.abs(z).mean() < 1.0, np.abs(z).mean()
    assert (np.abs(z) > 4.5).mean() < 1e-3
    np.testing.assert_allclose(g_vis[v].mean(), r_vis[v].mean(), atol=0.02)


def test_bake_deterministic_and_seeded(baker, room):
    s, _ = room
    baker.upload_scene(s)
    baker.rasterize(48, 48, 1, want=False)
    a = baker.bake(8, 2, seed=5)[0]
    b = baker.bake(8, 2, seed=5)[0]
    c = baker.bake(8, 2, seed=6)[0]
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert not np.array_equal(a, c)


def test_bake_batch_size_invariant(baker, room):
    s, _ = room
    baker.upload_scene(s)
    baker.rasterize(48, 48, 1, want=False)
    a = baker.bake(16, 3, seed=9)[0]
    old = baker.tuning()
    baker.set_tuning(batch_paths=4096)
    try:
        b = baker.bake(16, 3, seed=9)[0]
    finally:
        baker.set_tuning(**old)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("chunk,block", [(16, 64), (12, 256), (1, 128), (64, 256)])
def test_bake_persistent_traversal_bit_identical(baker, room, chunk, block):
    """The persistent traversal (refill threshold `chunk`, including the
    default 12 / 256) gives the one-ray-per-thread kernel's bits."""
    s, _ = room
    baker.upload_scene(s)
    baker.rasterize(48, 48, 1, want=False)
    old = baker.tuning()
    try:
        baker.set_tuning(queue_chunk=0, block_size=256)
        a, av, _ = baker.bake(16, 3, seed=21)
        baker.set_tuning(queue_chunk=chunk, block_size=block)
        b, bv, _ = baker.bake(16, 3, seed=21)
    finally:
        baker.set_tuning(**old)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.array_equal(av.view(np.uint32), bv.view(np.uint32))


@pytest.mark.parametrize("ss,dil", [(2, 2), (1, 3), (3, 0)])
def test_lightmap_pipeline_bit_exact(baker, room, ss, dil):
    s, _ = room
    baker.upload_scene(s)
    pts = baker.rasterize(32, 32, ss)
    raw_sh, raw_vis, _ = baker.bake(8, 2, seed=3)
    sh, vis, mask, _ = baker.bake_lightmap(8, 2, dil, seed=3)
    W = 32 * ss
    m = orc.validity_mask(pts)
    if ss > 1:
        e_sh = orc.downsample(raw_sh, W, W, ss)
        e_vis = orc.downsample(raw_vis, W, W, ss)[:, 0]
        m = orc.downsample_mask(m, 32, 32, ss)
    else:
        e_sh, e_vis = raw_sh, raw_vis
    if dil > 0:
        e_sh, m = orc.dilate(e_sh, m, 32, 32, dil)
        e_vis, m = orc.dilate(e_vis, m, 32, 32, dil)
    np.testing.assert_array_equal(mask, m)
    assert np.array_equal(sh.view(np.uint32), e_sh.reshape(sh.shape).view(np.uint32))
    assert np.array_equal(vis.view(np.uint32), e_vis.reshape(vis.shape).view(np.uint32))


def test_lightmap_device_resident_read_back(baker, room):
    s, _ = room
    baker.upload_scene(s)
    baker.rasterize(32, 32, 2, want=False)
    sh, vis, mask, _ = baker.bake_lightmap(8, 2, 1, seed=4)
    none = baker.bake_lightmap(8, 2, 1, seed=4, to_host=False)
    assert none[0] is None
    sh2, vis2, mask2 = baker.read_lightmap()
    assert np.array_equal(sh.view(np.uint32), sh2.view(np.uint32))
    assert np.array_equal(vis.view(np.uint32), vis2.view(np.uint32))
    np.testing.assert_array_equal(mask, mask2)
    baker.rasterize(32, 32, 2, want=False)  # invalidates the device result
    with pytest.raises(shb.ShbError):
        baker.read_lightmap()


def test_bake_errors(baker):
    s = scenes.box_room_scene(64, 64)
    fresh = shb.Baker(0)
    with pytest.raises(shb.ShbError) as e:
        fresh.bake(4, 1)
    assert e.value.code in (-2, -6)
    fresh.upload_scene(s)
    with pytest.raises(shb.ShbError) as e:
        fresh.bake(4, 1)
    assert e.value.code == -6
    fresh.rasterize(64, 64, 1, want=False)
    with pytest.raises(shb.ShbError) as e:
        fresh.bake(4, 1, num_light_samples=17)
    assert e.value.code == -1
    with pytest.raises(shb.ShbError) as e:
        fresh.bake(4, 1, num_light_samples=0)
    assert e.value.code == -1
    with pytest.raises(shb.ShbError) as e:  # the adjoint takes one light sample
        fresh.bake_adjoint(4, 1, np.zeros((64 * 64, 27), np.float32), num_light_samples=2)
    assert e.value.code == -7
    with pytest.raises(shb.ShbError) as e:
        fresh.bake(0, 1)
    assert e.value.code == -1
    with pytest.raises(shb.ShbError) as e:
        fresh.bake(4, 16)
    assert e.value.code == -1
    fresh.close()


def test_transparent_material_bakes(baker):
    # alpha < 1 is baked (the path split of tests/test_gpu_alpha.py)
    s = scenes.quad_scene()
    s.texels = np.array([255, 255, 255, 128], np.uint8)
    baker.upload_scene(s)
    baker.rasterize(4, 4, 1, want=False)
    sh, vis, st = baker.bake(4, 1)
    assert np.isfinite(sh).all() and st["valid_texels"] > 0
