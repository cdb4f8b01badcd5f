"""The reference's whole pipeline on its own glTF fixtures (main.cpp:51-163):
LoadScene -> CreateAtlasGeometries -> RasterizeScene -> BakeSHLightMap ->
downsample / dilate -> SaveSHLightMap, every step through this library
(shb_load_gltf, shb_atlas_create, the GPU raster and bake, shb_save_lightmap),
checked against the oracle on the same loaded and atlased scene: the raster
bit exact, the bake on the oracle's GPU-stream mode (same samples) and
within Monte Carlo error of its reference stream."""
import os

import numpy as np
import pytest

import shb
from oracle import oracle as orc
from shb import exr

pytestmark = pytest.mark.gpu

GL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "gltf")


def prepared(name, res, pad):
    s, env = shb.load_gltf(os.path.join(GL, name))
    a, info = shb.create_atlas(s, res, pad)
    return a, env


@pytest.mark.parametrize("name,res", [("DirectionalLight/scene.gltf", 96),
                                      ("LightsPunctualLamp/LightsPunctualLamp.gltf", 128),
                                      ("box/scene.gltf", 48)])
def test_loaded_scene_raster_and_bake_match_oracle(baker, name, res):
    s, _ = prepared(name, res, 2)
    baker.upload_scene(s)
    pts = baker.rasterize(res, res, 1)
    o = orc.Oracle(s)
    opts = o.rasterize(res, res, 1)
    np.testing.assert_array_equal(pts["material_id"], opts["material_id"])
    v = pts["material_id"] >= 0
    assert v.mean() > 0.05
    for k in ("position", "normal", "tangent"):
        assert np.array_equal(pts[k][v].view(np.uint32), opts[k][v].view(np.uint32)), k
    spp, B = 16, 3
    g_sh, g_vis, g_st = baker.bake(spp, B, seed=4321)
    o_sh, o_vis, _, o_st = o.bake(pts, res, res, spp, B, mode=orc.MODE_GPU, seed=4321)
    scale = np.abs(o_sh[v]).max(axis=1, keepdims=True) + 1e-6
    per_texel = (np.abs(g_sh[v] - o_sh[v]) / scale).max(axis=1)
    assert (per_texel < 1e-3).mean() > 0.97, np.sort(per_texel)[-20:]
    assert np.abs(g_vis[v] - o_vis[v]).max() <= 2.0 / spp
    assert np.isfinite(g_sh).all() and np.abs(g_sh[v]).max() > 0


@pytest.mark.parametrize("name", ["LightsPunctualLamp/LightsPunctualLamp.gltf"])
def test_loaded_scene_within_3_sigma_of_reference_stream(baker, name):
    # the lamp has five point lights and a textured shade with alpha < 1
    s, _ = prepared(name, 96, 2)
    baker.upload_scene(s)
    pts = baker.rasterize(96, 96, 1)
    o = orc.Oracle(s)
    spp, B = 64, 3
    g_sh, g_vis, _ = baker.bake(spp, B, seed=99)
    r_sh, r_vis, r_var, _ = o.bake(pts, 96, 96, spp, B, mode=orc.MODE_REF)
    _, _, g_var, _ = o.bake(pts, 96, 96, spp, B, mode=orc.MODE_GPU, seed=99)
    v = pts["material_id"] >= 0
    live = (r_var[v] + g_var[v]) > 0
    assert live.sum() > 100
    z = (g_sh[v] - r_sh[v])[live] / np.sqrt(r_var[v] + g_var[v])[live]
    assert 0.6 < np.abs(z).mean() < 1.0, np.abs(z).mean()
    assert (np.abs(z) > 4.5).mean() < 1e-3


def test_full_pipeline_to_exr(baker, tmp_path):
    s, env = prepared("DirectionalLight/scene.gltf", 128, 2)
    baker.upload_scene(s)
    baker.rasterize(128, 128, 2, want=False)
    sh, vis, mask, st = baker.bake_lightmap(16, 2, 2, seed=7)
    assert sh.shape == (128 * 128, 27)
    shb.save_lightmap(str(tmp_path / "lightmap.exr"), sh, vis, 128, 128, shb.SAVE_SPLIT)
    _, ch = exr.read_exr(str(tmp_path / "lightmap_L0.exr"))
    half = lambda a: np.asarray(a, np.float32).astype(np.float16).astype(np.float32)
    np.testing.assert_array_equal(ch["R"].ravel(), half(sh[:, 0]))
    np.testing.assert_array_equal(ch["A"].ravel(), half(vis))
    # the box's sun gives the Preetham sky whose SH the viewer pairs with vis
    assert env["type"] == shb.ENV_PREETHAM
    sky = shb.env_sh_preetham(env["sun_direction"])
    assert sky[2] > 0 and np.isfinite(sky).all()
