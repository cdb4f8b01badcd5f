"""GPU parity of spot and area-light NEE (light.h:117-144, 166-188;
light.cpp:16-78, 82-293) on the seeded synthetic room with every light
type: 4 point, 4 spot, 1 directional and 6 area lights, each emitter a fan
of 3 triangles of uneven areas (the uniform triangle pick, rng.index) with
a textured emission (emission at the sampled point's uv). Same stream as
the oracle's GPU mode, and per texel within the Monte Carlo error of the
oracle's reference stream. Parts of the ceiling lie within reach (< 1) of the
emitters (0.2-0.7 below it), so area-light shadow rays of texels and of
path vertices there arrive (farther ones end on their own emitter)."""
import numpy as np
import pytest

import shb
from shb import scenes
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

RES = 160


def mixed_light_room():
    return scenes.scaled_synthetic_scene(10000, RES, 77, n_spot=4, emitter_tris=3,
                                         textured_emission=True, emitter_y=(14.3, 14.8))


@pytest.fixture(scope="module")
def room():
    s = mixed_light_room()
    types = [l["type"] for l in s.lights]
    assert types.count(shb.LIGHT_SPOT) == 4 and types.count(shb.LIGHT_AREA) == 6
    for l in s.lights:
        if l["type"] == shb.LIGHT_AREA:
            g = s.geometries[l["geometry_index"]]
            assert g["index_count"] == 9  # three triangles
            assert s.materials[l["material_index"]]["width"] == 4  # textured emission
    return s, orc.Oracle(s)


def test_mixed_lights_same_stream_as_oracle(baker, room):
    s, o = room
    baker.upload_scene(s)
    pts = baker.rasterize(RES, RES, 1)
    spp, B = 16, 3
    g_sh, g_vis, g_st = baker.bake(spp, B, seed=4321)
    o_sh, o_vis, _, o_st = o.bake(pts, RES, RES, spp, B, mode=orc.MODE_GPU, seed=4321)
    v = pts["material_id"] >= 0
    assert v.sum() > 3000
    scale = np.abs(o_sh[v]).max(axis=1, keepdims=True) + 1e-6
    per_texel = (np.abs(g_sh[v] - o_sh[v]) / scale).max(axis=1)
    print(f"texels {v.sum()}: share within 1e-3 {np.mean(per_texel < 1e-3):.4f}", flush=True)
    assert (per_texel < 1e-3).mean() > 0.97, np.sort(per_texel)[-20:]
    assert np.abs(g_vis[v] - o_vis[v]).max() <= 2.0 / spp
    n_sh = o_st["rays_shadow"]
    assert abs(g_st["rays_shadow"] - n_sh) <= 2e-3 * n_sh


def test_mixed_lights_within_3_sigma_of_reference_stream(baker, room):
    s, o = room
    baker.upload_scene(s)
    pts = baker.rasterize(RES, RES, 1)
    spp, B = 64, 3
    g_sh, g_vis, _ = baker.bake(spp, B, seed=55)
    r_sh, r_vis, r_var, _ = o.bake(pts, RES, RES, spp, B, mode=orc.MODE_REF)
    _, _, g_var, _ = o.bake(pts, RES, RES, spp, B, mode=orc.MODE_GPU, seed=55)
    v = pts["material_id"] >= 0
    live = (r_var[v] + g_var[v]) > 0
    z = (g_sh[v] - r_sh[v])[live] / np.sqrt(r_var[v] + g_var[v])[live]
    print(f"mean |z| {np.abs(z).mean():.3f}, share |z| > 3: {(np.abs(z) > 3).mean():.4f}",
          flush=True)
    assert 0.6 < np.abs(z).mean() < 1.0, np.abs(z).mean()
    assert (np.abs(z) > 4.5).mean() < 2e-3
    np.testing.assert_allclose(g_vis[v].mean(), r_vis[v].mean(), atol=0.02)


@pytest.mark.parametrize("nls", [2, 5])
def test_several_light_samples_same_stream_as_oracle(baker, room, nls):
    """num_light_samples > 1: the GPU's n picks per NEE (inverse CDF over
    one set of per-light samples) against the oracle's GPU-stream
    restatement, and within the Monte Carlo error of the reference stream
    (std::discrete_distribution drawn n times, light.cpp:164-188)."""
    s, o = room
    baker.upload_scene(s)
    pts = baker.rasterize(RES, RES, 1)
    spp, B = 16, 3
    g_sh, g_vis, g_st = baker.bake(spp, B, seed=808, num_light_samples=nls)
    o_sh, o_vis, o_var, o_st = o.bake(pts, RES, RES, spp, B, mode=orc.MODE_GPU, seed=808,
                                      num_light_samples=nls)
    v = pts["material_id"] >= 0
    scale = np.abs(o_sh[v]).max(axis=1, keepdims=True) + 1e-6
    per_texel = (np.abs(g_sh[v] - o_sh[v]) / scale).max(axis=1)
    print(f"n={nls}: share within 1e-3 {np.mean(per_texel < 1e-3):.4f}", flush=True)
    assert (per_texel < 1e-3).mean() > 0.97, np.sort(per_texel)[-20:]
    n_sh = o_st["rays_shadow"]
    assert abs(g_st["rays_shadow"] - n_sh) <= 2e-3 * n_sh
    r_sh, _, r_var, _ = o.bake(pts, RES, RES, spp, B, mode=orc.MODE_REF, num_light_samples=nls)
    live = (r_var[v] + o_var[v]) > 0
    z = (g_sh[v] - r_sh[v])[live] / np.sqrt(r_var[v] + o_var[v])[live]
    print(f"  against the reference stream: mean |z| {np.abs(z).mean():.3f}")
    assert 0.6 < np.abs(z).mean() < 1.0
    assert (np.abs(z) > 4.5).mean() < 2e-3
