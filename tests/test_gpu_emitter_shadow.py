"""Shadow queries answered by the emitter's own triangle test.

An area light's shadow ray (light.h:176-188) runs from the shading point
toward a sampled point of the emitter, with t in units of |d| (the point at
t = 1) and tfar = |d| - 0.001: past the emitter once the light is farther
than 1.001, so the any-hit query (occlusion.cpp:88-107) is occluded by the
emitter itself if nothing comes first. The shading kernels answer such a
query with the traversal's own watertight test on that triangle (accepted
only well inside the ray's range) instead of queueing the ray. These tests
bake with and without SHB_BAKE_TRACE_ALL_SHADOW and require every output bit
and every query count to be the same, on the forms that reach the test:
one light sample, several, alpha splits, multi-triangle textured emitters,
the path-replay adjoint and the 262k-triangle room at its full geometry."""
import numpy as np
import pytest

import shb
from shb import scenes

pytestmark = pytest.mark.gpu


def bake_both(baker, spp, B, seed, nls=1):
    a_sh, a_vis, a_st = baker.bake(spp, B, seed=seed, num_light_samples=nls)
    b_sh, b_vis, b_st = baker.bake(spp, B, seed=seed, num_light_samples=nls,
                                   flags=shb.BAKE_TRACE_ALL_SHADOW)
    print(f"spp {spp} B {B} nls {nls}: shadow queries {a_st['rays_shadow']}, "
          f"answered by the emitter test {a_st['rays_shadow_resolved']}", flush=True)
    assert b_st["rays_shadow_resolved"] == 0
    assert a_st["rays_shadow"] == b_st["rays_shadow"]
    assert a_st["rays_extension"] == b_st["rays_extension"]
    assert np.array_equal(a_sh.view(np.uint32), b_sh.view(np.uint32))
    assert np.array_equal(a_vis.view(np.uint32), b_vis.view(np.uint32))
    return a_st


def with_alpha(s, every=4, alpha=150):
    """Every `every`-th 1x1 RGBA albedo material of s at the given alpha."""
    t = bytearray(s.texels)
    n = 0
    for k, m in enumerate(s.materials):
        if m["channels"] == 4 and m["width"] == 1 and m["emission"] == 0 and k % every == 0:
            t[m["offset"] + 3] = alpha
            n += 1
    s.texels = (np.frombuffer(bytes(t), np.uint8).copy() if isinstance(s.texels, np.ndarray)
                else type(s.texels)(t))
    assert n > 0
    return s


@pytest.mark.parametrize("nls", [1, 3])
def test_area_room_same_bits_with_and_without_the_emitter_test(baker, nls):
    s = scenes.scaled_synthetic_scene(20000, 192, 11)
    baker.upload_scene(s)
    baker.rasterize(192, 192, 1)
    st = bake_both(baker, 16, 3, 99, nls)
    # most NEE picks go to the emitters, and most of those are farther than 1
    assert st["rays_shadow_resolved"] > 0.2 * st["rays_shadow"]


def test_multi_triangle_textured_emitters_and_spots(baker):
    s = scenes.scaled_synthetic_scene(10000, 160, 77, n_spot=4, emitter_tris=3,
                                      textured_emission=True, emitter_y=(14.3, 14.8))
    baker.upload_scene(s)
    baker.rasterize(160, 160, 1)
    st = bake_both(baker, 16, 3, 4321)
    assert st["rays_shadow_resolved"] > 0


def test_alpha_splits(baker):
    s = with_alpha(scenes.scaled_synthetic_scene(10000, 128, 5))
    baker.upload_scene(s)
    baker.rasterize(128, 128, 1)
    st = bake_both(baker, 16, 3, 7)
    assert st["branches"] > 0 and st["rays_shadow_resolved"] > 0


def test_adjoint_same_gradient(baker):
    # the adjoint's gradient sums are float64 atomics (their order varies
    # from run to run), so the two gradients agree to rounding, not bitwise
    s = scenes.scaled_synthetic_scene(10000, 96, 3)
    baker.upload_scene(s)
    baker.rasterize(96, 96, 1)
    dsh = np.random.default_rng(1).standard_normal((96 * 96, 27)).astype(np.float32)
    la, ma, sa = baker.bake_adjoint(16, 3, dsh, seed=5)
    lb, mb, sb = baker.bake_adjoint(16, 3, dsh, seed=5, flags=shb.BAKE_TRACE_ALL_SHADOW)
    assert sa["rays_shadow_resolved"] > 0 and sb["rays_shadow_resolved"] == 0
    assert sa["rays_shadow"] == sb["rays_shadow"]
    for x, y in ((la, lb), (ma, mb)):
        assert np.abs(x).max() > 0
        np.testing.assert_allclose(x, y, rtol=1e-5, atol=1e-6 * np.abs(y).max())


def test_full_room_geometry(baker):
    # the 262k-triangle room with its 69 lights, at a low sample count
    s = scenes.synthetic_room_scene(512, 512)
    baker.upload_scene(s)
    baker.rasterize(512, 512, 2)
    st = bake_both(baker, 8, 3, 12345)
    assert st["rays_shadow_resolved"] > 0.2 * st["rays_shadow"]
