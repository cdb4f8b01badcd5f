"""The GPU's SH9 projection (k_project, sh_coeffs.cpp:50-78 with the
averaging of baker.cpp:284-294) against an FP64 closed form, with the error
stated. A plane lit by one point light and nothing else to hit: every sample
of a texel adds the same texel-NEE term L * Y_k(d), d = light - P
(unnormalised, from the un-offset texel position), L = I / |d|^2, and every
primary ray escapes. So each texel's SH is L * Y_k(d) exactly, evaluated
here in FP64 from the same float inputs, over a grid of directions."""
import numpy as np
import pytest

import shb
from shb import scenes

pytestmark = pytest.mark.gpu

K = np.array([0.282095, 0.488603, 0.488603, 0.488603, 1.092548, 1.092548, 0.315392, 1.092548,
              0.546274])


def basis64(d):
    x, y, z = d[:, 0], d[:, 1], d[:, 2]
    return np.stack([np.full_like(x, K[0]), K[1] * y, K[2] * z, K[3] * x, K[4] * x * y,
                     K[5] * y * z, K[6] * (3 * z * z - 1), K[7] * x * z, K[8] * (x * x - y * y)], 1)


def lit_plane(light_pos, intensity):
    s = scenes.quad_scene()
    s.positions = s.positions * np.float32(2.0)  # 4 x 4 m plane z = 0
    s.lights = [dict(type=shb.LIGHT_POINT, position=tuple(light_pos), intensity=intensity,
                     color=(1, 1, 1))]
    return s


@pytest.mark.parametrize("light", [(0.3, -0.2, 1.5), (-1.0, 0.7, 0.4)])
def test_projection_against_fp64(baker, light):
    s = lit_plane(light, 10.0)
    baker.upload_scene(s)
    pts = baker.rasterize(64, 64, 1)
    spp = 16
    sh, vis, st = baker.bake(spp, 0, seed=3)
    v = pts["material_id"] >= 0
    assert v.sum() > 3000
    P = pts["position"][v].astype(np.float64)
    d = np.asarray(light, np.float32).astype(np.float64) - P
    L = 10.0 / (d * d).sum(1)
    exp = (L[:, None] * basis64(d))  # (n, 9): same value in R, G, B
    got = sh[v].reshape(-1, 9, 3).astype(np.float64)
    scale = np.abs(exp).max(axis=1, keepdims=True)
    rel = np.abs(got - exp[:, :, None]) / scale[:, :, None]
    print(f"max relative error of the GPU projection vs FP64: {rel.max():.3e}")
    assert rel.max() < 2e-6
    np.testing.assert_array_equal(vis[v], 1.0)  # every primary ray escapes
