"""Environment SH (src/scene.cpp:22-160, ProjectEnvironmentToSH): the oracle's
LatLong projection and the library's host Preetham projection, pinned by
closed forms. The GPU LatLong kernel is checked in tests/test_gpu_env_sh.py.

LatLong pins: the uniform-white KAT of scene_test.cpp:96-126; a map equal to
one real SH basis function projects to that coefficient alone
(orthonormality; midpoint quadrature and the 6-digit basis constants bound
the error); the reference's serial float sum against the FP64 sum of the
same terms. Preetham pins: every coefficient is (0.2, 0.5, 0.9) times one
scalar; a zenith sun is symmetric about +Y, so only Y_0, y, 3z^2-1 and
x^2-y^2 survive, with c6 / 0.315392 = c8 / 0.546274; a sun turned 90
degrees about +Y turns the coefficients by the matching SH rotation (the
phi grid maps onto itself); a 16x finer FP64 quadrature agrees to the
midpoint rule's error."""
import math

import numpy as np
import pytest

import shb
from oracle import oracle as orc

K = [0.282095, 0.488603, 0.488603, 0.488603, 1.092548, 1.092548, 0.315392, 1.092548, 0.546274]


def basis(d):
    x, y, z = d[..., 0], d[..., 1], d[..., 2]
    return np.stack([np.full_like(x, K[0]), K[1] * y, K[2] * z, K[3] * x, K[4] * x * y,
                     K[5] * y * z, K[6] * (3 * z * z - 1), K[7] * x * z, K[8] * (x * x - y * y)],
                    -1)


def latlong_dirs(w, h):
    th = (np.arange(h) + 0.5) * math.pi / h
    ph = (np.arange(w) + 0.5) * 2 * math.pi / w
    T, P = np.meshgrid(th, ph, indexing="ij")
    return np.stack([np.sin(T) * np.sin(P), np.cos(T), np.sin(T) * np.cos(P)], -1)


def test_latlong_uniform_white_kat():
    # scene_test.cpp:96-126: 64 x 32 white map -> L0 ~ 4 pi Y00, others ~ 0
    ref, s64 = orc.env_latlong(np.ones((32, 64, 3), np.float32))
    exp = 0.282095 * 4 * math.pi
    for c in range(3):
        assert abs(ref[c] - exp) < 0.1 and abs(s64[c] - exp) < 0.1
    assert np.abs(ref[3:]).max() < 0.1 and np.abs(s64[3:]).max() < 0.1
    # the discrete solid angles sum to 4 pi to O(1/h^2)
    assert abs(s64[0] / 0.282095 - 4 * math.pi) < 4 * math.pi * 2e-3


@pytest.mark.parametrize("k", range(9))
def test_latlong_basis_map_projects_to_one_coefficient(k):
    w, h = 256, 128
    Y = basis(latlong_dirs(w, h))[..., k]
    px = np.zeros((h, w, 4), np.float32)
    px[..., 0] = Y          # R carries Y_k, G carries 2 Y_k, B nothing,
    px[..., 1] = 2 * Y      # A is ignored
    px[..., 3] = 1e6
    _, s64 = orc.env_latlong(px)
    c = s64.reshape(9, 3)
    e = np.zeros(9)
    e[k] = 1.0
    np.testing.assert_allclose(c[:, 0], e, atol=3e-4)
    np.testing.assert_allclose(c[:, 1], 2 * e, atol=6e-4)
    assert not c[:, 2].any()


def test_latlong_serial_float_sum_against_fp64():
    rng = np.random.default_rng(1)
    px = rng.uniform(0, 5, (200, 400, 3)).astype(np.float32)
    ref, s64 = orc.env_latlong(px)
    # a serial float sum of n terms errs by up to ~n eps |sum|
    n = 200 * 400
    assert np.abs(ref - s64).max() <= n * 6e-8 * np.abs(s64).max()
    assert np.abs(ref - s64).max() > 0  # and it does differ


def test_preetham_matches_oracle_and_color_ratio():
    for sun in [(0, 1, 0), (0.3, 0.8, 0.2), (1, -0.2, 0), (0, 0, 0), (2, 5, -1)]:
        a = shb.env_sh_preetham(sun)
        assert np.array_equal(a, orc.env_preetham(sun)), sun
        c = a.reshape(9, 3).astype(np.float64) / [0.2, 0.5, 0.9]
        # float sums of 4096 terms of both signs: up to a few 1e-5 of the
        # largest coefficient (n eps = 2.4e-4 bounds it)
        assert np.abs(c - c[:, 2:3]).max() < 1e-4 * np.abs(c).max(), sun


def test_preetham_zenith_sun_symmetry():
    c = shb.env_sh_preetham((0, 1, 0)).reshape(9, 3)[:, 0].astype(np.float64)
    assert np.abs(c[[2, 3, 4, 5, 7]]).max() < 1e-5 * c[0]
    assert c[0] > 0 and c[1] > 0  # brighter above
    assert abs(c[6] / K[6] - c[8] / K[8]) < 1e-5 * c[0]


def test_preetham_rotation_about_y():
    # turning the sun by +90 degrees about Y maps (x, z) -> (z, -x); the
    # projection rotates the same way: y, z, x -> y, -x, z for band 1 and
    # xy, yz, 3z^2-1, xz, x^2-y^2 -> yz, -xy, 3x^2-1, -xz, z^2-y^2 in band 2
    s = np.array([0.6, 0.5, 0.2])
    a = shb.env_sh_preetham(s).reshape(9, 3)[:, 0].astype(np.float64)
    b = shb.env_sh_preetham((s[2], s[1], -s[0])).reshape(9, 3)[:, 0].astype(np.float64)
    tol = 2e-5 * abs(a[0])
    # dir d' = R d with R: (x, y, z) -> (z, y, -x); f'(d') = f(d), so
    # c'_k = sum f(d) Y_k(R d)
    assert abs(b[0] - a[0]) < tol and abs(b[1] - a[1]) < tol
    assert abs(b[2] + a[3]) < tol and abs(b[3] - a[2]) < tol   # z' = -x, x' = z
    assert abs(b[4] / K[4] - a[5] / K[5]) < tol                 # x'y' = zy
    assert abs(b[5] / K[5] + a[4] / K[4]) < tol                 # y'z' = -xy
    assert abs(b[7] / K[7] + a[7] / K[7]) < tol                 # x'z' = -zx
    # 3z'^2 - 1 = 3x^2 - 1 and x'^2 - y'^2 = z^2 - y^2 follow from a's
    # quadratic moments: with q = (xx, yy, zz) sums, c6 = K6 (3 zz - S),
    # c8 = K8 (xx - yy), and S = xx + yy + zz = c0 / K0
    S = a[0] / K[0]
    zz = (a[6] / K[6] + S) / 3
    xx_minus_yy = a[8] / K[8]
    # xx + yy = S - zz
    xx = (S - zz + xx_minus_yy) / 2
    yy = S - zz - xx
    assert abs(b[6] / K[6] - (3 * xx - S)) < 3 * tol / K[6]
    assert abs(b[8] / K[8] - (zz - yy)) < 3 * tol / K[8]


def perez64(d, s):
    cz = np.maximum(0.001, d[..., 1])
    cs = np.clip(d @ s, -1, 1)
    g = np.arccos(cs)
    return (1 - np.exp(-0.32 / cz)) * (1 + 10 * np.exp(-3 * g) + 0.45 * cs * cs)


def test_preetham_against_fine_fp64_quadrature():
    s = np.array([0.4, 0.7, -0.3])
    s = s / np.linalg.norm(s)
    zen = perez64(np.array([0.0, 1.0, 0.0]), s)
    nt, nph = 64 * 16, 128 * 16
    th = (np.arange(nt) + 0.5) * math.pi / nt
    ph = (np.arange(nph) + 0.5) * 2 * math.pi / nph
    T, P = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.sin(P), np.cos(T), np.sin(T) * np.cos(P)], -1)
    f = np.minimum(perez64(d, s), 30 * zen) / zen * (d[..., 1] >= 0)
    w = np.sin(T) * (math.pi / nt) * (2 * math.pi / nph)
    fine = np.einsum("ij,ijk->k", f * w, basis(d))
    got = shb.env_sh_preetham(s).reshape(9, 3)[:, 2].astype(np.float64) / 0.9
    # midpoint rule on a clamped, horizon-cut integrand: a few 1e-3 of L0
    assert np.abs(got - fine).max() < 1e-2 * fine[0], (got, fine)
