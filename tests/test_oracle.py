"""Pins the CPU oracle (oracle/oracle.cpp) before it is trusted as the
checker: the reference's own known-answer tests (src/*_test.cpp), published
known answers (Philox4x32-10, xoshiro128**), a library routine (scipy's
spherical harmonics), independent numpy restatements (raster, masks,
downsample, dilation), FP64 brute force (traversal), closed forms (lights,
BakeSimpleQuad) and the statistical equivalence of the oracle's two random
streams. Runs on the CPU."""
import math
import os

import numpy as np
import pytest

from shb import RAY_DTYPE, SURFACE_DTYPE, scenes
from oracle import oracle as orc


# ------------------------------------------------------------------ RNG
def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32 with 10 rounds
    assert list(orc.philox([0, 0, 0, 0], [0, 0])) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert list(orc.philox([0xffffffff] * 4, [0xffffffff] * 2)) == \
        [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert list(orc.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                           [0xa4093822, 0x299f31d0])) == [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def _xoshiro_py(s, n):
    s = list(s)
    out = []
    M = 0xffffffff
    rotl = lambda x, k: ((x << k) | (x >> (32 - k))) & M
    for _ in range(n):
        out.append((rotl((s[1] * 5) & M, 7) * 9) & M)
        t = (s[1] << 9) & M
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t
        s[3] = rotl(s[3], 11)
    return out


def test_xoshiro_known_answers():
    got = list(orc.xoshiro([1, 2, 3, 4], 64))
    assert got[:3] == [11520, 0, 5927040]  # worked by hand from the definition
    assert got == _xoshiro_py([1, 2, 3, 4], 64)


def test_site_uniforms_range_and_independence():
    u = np.concatenate([orc.site_uniforms(12345, t, s, 0, 1, 0, 0, 4) for t in range(200) for s in range(8)])
    assert u.min() >= 0 and u.max() < 1
    assert abs(u.mean() - 0.5) < 0.01 and abs(u.var() - 1 / 12) < 0.005
    a = orc.site_uniforms(1, 7, 3, 2, 1, 1, 0, 8)
    b = orc.site_uniforms(1, 7, 3, 2, 2, 1, 0, 8)
    assert not np.array_equal(a, b)


# --------------------------------------------------------------- SH9
def test_sh_basis_kat():
    sh = orc.sh_accumulate([1, 1, 1], [0, 0, 1])  # sh_coeffs_test.cpp:7-38
    exp = [0.282095, 0, 0.488603, 0, 0, 0, 0.630784, 0, 0]
    np.testing.assert_allclose(sh.reshape(9, 3)[:, 0], exp, atol=1e-5)


def test_sh_basis_against_scipy_real_sh():
    from scipy.special import sph_harm_y
    rng = np.random.default_rng(0)
    worst = 0.0
    for _ in range(500):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        sh = orc.sh_accumulate([1, 0, 0], d).reshape(9, 3)[:, 0].astype(np.float64)
        theta, phi = math.acos(d[2]), math.atan2(d[1], d[0])
        ref = []
        for l, m in [(0, 0), (1, -1), (1, 0), (1, 1), (2, -2), (2, -1), (2, 0), (2, 1), (2, 2)]:
            y = sph_harm_y(l, abs(m), theta, phi)
            if m > 0:
                r = math.sqrt(2) * (-1) ** m * y.real
            elif m < 0:
                r = math.sqrt(2) * (-1) ** m * y.imag
            else:
                r = y.real
            ref.append(r)
        ref = np.array(ref)  # standard real SH from scipy's complex Y_l^m
        worst = max(worst, np.abs(sh - ref).max())
    assert worst < 2e-6, worst  # 6-digit constants of sh_coeffs.cpp:27-31


def test_sh_accumulates_unnormalised_direction():
    # AccumulateRadiance does not normalise (sh_coeffs.cpp:50-78)
    sh = orc.sh_accumulate([1, 0, 0], [0, 0, 2]).reshape(9, 3)[:, 0]
    assert abs(sh[2] - 2 * 0.488603) < 1e-5 and abs(sh[6] - 0.315392 * 11) < 1e-4


# ------------------------------------------------------------ occlusion
def test_occlusion_kats():
    s = scenes.triangles_scene([[[-1, -1, 0], [1, -1, 0], [0, 1, 0]]])
    o = orc.Oracle(s)
    r = np.zeros(6, RAY_DTYPE)
    r["origin"] = [[0, 0, 1], [2, 0, 1], [0, 0, 1], [0, 0, 1], [0, 0, 1], [0, 0, 1]]
    r["direction"] = [[0, 0, -1], [0, 0, -1], [0, 0, 1], [0, 0, -1], [0, 0, -1], [0, 0, -1]]
    r["tnear"] = [0, 0, 0, 0, 1.2, 0.1]
    r["tfar"] = [np.inf, np.inf, np.inf, 0.5, np.inf, 10.0]
    h = o.trace(r)
    assert h["prim_id"][0] == 0 and abs(h["position"][0][2]) < 1e-4
    assert abs(h["normal"][0][2] - 1) < 1e-4  # geometric normal (v1-v0)x(v2-v0)
    assert list(h["prim_id"][1:5] < 0) == [True] * 4
    assert h["prim_id"][5] == 0
    qa = ([[-3, -1, 0], [-1, -1, 0], [-3, 1, 0], [-1, 1, 0]], [0, 1, 2, 1, 3, 2])
    qb = ([[1, -1, 0], [3, -1, 0], [1, 1, 0], [3, 1, 0]], [0, 1, 2, 1, 3, 2])
    o = orc.Oracle(scenes.triangles_scene([qa, qb], materials=[10, 20]))
    r = np.zeros(1, RAY_DTYPE)
    r["origin"] = [2.8, 0.8, 5]
    r["direction"] = [0, 0, -1]
    r["tfar"] = np.inf
    h = o.trace(r)
    assert h["material_id"][0] == 20 and h["geom_id"][0] == 1
    np.testing.assert_allclose(h["position"][0], [2.8, 0.8, 0], atol=1e-4)


def _rays(scene, n, seed):
    rng = np.random.default_rng(seed)
    lo, hi = scene.positions.min(0), scene.positions.max(0)
    r = np.zeros(n, RAY_DTYPE)
    r["origin"] = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    r["direction"] = d / np.linalg.norm(d, axis=1, keepdims=True) * rng.uniform(0.2, 5, (n, 1))
    r["tnear"] = 0.001
    r["tfar"] = np.where(rng.random(n) < 0.5, np.inf, rng.uniform(0.5, 20, n))
    return r


def test_bvh_traversal_equals_brute_force():
    s = scenes.scaled_synthetic_scene(4000, 128, 2)
    rays = _rays(s, 3000, 1)
    a = orc.Oracle(s, bvh=True).trace(rays)
    b = orc.Oracle(s, bvh=False).trace(rays)
    np.testing.assert_array_equal(a["prim_id"], b["prim_id"])
    hit = a["prim_id"] >= 0
    assert np.array_equal(a["t"][hit].view(np.uint32), b["t"][hit].view(np.uint32))


def test_fp32_watertight_vs_fp64():
    s = scenes.scaled_synthetic_scene(4000, 128, 3)
    rays = _rays(s, 3000, 2)
    a = orc.Oracle(s, bvh=False).trace(rays)
    b = orc.Oracle(s, bvh=False).trace(rays, mode=1)
    same = a["prim_id"] == b["prim_id"]
    assert same.mean() > 0.995
    both = same & (a["prim_id"] >= 0)
    rel = np.abs(a["t"][both] - b["t"][both]) / np.abs(b["t"][both])
    assert np.quantile(rel, 0.999) < 1e-5
    # hit points agree to 1e-5 of the coordinate scale (tiny t near the
    # origin make the relative t error alone ill-conditioned)
    dlen = np.linalg.norm(rays["direction"][both], axis=1)
    scale = np.maximum(1.0, np.abs(rays["origin"][both]).max(axis=1))
    assert (np.abs(a["t"][both] - b["t"][both]) * dlen / scale).max() < 1e-5


def test_watertight_shared_edge_has_no_gap():
    # rays through the shared edge of two triangles always hit one of them
    s = scenes.triangles_scene([([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], [0, 1, 2, 1, 3, 2])])
    o = orc.Oracle(s, bvh=False)
    t = np.linspace(0, 1, 2001)[1:-1]
    r = np.zeros(t.size, RAY_DTYPE)
    r["origin"] = np.stack([1 - t, t, np.ones_like(t)], 1)
    r["direction"] = [0.0, 0.0, -1.0]
    r["tfar"] = np.inf
    assert (o.trace(r)["prim_id"] >= 0).all()


# ------------------------------------------------------------- raster
def _np_raster(scene, W, H):
    """Independent numpy restatement of RasterizeScene (rasterizer.cpp:68-211)
    in float32, vectorised over texels, serial triangle order."""
    f = np.float32
    out_mat = np.full(W * H, -1, np.int32)
    out_tri = np.full(W * H, -1, np.int64)
    out_bary = np.zeros((W * H, 3), np.float32)
    fb_tri = []
    gid = 0
    for g in scene.geometries:
        I = scene.indices[g["index_offset"]:g["index_offset"] + g["index_count"]].reshape(-1, 3) \
            .astype(np.int64) + g["vertex_offset"]
        for t in range(len(I)):
            a, b, c = (scene.lightmap_uvs[I[t, k]].astype(f) for k in range(3))
            mn = np.minimum(np.minimum(a, b), c)
            mx = np.maximum(np.maximum(a, b), c)
            x0 = max(0, int(f(mn[0] * f(W))))
            x1 = min(W - 1, int(math.ceil(f(mx[0] * f(W)))))
            y0 = max(0, int(f(mn[1] * f(H))))
            y1 = min(H - 1, int(math.ceil(f(mx[1] * f(H)))))
            hit = False
            if x1 >= x0 and y1 >= y0:
                xs, ys = np.meshgrid(np.arange(x0, x1 + 1), np.arange(y0, y1 + 1))
                px = (xs.astype(f) + f(0.5)) / f(W)
                py = (ys.astype(f) + f(0.5)) / f(H)
                v0, v1 = b - a, c - a
                v2x, v2y = px - a[0], py - a[1]
                d00 = v0[0] * v0[0] + v0[1] * v0[1]
                d01 = v0[0] * v1[0] + v0[1] * v1[1]
                d11 = v1[0] * v1[0] + v1[1] * v1[1]
                d20 = v2x * v0[0] + v2y * v0[1]
                d21 = v2x * v1[0] + v2y * v1[1]
                den = d00 * d11 - d01 * d01
                if abs(den) >= f(1e-8):
                    vv = (d11 * d20 - d01 * d21) / den
                    ww = (d00 * d21 - d01 * d20) / den
                    uu = f(1.0) - vv - ww
                    ins = (vv >= 0) & (ww >= 0) & (uu >= 0)
                    if ins.any():
                        hit = True
                        idx = (ys * W + xs)[ins]
                        out_mat[idx] = g["material_id"]
                        out_tri[idx] = gid + t
                        out_bary[idx] = np.stack([uu[ins], vv[ins], ww[ins]], 1)
            if not hit:
                cu = (a[0] + b[0] + c[0]) / f(3.0)
                cv = (a[1] + b[1] + c[1]) / f(3.0)
                cx, cy = int(f(cu * f(W))), int(f(cv * f(H)))
                if 0 <= cx < W and 0 <= cy < H and out_mat[cy * W + cx] < 0:
                    out_mat[cy * W + cx] = g["material_id"]
                    out_tri[cy * W + cx] = gid + t
                    out_bary[cy * W + cx] = f(1.0) / f(3.0)
        gid += len(I)
    return out_mat, out_tri, out_bary


@pytest.mark.parametrize("res,ss", [(64, 1), (32, 2)])
def test_raster_matches_numpy_restatement(res, ss):
    s = scenes.box_room_scene(res, res)
    o = orc.Oracle(s).rasterize(res, res, ss)
    mat, tri, bary = _np_raster(s, res * ss, res * ss)
    np.testing.assert_array_equal(o["material_id"], mat)
    # positions from the numpy barycentrics (identity transforms)
    P = s.positions
    v = mat >= 0
    ids = tri[v]
    I = np.concatenate([s.indices[g["index_offset"]:g["index_offset"] + g["index_count"]].reshape(-1, 3)
                        .astype(np.int64) + g["vertex_offset"] for g in s.geometries])[ids]
    b = bary[v]
    pos = (P[I[:, 0]] * b[:, :1] + P[I[:, 1]] * b[:, 1:2]) + P[I[:, 2]] * b[:, 2:3]
    assert np.array_equal(o["position"][v].view(np.uint32), pos.astype(np.float32).view(np.uint32))


def test_raster_reference_kats():
    s = scenes.quad_scene()
    o = orc.Oracle(s)
    r = o.rasterize(4, 4, 1)  # rasterizer_test.cpp:7-43
    assert (r["material_id"] >= 0).all()
    np.testing.assert_allclose(r["position"][5][:2], [-0.25, -0.25], atol=1e-3)
    assert abs(r["normal"][5][2] - 1) < 1e-3
    r = o.rasterize(4, 4, 2)  # rasterizer_test.cpp:45-72
    assert r.size == 64 and (r["material_id"] >= 0).all()
    pts = np.zeros(3, SURFACE_DTYPE)  # rasterizer_test.cpp:74-87
    pts["material_id"] = [0, -1, 0]
    assert list(orc.validity_mask(pts)) == [1, 0, 1]
    m = np.zeros(16, np.uint8)  # rasterizer_test.cpp:89-126
    m[[2, 8, 9, 12, 13, 15]] = 1
    assert list(orc.downsample_mask(m, 2, 2, 2)) == [0, 1, 1, 1]


def test_raster_tiny_triangles_use_centroid_fallback():
    # |denom| < 1e-8 in uv^2 rejects small triangles (rasterizer.cpp:78); they
    # land on one centroid texel each (rasterizer.cpp:160-205)
    s = scenes.scaled_synthetic_scene(3000, 256, 1)
    r = orc.Oracle(s).rasterize(256, 256, 1)
    mat, _, _ = _np_raster(s, 256, 256)
    np.testing.assert_array_equal(r["material_id"], mat)


# ---------------------------------------------------- downsample / dilate
def _np_downsample(img, W, H, s):
    c = img.shape[1]
    a = img.reshape(H // s, s, W // s, s, c)
    acc = np.zeros((H // s, W // s, c), np.float32)
    for dy in range(s):
        for dx in range(s):
            acc = acc + a[:, dy, :, dx, :]
    return (acc * np.float32(1.0 / (s * s))).reshape(-1, c)


def test_downsample_matches_numpy():
    rng = np.random.default_rng(1)
    img = rng.normal(size=(30 * 3 * 20 * 3, 27)).astype(np.float32)
    o = orc.downsample(img, 90, 60, 3)
    assert np.array_equal(o.view(np.uint32), _np_downsample(img, 90, 60, 3).view(np.uint32))
    one = np.ones((4, 27), np.float32)  # baker_test.cpp:67-88
    assert (orc.downsample(one, 2, 2, 2) == 1).all()


def _np_dilate(data, mask, W, H, passes):
    d = data.reshape(H, W, -1).copy()
    m = mask.reshape(H, W).astype(bool).copy()
    for _ in range(passes):
        acc = np.zeros_like(d)
        cnt = np.zeros((H, W), np.int32)
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                if dx == 0 and dy == 0:
                    continue
                sh_m = np.zeros((H, W), bool)
                sh_d = np.zeros_like(d)
                ys = slice(max(0, dy), H + min(0, dy))
                yd = slice(max(0, -dy), H + min(0, -dy))
                xs = slice(max(0, dx), W + min(0, dx))
                xd = slice(max(0, -dx), W + min(0, -dx))
                sh_m[yd, xd] = m[ys, xs]
                sh_d[yd, xd] = d[ys, xs]
                acc = np.where(sh_m[..., None], acc + sh_d, acc)
                cnt += sh_m
        fill = (~m) & (cnt > 0)
        if not fill.any():
            break
        rec = (np.float32(1.0) / cnt[fill].astype(np.float32))[:, None]
        d[fill] = acc[fill] * rec
        m = m | fill
    return d.reshape(data.shape), m.reshape(-1).astype(np.uint8)


@pytest.mark.parametrize("ch,passes", [(27, 1), (27, 3), (1, 5)])
def test_dilate_matches_numpy(ch, passes):
    rng = np.random.default_rng(ch + passes)
    W, H = 37, 29
    mask = (rng.random(W * H) < 0.1).astype(np.uint8)
    data = rng.normal(size=(W * H, ch)).astype(np.float32)
    od, om = orc.dilate(data, mask, W, H, passes)
    nd, nm = _np_dilate(data, mask, W, H, passes)
    np.testing.assert_array_equal(om, nm)
    assert np.array_equal(od.view(np.uint32), nd.astype(np.float32).view(np.uint32))


def test_dilate_reference_kats():
    d = np.zeros((9, 27), np.float32)
    m = np.zeros(9, np.uint8)
    m[4] = 1
    d[4, 0] = 1
    od, om = orc.dilate(d, m, 3, 3, 1)  # dilation_test.cpp:11-33
    assert om.all() and abs(od[1, 0] - 1) < 1e-3 and abs(od[0, 0] - 1) < 1e-3
    d = np.zeros((9, 27), np.float32)
    m = np.zeros(9, np.uint8)
    m[3] = m[5] = 1
    d[3, 0] = 1
    d[5, 2] = 1
    od, om = orc.dilate(d, m, 3, 3, 1)  # dilation_test.cpp:35-58
    assert abs(od[4, 0] - 0.5) < 1e-3 and abs(od[4, 2] - 0.5) < 1e-3


# --------------------------------------------------- lights / material
def _light_scene(lights, mats=None, geoms=None):
    s = scenes.triangles_scene(geoms or [[[10, -1, 1], [10, 1, 0], [10, -1, -1]]])
    if geoms is None:
        s.normals = np.array([[-1, 0, 0]] * 3, np.float32)
        s.geometries[0]["flags"] = 1
    s.materials = mats or [dict(width=0, height=0, channels=0, offset=0, emission=0.0)]
    s.lights = lights
    return s


def test_light_closed_forms():
    from shb import LIGHT_AREA, LIGHT_DIRECTIONAL, LIGHT_POINT, LIGHT_SPOT
    # light_test.cpp:74-109: I=100 at d=10 over default 0.8 albedo -> 0.8/pi
    s = _light_scene([dict(type=LIGHT_POINT, position=(0, 10, 0), intensity=100.0)])
    o = orc.Oracle(s)
    r = o.evaluate_light_samples([0, 0, 0], [0, 1, 0], [0, 1, 0], 0, [0, 0], 1, 12345)
    assert abs(r[0] - 0.8 / math.pi) < 1e-6 and abs(r[0] - 0.249) < 0.01
    # light_test.cpp:171-201 DirectionalLightRadiance
    s = _light_scene([dict(type=LIGHT_DIRECTIONAL, direction=(0, -1, 0), intensity=2.0)])
    rad, ro, rd, tn, tf = orc.Oracle(s).light_radiance(0, [0, 0, 0], [0, 1, 0], with_brdf=-1)
    assert abs(rad[0] - 2.0) < 1e-5 and list(rd) == [0, 1, 0] and tf == np.float32(1e10)
    # light_test.cpp:203-276 SpotLightRadiance_Falloff
    spot = dict(type=LIGHT_SPOT, position=(0, 10, 0), direction=(0, -1, 0), intensity=100.0,
                cos_inner_cone=0.70710678, cos_outer_cone=0.5)
    o = orc.Oracle(_light_scene([spot]))
    assert abs(o.light_radiance(0, [0, 0, 0], [0, 1, 0], with_brdf=-1)[0][0] - 1.0) < 1e-5
    for deg, expect in ((50.0, None), (70.0, 0.0)):
        a = deg * math.pi / 180
        L = np.array([math.sin(a), -math.cos(a), 0])
        P = np.array([0, 10, 0]) + L * 10
        got = o.light_radiance(0, P, -L, with_brdf=-1)[0][0]
        if expect is None:
            expect = 100 * ((math.cos(a) - 0.5) / (0.70710678 - 0.5)) / 100
        assert abs(got - expect) < 1e-4
    # light_test.cpp:278-306 SampleAreaLight: x=10 plane, pdf 1/(1*2), 10*0.8
    s = _light_scene([dict(type=LIGHT_AREA, geometry_index=0, material_index=0, intensity=10.0)],
                     mats=[dict(width=0, height=0, channels=0, offset=0, emission=10.0)])
    p, n, rad, pdf = orc.Oracle(s).sample_area_light(0, 0, 0.3, 0.4)
    assert abs(p[0] - 10) < 1e-4 and abs(n[0] + 1) < 1e-4 and abs(pdf - 0.5) < 1e-6
    assert abs(rad[0] - 8.0) < 1e-4


def test_material_kats():
    tex = np.array([255, 128, 0, 255, 255, 255, 255, 255, 255, 128], np.uint8)
    s = scenes.triangles_scene([[[0, 0, 0], [1, 0, 0], [0, 1, 0]]])
    s.texels = tex
    s.materials = [dict(width=1, height=1, channels=3, offset=0, emission=0.0),
                   dict(width=1, height=1, channels=3, offset=3, emission=5.0),
                   dict(width=1, height=1, channels=4, offset=6, emission=0.0)]
    o = orc.Oracle(s)
    a, alpha, _ = o.material(0, [0.5, 0.5])  # material_test.cpp:43-53
    assert abs(a[0] - 1) < 1e-5 and abs(a[1] - 0.21586) < 1e-3 and abs(a[2]) < 1e-5 and alpha == 1
    _, _, e = o.material(1, [0.5, 0.5])  # material_test.cpp:88-99
    assert np.allclose(e, 5.0, atol=1e-5)
    _, alpha, _ = o.material(2, [0.5, 0.5])  # material_test.cpp:101-110
    assert abs(alpha - 128 / 255) < 1e-5
    assert orc.srgb_to_linear(255) == 1.0 and orc.srgb_to_linear(0) == 0.0


def test_cosine_sampling():
    rng = np.random.default_rng(3)
    n = np.array([0.3, -0.5, 0.8])
    n /= np.linalg.norm(n)
    cos = []
    for r1, r2 in rng.random((4000, 2)):
        d, pdf = orc.sample_material(n, r1, r2)
        c = float(np.dot(d, n))
        assert abs(np.linalg.norm(d) - 1) < 1e-5 and c >= -1e-6
        assert abs(pdf - c / math.pi) < 1e-5
        cos.append(c)
    assert abs(np.mean(cos) - 2 / 3) < 0.02  # E[cos] under cos/pi is 2/3


# ---------------------------------------------------------------- bake
@pytest.mark.parametrize("mode", [orc.MODE_REF, orc.MODE_GPU])
def test_bake_simple_quad_closed_form(mode):
    s = scenes.quad_scene()
    o = orc.Oracle(s)
    pts = np.zeros(16, SURFACE_DTYPE)
    pts["material_id"] = -1
    pts[5]["material_id"] = 0
    pts[5]["position"] = (-0.25, -0.25, 0)
    pts[5]["normal"] = (0, 0, 1)
    pts[5]["tangent"] = (1, 0, 0, 1)
    sh, vis, _, st = o.bake(pts, 4, 4, 64, 0, mode=mode)
    exp = np.repeat([0.282095, 0, 0.488603, 0, 0, 0, 0.630784, 0, 0], 3)
    np.testing.assert_allclose(sh[5], exp, rtol=2e-6, atol=1e-7)
    assert vis[5] == 1.0 and st["rays_extension"] == 64 and st["rays_shadow"] == 64


def test_two_streams_are_the_same_estimator():
    """The GPU's stream (counter-based, reservoir light pick) and the
    reference stream (mt19937, discrete_distribution) give per-texel means
    that agree within their joint Monte Carlo error."""
    s = scenes.box_room_scene(40, 40)
    o = orc.Oracle(s)
    pts = o.rasterize(40, 40, 1)
    v = pts["material_id"] >= 0
    g, _, gv, _ = o.bake(pts, 40, 40, 64, 3, mode=orc.MODE_GPU, seed=99)
    r, _, rv, _ = o.bake(pts, 40, 40, 64, 3, mode=orc.MODE_REF)
    live = (gv[v] + rv[v]) > 0
    z = (g[v] - r[v])[live] / np.sqrt(gv[v] + rv[v])[live]
    assert 0.6 < np.abs(z).mean() < 1.0
    assert (np.abs(z) > 4.5).mean() < 1e-3


def test_bake_skips_invalid_texels_and_is_deterministic():
    s = scenes.box_room_scene(24, 24)
    o = orc.Oracle(s)
    pts = o.rasterize(24, 24, 1)
    a = o.bake(pts, 24, 24, 4, 2, threads=1)[0]
    b = o.bake(pts, 24, 24, 4, 2, threads=8)[0]
    assert np.array_equal(a, b)
    assert (a[pts["material_id"] < 0] == 0).all()


def _alpha_tests():
    import sys
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import test_gpu_alpha
    return test_gpu_alpha


def _cutout(res, alphas, rgb=200):
    return _alpha_tests().cutout_room(res, alphas, rgb)


def test_gpu_stream_split_matches_the_reference_split():
    """Both modes split a path at alpha < 1 into the transmitted ray (weight
    1 - alpha) and the reflected part (weight alpha), baker.cpp:97-109, with
    different random streams: per texel within Monte Carlo error, and no bias
    in the mean over texels."""
    s = _cutout(48, _alpha_tests().ALPHAS)
    o = orc.Oracle(s)
    pts = o.rasterize(48, 48, 1)
    v = pts["material_id"] >= 0
    g_sh, _, g_var, _ = o.bake(pts, 48, 48, 128, 3, mode=orc.MODE_GPU, seed=31)
    r_sh, _, r_var, _ = o.bake(pts, 48, 48, 128, 3, mode=orc.MODE_REF)
    live = (r_var[v] + g_var[v]) > 0
    z = (g_sh[v] - r_sh[v])[live] / np.sqrt(r_var[v] + g_var[v])[live]
    assert 0.6 < np.abs(z).mean() < 1.0 and (np.abs(z) > 4.5).mean() < 1e-3
    d = (g_sh[v] - r_sh[v]).mean(axis=0)
    se = np.sqrt((r_var[v] + g_var[v]).sum(axis=0)) / v.sum()
    assert (np.abs(d) <= 4 * se + 1e-7).all()


@pytest.mark.parametrize("mode", ["gpu", "ref"])
def test_split_is_a_polynomial_in_alpha(mode):
    """With the deterministic split the traced rays and every random draw are
    the same for any 0 < alpha < 1, and each path vertex contributes a factor
    alpha or 1 - alpha: at one bounce the bake is a quadratic in a uniform
    alpha, so the third difference over four equally spaced alphas is zero
    up to rounding. Choosing one branch (piecewise constant in alpha) fails
    this."""
    m = orc.MODE_GPU if mode == "gpu" else orc.MODE_REF
    outs = []
    for a8 in (51, 102, 153, 204):
        s = _cutout(32, [a8] * 16)
        o = orc.Oracle(s)
        pts = o.rasterize(32, 32, 1)
        outs.append(o.bake(pts, 32, 32, 8, 1, mode=m, seed=3)[0].astype(np.float64))
    v = pts["material_id"] >= 0
    f0, f1, f2, f3 = (x[v] for x in outs)
    d3 = f0 - 3 * f1 + 3 * f2 - f3
    d1 = np.abs(f3 - f0)
    scale = max(np.abs(f0).max(), np.abs(f3).max())
    assert np.abs(d3).max() <= 2e-5 * scale, np.abs(d3).max() / scale
    # alpha matters: the bakes differ well beyond that tolerance
    assert d1.max() > 0.05 * scale


def test_transmission_weight_closed_form():
    """A black cube of uniform alpha at 2 bounces: every texel obeys
    f(a) = f(1) + (1 - a)^2 (f(0) - f(1)) (alpha_weight_residual says why)."""
    f = {}
    for a8 in (0, 255, 77, 179):
        s = _cutout(32, [a8] * 16, rgb=0)
        o = orc.Oracle(s)
        pts = o.rasterize(32, 32, 1)
        f[a8] = o.bake(pts, 32, 32, 8, 2, mode=orc.MODE_GPU, seed=9)[0].astype(np.float64)
    for a8 in (77, 179):
        res, spread = _alpha_tests().alpha_weight_residual(f[0], f[255], f[a8], a8 / 255.0)
        assert res < 1e-5, res
        assert spread > 0.01  # the light transmitted through the cube matters


def _cw_node(p, e, imask, child_base, tri_base, meta, lo, hi):
    """80 bytes of a compressed 8-wide node: origin p, exponents e, internal
    mask, bases, meta[8], planes lo/hi[axis][slot] (child plane = p + q 2^e)."""
    b = np.zeros(80, np.uint8)
    b[0:12] = np.frombuffer(np.asarray(p, np.float32).tobytes(), np.uint8)
    b[12:15] = np.asarray(e, np.int8).view(np.uint8)
    b[15] = imask
    b[16:20] = np.frombuffer(np.uint32(child_base).tobytes(), np.uint8)
    b[20:24] = np.frombuffer(np.uint32(tri_base).tobytes(), np.uint8)
    b[24:32] = meta
    for ax in range(3):
        b[32 + 8 * ax:40 + 8 * ax] = lo[ax]
        b[56 + 8 * ax:64 + 8 * ax] = hi[ax]
    return b


def test_cwbvh_recount_on_a_hand_built_tree():
    """The host recount of the compressed 8-wide traversal on a two-node tree
    written byte by byte: root with two one-triangle leaves over [0,1]^2 (z
    near 0.03 and 2.03) and an internal child over [3,4]^3 whose node holds a
    triangle at z = 3.5; planes on a 1/16 grid. Rays straight down: closest
    hits, node visits and triangle tests follow from the layout alone."""
    z = np.zeros(8, np.uint8)

    def planes(boxes):  # boxes per slot: ((x0,y0,z0), (x1,y1,z1)) in 1/16 units
        lo = [z.copy(), z.copy(), z.copy()]
        hi = [z.copy(), z.copy(), z.copy()]
        for s, (a, b) in boxes.items():
            for ax in range(3):
                lo[ax][s], hi[ax][s] = a[ax], b[ax]
        return lo, hi

    lo0, hi0 = planes({0: ((0, 0, 0), (16, 16, 1)), 1: ((0, 0, 32), (16, 16, 33)),
                       2: ((48, 48, 48), (64, 64, 64))})
    meta0 = z.copy()
    meta0[0], meta0[1], meta0[2] = 0x20, 0x21, 0x80
    root = _cw_node((0, 0, 0), (-4, -4, -4), 0x04, 1, 0, meta0, lo0, hi0)
    lo1, hi1 = planes({0: ((48, 48, 55), (64, 64, 57))})
    meta1 = z.copy()
    meta1[0] = 0x20
    node1 = _cw_node((0, 0, 0), (-4, -4, -4), 0x00, 0, 2, meta1, lo1, hi1)
    nodes = np.stack([root, node1])
    tri = [[(0.1, 0.1, 0.03), (0.9, 0.1, 0.03), (0.1, 0.9, 0.03)],
           [(0.1, 0.1, 2.03), (0.9, 0.1, 2.03), (0.1, 0.9, 2.03)],
           [(3.1, 3.1, 3.5), (3.9, 3.1, 3.5), (3.1, 3.9, 3.5)]]
    tris = np.zeros((3, 12), np.float32)
    for i, t in enumerate(tri):
        for k in range(3):
            tris[i, 4 * k:4 * k + 3] = t[k]
        tris[i, 3] = np.array([i], np.uint32).view(np.float32)[0]
    rays = np.zeros(4, RAY_DTYPE)
    for i, (x, y) in enumerate([(0.3, 0.3), (3.3, 3.3), (5.0, 5.0), (0.3, 0.3)]):
        rays[i]["origin"] = (x, y, 10.0)
        rays[i]["direction"] = (0.0, 0.0, -1.0)
        rays[i]["tnear"] = 0.0
        rays[i]["tfar"] = 1e30
    rays[3]["tfar"] = 9.0  # ends above the z = 0.03 triangle
    h, nv, tt = orc.cwbvh_trace(nodes, tris, rays)
    assert list(h["prim_id"]) == [1, 2, -1, 1]
    assert np.allclose(h["t"][[0, 1, 3]], [10 - 2.03, 6.5, 10 - 2.03], rtol=1e-6)
    assert list(nv) == [1, 2, 1, 1] and list(tt) == [2, 1, 0, 1]
    ha, nva, tta = orc.cwbvh_trace(nodes, tris, rays, any_hit=True)
    assert list(ha["prim_id"] >= 0) == [True, True, False, True]
    assert list(nva) == [1, 2, 1, 1] and list(tta) == [2, 1, 0, 1]


def test_two_streams_agree_with_spot_and_area_lights():
    """The two streams on the room with spot lights and 3-triangle textured
    emitters (the rng.index triangle pick of the GPU stream against the
    reference's uniform_int_distribution)."""
    s = scenes.scaled_synthetic_scene(10000, 96, 77, n_spot=4, emitter_tris=3,
                                      textured_emission=True, emitter_y=(14.3, 14.8))
    o = orc.Oracle(s)
    pts = o.rasterize(96, 96, 1)
    v = pts["material_id"] >= 0
    g, _, gv, _ = o.bake(pts, 96, 96, 64, 3, mode=orc.MODE_GPU, seed=99)
    r, _, rv, _ = o.bake(pts, 96, 96, 64, 3, mode=orc.MODE_REF)
    live = (gv[v] + rv[v]) > 0
    z = (g[v] - r[v])[live] / np.sqrt(gv[v] + rv[v])[live]
    assert 0.6 < np.abs(z).mean() < 1.0, np.abs(z).mean()
    assert (np.abs(z) > 4.5).mean() < 2e-3


def test_two_streams_agree_with_several_light_samples():
    """num_light_samples = 3: the GPU stream's inverse-CDF picks against
    the reference's discrete_distribution drawn three times."""
    s = scenes.scaled_synthetic_scene(10000, 96, 77, n_spot=4, emitter_tris=3,
                                      textured_emission=True, emitter_y=(14.3, 14.8))
    o = orc.Oracle(s)
    pts = o.rasterize(96, 96, 1)
    v = pts["material_id"] >= 0
    g, _, gv, _ = o.bake(pts, 96, 96, 64, 3, mode=orc.MODE_GPU, seed=99, num_light_samples=3)
    r, _, rv, _ = o.bake(pts, 96, 96, 64, 3, mode=orc.MODE_REF, num_light_samples=3)
    live = (gv[v] + rv[v]) > 0
    z = (g[v] - r[v])[live] / np.sqrt(gv[v] + rv[v])[live]
    assert 0.6 < np.abs(z).mean() < 1.0, np.abs(z).mean()
    assert (np.abs(z) > 4.5).mean() < 2e-3
