"""Lightmap atlas (shb_atlas_*, replacing CreateAtlasGeometries and
atlas_internal::CalculateGeometryScales, src/atlas.cpp:22-250). The
reference charts and packs with xatlas (git master, not vendored): the
chart layout is this build's own, so it is checked by the properties the
reference relies on, not against xatlas output.

Pinned by: the atlas_test.cpp KATs (the four CalculateGeometryScales cases,
SimpleQuad); on the reference's glTF fixtures and seeded synthetic meshes:
every output vertex is a copy of its xref source, faces keep their order and
corners, lightmap UVs lie in [0, 1], the bounding boxes of different charts are
at least `padding` texels apart, and every face's texel area
is its scaled world area times density^2 times cos(angle to the chart's
projection axis) with that angle within 25 degrees. Host code: CPU."""
import math
import os

import numpy as np
import pytest

import shb
from shb import scenes

GL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "gltf")


def scene_of(geos, mats):
    """geos: list of dict(P, N, UV, T, I, material); mats: list of (w, h)."""
    s = scenes.Scene()
    P, N, UV, T, I, G = [], [], [], [], [], []
    nv = ni = 0
    for g in geos:
        n = len(g["P"])
        P.append(np.asarray(g["P"], np.float32))
        N.append(np.asarray(g.get("N", np.tile([0, 0, 1], (n, 1))), np.float32))
        UV.append(np.asarray(g["UV"], np.float32))
        T.append(np.asarray(g.get("T", np.tile([1, 0, 0, 1], (n, 1))), np.float32))
        idx = np.asarray(g["I"], np.uint32)
        I.append(idx)
        G.append(dict(vertex_offset=nv, vertex_count=n, index_offset=ni, index_count=len(idx),
                      material_id=g.get("material", 0),
                      flags=shb.GEOM_HAS_NORMALS | shb.GEOM_HAS_TEXTURE_UVS | shb.GEOM_HAS_TANGENTS,
                      transform=g.get("transform", scenes.IDENTITY)))
        nv += n
        ni += len(idx)
    s.positions, s.normals = np.concatenate(P), np.concatenate(N)
    s.texture_uvs, s.tangents = np.concatenate(UV), np.concatenate(T)
    s.lightmap_uvs = np.zeros((nv, 2), np.float32)
    s.indices = np.concatenate(I)
    s.geometries = G
    s.materials = [dict(width=w, height=h, channels=0, offset=0, emission=0.0) for w, h in mats]
    s.texels = np.zeros(0, np.uint8)
    s.lights = []
    return s


# ------------------------------------------------ atlas_test.cpp KATs
def test_scales_basic():  # atlas_test.cpp:67-88
    s = scene_of([dict(P=np.zeros((3, 3)), UV=[[0, 0], [1, 0], [0, 1]], I=[0, 1, 2])], [(100, 100)])
    assert shb.geometry_scales(s)[0] == pytest.approx(100.0, abs=1e-4)


def test_scales_tiling():  # atlas_test.cpp:90-111
    s = scene_of([dict(P=np.zeros((3, 3)), UV=[[0, 0], [2, 0], [0, 2]], I=[0, 1, 2])], [(10, 10)])
    assert shb.geometry_scales(s)[0] == pytest.approx(20.0, abs=1e-4)


def test_scales_clamping():  # atlas_test.cpp:113-148
    g = [dict(P=np.zeros((2, 3)), UV=[[0, 0], [1, 1]], I=[], material=k) for k in range(3)]
    s = scene_of(g, [(10, 10), (10, 10), (1000, 1000)])
    np.testing.assert_allclose(shb.geometry_scales(s), [10, 10, 50], atol=1e-4)


def test_scales_density_multiplier():  # atlas_test.cpp:150-166
    s = scene_of([dict(P=np.zeros((2, 3)), UV=[[0, 0], [1, 1]], I=[])], [(100, 100)])
    assert shb.geometry_scales(s, 2.0)[0] == pytest.approx(200.0, abs=1e-4)


def test_simple_quad():  # atlas_test.cpp:9-65
    g = dict(P=[[-1, -1, 0], [1, -1, 0], [-1, 1, 0], [1, 1, 0]],
             UV=[[0, 0], [1, 0], [0, 1], [1, 1]], I=[0, 1, 2, 2, 1, 3])
    s = scene_of([g], [(256, 256)])
    a, info = shb.create_atlas(s, 256, 2)
    assert info["width"] >= 0 and info["height"] >= 0
    assert len(a.geometries) == 1
    uv = a.lightmap_uvs
    assert uv.shape[0] == a.positions.shape[0]
    assert (uv >= 0).all() and (uv <= 1).all()
    assert a.geometries[0]["index_count"] == 6
    assert a.positions.shape[0] >= 4
    assert info["charts"] == 1  # a planar quad is one chart


# ---------------------------------------------------------- properties
def face_normals(P, I):
    t = P[I.reshape(-1, 3)].astype(np.float64)
    c = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    return c, 0.5 * np.linalg.norm(c, axis=1)


def check_atlas(s, res, pad, dm=1.0):
    a, info = shb.create_atlas(s, res, pad, dm)
    scales = shb.geometry_scales(s, dm)
    d = info["texels_per_unit"]
    xref = info["xref"]
    assert (a.lightmap_uvs >= 0).all() and (a.lightmap_uvs <= 1).all()
    boxes = []
    for gi, (gin, gout) in enumerate(zip(s.geometries, a.geometries)):
        assert gout["material_id"] == gin["material_id"] and gout["transform"] == gin["transform"]
        assert gout["flags"] & shb.GEOM_HAS_LIGHTMAP_UVS
        vo, vc = gout["vertex_offset"], gout["vertex_count"]
        src = gin["vertex_offset"] + xref[vo:vo + vc]
        for k in ("positions", "normals", "texture_uvs", "tangents"):
            np.testing.assert_array_equal(getattr(a, k)[vo:vo + vc], getattr(s, k)[src], err_msg=k)
        Iin = s.indices[gin["index_offset"]:gin["index_offset"] + gin["index_count"]]
        Iout = a.indices[gout["index_offset"]:gout["index_offset"] + gout["index_count"]]
        np.testing.assert_array_equal(xref[vo + Iout], Iin)
        # world (scaled) and texel areas per face
        T = np.array(gin["transform"], np.float64).reshape(3, 4)
        Pw = (s.positions[gin["vertex_offset"]:gin["vertex_offset"] + gin["vertex_count"]].astype(np.float64)
              @ T[:, :3].T + T[:, 3]) * scales[gi]
        c, area = face_normals(Pw, Iin.astype(np.int64))
        L = a.lightmap_uvs[vo:vo + vc].astype(np.float64) * res
        tl = L[Iout.reshape(-1, 3)]
        tarea = 0.5 * np.abs((tl[:, 1, 0] - tl[:, 0, 0]) * (tl[:, 2, 1] - tl[:, 0, 1]) -
                             (tl[:, 2, 0] - tl[:, 0, 0]) * (tl[:, 1, 1] - tl[:, 0, 1]))
        # float32 UVs carry ~1e-4 texel of rounding: ratios on faces of a
        # texel or more, an absolute bound on the rest
        exp = area * d * d
        big = exp >= 1.0
        ratio = tarea[big] / exp[big]
        if big.any():
            assert ratio.min() >= math.cos(math.radians(25)) - 1e-3, ratio.min()
            assert ratio.max() <= 1 + 1e-3, ratio.max()
        assert (tarea <= exp + 1e-3).all()
        assert (tarea >= math.cos(math.radians(25)) * exp - 1e-3).all()
        ch = info["chart"][vo:vo + vc]
        used = np.unique(Iout)
        for cid in np.unique(ch[used]):
            v = used[ch[used] == cid]
            boxes.append((L[v, 0].min(), L[v, 1].min(), L[v, 0].max(), L[v, 1].max()))
        # a face's three corners belong to one chart
        t3 = ch[Iout.reshape(-1, 3)]
        assert (t3 == t3[:, :1]).all()
    assert len(boxes) == info["charts"]
    # chart boxes at least `pad` texels apart
    B = np.array(boxes)
    n = len(B)
    gapx = np.maximum(B[None, :, 0] - B[:, None, 2], B[:, None, 0] - B[None, :, 2])
    gapy = np.maximum(B[None, :, 1] - B[:, None, 3], B[:, None, 1] - B[None, :, 3])
    gap = np.maximum(gapx, gapy)
    np.fill_diagonal(gap, np.inf)
    assert gap.min() >= pad - 1e-3 or n == 1, gap.min()
    return a, info


@pytest.mark.parametrize("fx", ["box/scene.gltf", "plane/scene.gltf", "cube/Cube.gltf",
                                "DirectionalLight/scene.gltf"])
def test_fixture_atlases(fx):
    s, _ = shb.load_gltf(os.path.join(GL, fx))
    check_atlas(s, 256, 2)


def test_lamp_fixture_atlas():
    s, _ = shb.load_gltf(os.path.join(GL, "LightsPunctualLamp", "LightsPunctualLamp.gltf"))
    check_atlas(s, 1024, 4)


def sphere(n_lat, n_lon, r=1.0, c=(0, 0, 0)):
    th = np.linspace(0, np.pi, n_lat + 1)
    ph = np.linspace(0, 2 * np.pi, n_lon + 1)
    T, Ph = np.meshgrid(th, ph, indexing="ij")
    P = np.stack([np.sin(T) * np.cos(Ph), np.cos(T), np.sin(T) * np.sin(Ph)], -1).reshape(-1, 3)
    I = []
    for i in range(n_lat):
        for j in range(n_lon):
            a, b = i * (n_lon + 1) + j, i * (n_lon + 1) + j + 1
            cc, d = a + n_lon + 1, b + n_lon + 1
            if i > 0:
                I += [a, cc, b]
            if i < n_lat - 1:
                I += [b, cc, d]
    UV = np.stack([Ph.ravel() / (2 * np.pi), T.ravel() / np.pi], -1)
    return dict(P=P * r + c, N=P, UV=UV, I=I)


def test_curved_and_transformed_meshes():
    rng = np.random.default_rng(4)
    geos = [sphere(12, 24, 1.0, (0, 0, 0)), sphere(6, 10, 0.3, (3, 0, 0))]
    # a random triangle soup, rotated and translated by its transform
    P = rng.normal(size=(90, 3))
    geos.append(dict(P=P, N=np.tile([0, 1, 0], (90, 1)), UV=rng.random((90, 2)),
                     I=np.arange(90), transform=(0, -1, 0, 5, 1, 0, 0, 1, 0, 0, 1, 2)))
    s = scene_of(geos, [(64, 64)])
    check_atlas(s, 512, 3)
    check_atlas(s, 512, 0, dm=2.0)


def test_too_many_charts_is_a_limit_error():
    rng = np.random.default_rng(5)
    P = rng.normal(size=(3 * 400, 3))
    s = scene_of([dict(P=P, UV=rng.random((1200, 2)), I=np.arange(1200))], [(8, 8)])
    with pytest.raises(shb.ShbError) as e:
        shb.create_atlas(s, 16, 2)
    assert e.value.code == -5  # SHB_ERR_LIMIT
