"""GPU traversal parity: shb_trace_rays (compressed 8-wide BVH, watertight
FP32 test) against the oracle's FP32 restatement (bit exact: same triangle,
same t/u/v bits) and its FP64 brute force (same triangle except edge ties,
t within 1e-5 relative). Reference: FindOcclusion, src/occlusion.cpp:9-107,
and the KATs of src/occlusion_test.cpp."""
import numpy as np
import pytest

import shb
from shb import RAY_DTYPE, scenes
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def seeded_rays(scene, n, seed, inside_box=None, unnormalized=True):
    rng = np.random.default_rng(seed)
    lo = scenes_bounds(scene)[0] if inside_box is None else np.array(inside_box[0])
    hi = scenes_bounds(scene)[1] if inside_box is None else np.array(inside_box[1])
    r = np.zeros(n, RAY_DTYPE)
    r["origin"] = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    if unnormalized:
        d *= rng.uniform(0.1, 10.0, (n, 1))
    d[: n // 50, 1] = 0.0  # axis-parallel components
    r["direction"] = d
    r["tnear"] = np.where(rng.random(n) < 0.5, 0.0, 0.001)
    r["tfar"] = np.where(rng.random(n) < 0.7, np.inf, rng.uniform(0.1, 20.0, n))
    return r


def scenes_bounds(scene):
    p = scene.positions
    return p.min(0), p.max(0)


def compare_exact(hg, ho):
    np.testing.assert_array_equal(hg["prim_id"], ho["prim_id"])
    hit = ho["prim_id"] >= 0
    for k in ("t", "u", "v", "position", "normal", "uv"):
        a, b = hg[k][hit], ho[k][hit]
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), k
    np.testing.assert_array_equal(hg["material_id"], ho["material_id"])
    np.testing.assert_array_equal(hg["geom_id"], ho["geom_id"])


@pytest.mark.parametrize("which", ["box_room", "synth20k"])
def test_closest_hit_bit_exact_vs_oracle(baker, which):
    s = scenes.box_room_scene() if which == "box_room" else scenes.scaled_synthetic_scene(20000, 256, 7)
    baker.upload_scene(s)
    rays = seeded_rays(s, 20000, 1)
    hg = baker.trace_rays(rays)
    o = orc.Oracle(s)
    ho = o.trace(rays, mode=0)
    compare_exact(hg, ho)
    assert (hg["prim_id"] >= 0).mean() > 0.3


@pytest.mark.parametrize("scale", [1e-9, 1.0, 1e10])
def test_closest_hit_bit_exact_at_extreme_scales(baker, scale):
    """The triangle test's 1/det comes from the corrected hardware
    reciprocal for |det| in [2^-48, 2^63] and from the IEEE division
    outside it: the box room scaled by 1e-9 (determinants ~1e-16) and by
    1e10 (~1e20) takes the division, scale 1 the reciprocal; hits stay
    bit-exact against the oracle's FP32 restatement either way."""
    s = scenes.box_room_scene()
    s.positions = (s.positions.astype(np.float64) * scale).astype(np.float32)
    baker.upload_scene(s)
    rays = seeded_rays(s, 20000, 5)
    rays["tnear"] = (rays["tnear"].astype(np.float64) * scale).astype(np.float32)
    rays["tfar"] = (rays["tfar"].astype(np.float64) * scale).astype(np.float32)
    hg = baker.trace_rays(rays)
    ho = orc.Oracle(s).trace(rays, mode=0)
    compare_exact(hg, ho)
    assert (hg["prim_id"] >= 0).mean() > 0.3
    # the determinants of the hits really lie where the case says
    P = (s.positions[s.indices].astype(np.float64)).reshape(-1, 3, 3)
    area2 = np.linalg.norm(np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]), axis=1)
    if scale < 1:
        assert area2.max() < 2.0 ** -48
    elif scale > 1e9:
        assert area2.min() > 2.0 ** 63


def test_closest_hit_vs_fp64_bruteforce(baker):
    s = scenes.scaled_synthetic_scene(6000, 128, 3)
    baker.upload_scene(s)
    rays = seeded_rays(s, 4000, 2)
    hg = baker.trace_rays(rays)
    h64 = orc.Oracle(s, bvh=False).trace(rays, mode=1)
    same = hg["prim_id"] == h64["prim_id"]
    # edge / vertex ties and grazing hits may pick a neighbour triangle
    assert same.mean() > 0.995, same.mean()
    both = same & (hg["prim_id"] >= 0)
    rel = np.abs(hg["t"][both] - h64["t"][both]) / np.maximum(np.abs(h64["t"][both]), 1e-30)
    assert np.quantile(rel, 0.999) < 1e-5, np.quantile(rel, 0.999)
    dlen = np.linalg.norm(rays["direction"][both], axis=1)
    scale = np.maximum(1.0, np.abs(rays["origin"][both]).max(axis=1))
    assert (np.abs(hg["t"][both] - h64["t"][both]) * dlen / scale).max() < 1e-5
    # a mismatch must be a near tie: the FP64 t of both candidates agrees
    mism = ~same & (hg["prim_id"] >= 0) & (h64["prim_id"] >= 0)
    if mism.any():
        r2 = np.abs(hg["t"][mism] - h64["t"][mism]) / np.abs(h64["t"][mism])
        assert r2.max() < 1e-4


def test_any_hit_matches_closest(baker):
    s = scenes.scaled_synthetic_scene(20000, 256, 9)
    baker.upload_scene(s)
    rays = seeded_rays(s, 20000, 3)
    rays["tfar"] = np.random.default_rng(4).uniform(0.01, 30.0, rays.size)
    ha = baker.trace_rays(rays, any_hit=True)
    ho = orc.Oracle(s).trace(rays, mode=2)
    np.testing.assert_array_equal(ha["prim_id"] >= 0, ho["prim_id"] >= 0)


def test_occlusion_kats(baker):
    # occlusion_test.cpp:18-62 TriangleIntersection
    s = scenes.triangles_scene([[[-1, -1, 0], [1, -1, 0], [0, 1, 0]]])
    baker.upload_scene(s)
    r = np.zeros(3, RAY_DTYPE)
    r["origin"] = [[0, 0, 1], [2, 0, 1], [0, 0, 1]]
    r["direction"] = [[0, 0, -1], [0, 0, -1], [0, 0, 1]]
    r["tfar"] = np.inf
    h = baker.trace_rays(r)
    assert h["prim_id"][0] == 0 and abs(h["position"][0][2]) < 1e-4
    assert abs(h["normal"][0][2] - 1.0) < 1e-4
    assert h["prim_id"][1] < 0 and h["prim_id"][2] < 0
    # occlusion_test.cpp:64-115 RayClipping
    r = np.zeros(3, RAY_DTYPE)
    r["origin"] = [0, 0, 1]
    r["direction"] = [0, 0, -1]
    r["tnear"] = [0.0, 1.2, 0.1]
    r["tfar"] = [0.5, np.inf, 10.0]
    h = baker.trace_rays(r)
    assert h["prim_id"][0] < 0 and h["prim_id"][1] < 0
    assert h["prim_id"][2] == 0 and abs(h["position"][2][2]) < 1e-4
    # occlusion_test.cpp:117-184 MultiGeometryMaterialTest
    quad_a = ([[-3, -1, 0], [-1, -1, 0], [-3, 1, 0], [-1, 1, 0]], [0, 1, 2, 1, 3, 2])
    quad_b = ([[1, -1, 0], [3, -1, 0], [1, 1, 0], [3, 1, 0]], [0, 1, 2, 1, 3, 2])
    s = scenes.triangles_scene([quad_a, quad_b], materials=[10, 20])
    baker.upload_scene(s)
    r = np.zeros(1, RAY_DTYPE)
    r["origin"] = [2.8, 0.8, 5.0]
    r["direction"] = [0, 0, -1]
    r["tfar"] = np.inf
    h = baker.trace_rays(r)
    assert h["material_id"][0] == 20
    np.testing.assert_allclose(h["position"][0], [2.8, 0.8, 0.0], atol=1e-4)


def test_empty_scene_misses(baker):
    s = scenes.triangles_scene([[[0, 0, 0], [1, 0, 0], [0, 1, 0]]])
    s.geometries[0]["index_count"] = 0
    baker.upload_scene(s)
    r = np.zeros(4, RAY_DTYPE)
    r["direction"] = [0, 0, 1]
    r["tfar"] = np.inf
    h = baker.trace_rays(r)
    assert (h["prim_id"] < 0).all()


def test_bvh_widths_agree(baker):
    s = scenes.scaled_synthetic_scene(8000, 128, 5)
    baker.upload_scene(s)
    rays = seeded_rays(s, 8000, 6)
    ref = orc.Oracle(s).trace(rays, mode=0)
    for w in (2, 4, 8):
        baker.set_tuning(bvh_width=w)
        compare_exact(baker.trace_rays(rays), ref)
    baker.set_tuning(bvh_width=8)


@pytest.mark.parametrize("which,width", [("box_room", 8), ("synth20k", 8), ("synth20k", 4),
                                         ("room_262k", 8)])
def test_visit_counts_equal_host_recount(baker, which, width):
    # the per-ray node visits and triangle tests behind the roofline's
    # algorithmic bytes equal a host traversal of the same exported BVH on
    # the same rays (closest hit and any hit); hits agree bit for bit
    if which == "box_room":
        s = scenes.box_room_scene()
    elif which == "synth20k":
        s = scenes.scaled_synthetic_scene(20000, 256, 7)
    else:
        s = scenes.synthetic_room_scene(64, 64)  # the 262k-triangle room
    baker.upload_scene(s)
    old = baker.tuning()
    baker.set_tuning(bvh_width=width)
    try:
        nodes, tris = baker.bvh_export()
        assert nodes.shape[0] >= 1 and tris.shape[0] == sum(
            g["index_count"] // 3 for g in s.geometries)
        rays = seeded_rays(s, 20000, 12)
        hg, nv, tt = baker.trace_rays(rays, counts=True)
        ho, nvo, tto = orc.cwbvh_trace(nodes, tris, rays)
        np.testing.assert_array_equal(hg["prim_id"], ho["prim_id"])
        hit = ho["prim_id"] >= 0
        assert np.array_equal(hg["t"][hit].view(np.uint32), ho["t"][hit].view(np.uint32))
        np.testing.assert_array_equal(nv, nvo)
        np.testing.assert_array_equal(tt, tto)
        assert nv.mean() >= 1.5 and tt.mean() > 1  # the rays do descend
        ha, nva, tta = baker.trace_rays(rays, any_hit=True, counts=True)
        hao, nvao, ttao = orc.cwbvh_trace(nodes, tris, rays, any_hit=True)
        np.testing.assert_array_equal(ha["prim_id"] >= 0, hao["prim_id"] >= 0)
        np.testing.assert_array_equal(nva, nvao)
        np.testing.assert_array_equal(tta, ttao)
    finally:
        baker.set_tuning(**old)


def test_visit_counts(baker):
    s = scenes.scaled_synthetic_scene(20000, 256, 7)
    baker.upload_scene(s)
    rays = seeded_rays(s, 5000, 8)
    hits, nv, tt = baker.trace_rays(rays, counts=True)
    assert (nv >= 1).all() and nv.mean() < 200 and tt.mean() < 400


@pytest.mark.parametrize("chunk,block", [(0, 256), (1, 256), (8, 64), (32, 64), (64, 256)])
def test_persistent_traversal_matches(baker, chunk, block):
    # the persistent dynamic-fetch form returns exactly what the one-ray-per-
    # thread form returns (same closest hits, same occlusion)
    s = scenes.scaled_synthetic_scene(20000, 256, 9)
    baker.upload_scene(s)
    rays = seeded_rays(s, 30011, 11)
    ref = orc.Oracle(s).trace(rays, mode=0)
    old = baker.tuning()
    baker.set_tuning(queue_chunk=chunk, block_size=block)
    try:
        compare_exact(baker.trace_rays(rays), ref)
        rays["tfar"] = np.random.default_rng(5).uniform(0.01, 30.0, rays.size)
        ha = baker.trace_rays(rays, any_hit=True)
        ho = orc.Oracle(s).trace(rays, mode=2)
        np.testing.assert_array_equal(ha["prim_id"] >= 0, ho["prim_id"] >= 0)
        hc, nv, tt = baker.trace_rays(rays[:5000], counts=True)
        assert (nv >= 1).all()
    finally:
        baker.set_tuning(**old)


@pytest.mark.parametrize("width", [8, 2])
def test_skewed_scene_uploads_and_traces(baker, width):
    """A chain of triangles whose sizes and positions halve (120 scales): the
    SAH alone builds a tree deeper than the traversal stack; the builder's
    depth budget brings it within, and the hits stay bit-exact."""
    n = 6000
    i = np.arange(n)
    sc = np.ldexp(1.0, -(i % 120)).astype(np.float32)
    x = (sc + np.float32(1e-7) * (i // 120)).astype(np.float32)
    tris = np.zeros((n, 3, 3), np.float32)
    tris[:, 0, 0] = x
    tris[:, 1, 0] = x + sc
    tris[:, 2, 0] = x
    tris[:, 2, 1] = sc
    tris[:, :, 2] = (i % 7)[:, None] * np.float32(1e-3)
    s = scenes.triangles_scene([t for t in tris])
    old = baker.tuning()
    try:
        baker.set_tuning(bvh_width=width)
        baker.upload_scene(s)
        rng = np.random.default_rng(5)
        r = np.zeros(20000, RAY_DTYPE)
        r["origin"] = np.stack([rng.uniform(-0.1, 2.2, 20000), rng.uniform(-0.1, 1.1, 20000),
                                np.full(20000, -1.0)], 1)
        d = np.stack([rng.normal(0, 0.05, 20000), rng.normal(0, 0.05, 20000),
                      np.ones(20000)], 1)
        r["direction"] = d
        r["tfar"] = np.inf
        hg = baker.trace_rays(r)
        ho = orc.Oracle(s).trace(r, mode=0)
        compare_exact(hg, ho)
        assert (ho["prim_id"] >= 0).mean() > 0.05
    finally:
        baker.set_tuning(**old)
