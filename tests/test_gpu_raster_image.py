"""GPU rasterizer, masks, downsample and dilation: bit exact against the
oracle's serial restatement (src/rasterizer.cpp:68-241,
src/baker.cpp:151-182, src/dilation.cpp:7-73) and the reference KATs."""
import numpy as np
import pytest

import shb
from shb import scenes
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def surf_equal(a, b):
    assert a.dtype == b.dtype
    np.testing.assert_array_equal(a["material_id"], b["material_id"])
    v = a["material_id"] >= 0
    for k in ("position", "normal", "tangent"):
        assert np.array_equal(a[k][v].view(np.uint32), b[k][v].view(np.uint32)), k


@pytest.mark.parametrize("w,ss", [(256, 1), (128, 2), (64, 3)])
def test_raster_box_room_bit_exact(baker, w, ss):
    s = scenes.box_room_scene(w, w)
    baker.upload_scene(s)
    g = baker.rasterize(w, w, ss)
    o = orc.Oracle(s).rasterize(w, w, ss)
    surf_equal(g, o)
    assert (g["material_id"] >= 0).mean() > 0.3


def test_raster_synthetic_bit_exact(baker):
    # many sub-texel triangles: exercises the centroid fallback and
    # overlapping-writer resolution
    s = scenes.scaled_synthetic_scene(30000, 256, 11)
    baker.upload_scene(s)
    g = baker.rasterize(256, 256, 1)
    o = orc.Oracle(s).rasterize(256, 256, 1)
    surf_equal(g, o)


def test_raster_kats(baker):
    s = scenes.quad_scene()
    baker.upload_scene(s)
    g = baker.rasterize(4, 4, 1)  # rasterizer_test.cpp:7-43
    assert (g["material_id"] >= 0).all()
    np.testing.assert_allclose(g["position"][5][:2], [-0.25, -0.25], atol=1e-3)
    assert abs(g["normal"][5][2] - 1) < 1e-3
    g = baker.rasterize(4, 4, 2)  # rasterizer_test.cpp:45-72
    assert g.size == 64 and (g["material_id"] >= 0).all()
    m = baker.validity_mask()
    assert m.all()


def test_raster_requires_tangents(baker):
    s = scenes.quad_scene(with_tangents=False)
    baker.upload_scene(s)
    with pytest.raises(shb.ShbError) as e:
        baker.rasterize(4, 4, 1)
    assert e.value.code == -1


def test_downsample_mask_kat(baker):
    pts = np.zeros(16, np.uint8)  # rasterizer_test.cpp:89-126
    pts[[2, 8, 9, 12, 13, 15]] = 1
    out = baker.downsample_mask(pts, 2, 2, 2)
    np.testing.assert_array_equal(out, [0, 1, 1, 1])


def test_downsample_bit_exact(baker):
    rng = np.random.default_rng(0)
    for ch, ss in ((27, 3), (1, 2), (27, 1)):
        img = rng.normal(size=(48 * ss * 40 * ss, ch)).astype(np.float32)
        g = baker.downsample(img, 48, 40, ss)
        o = orc.downsample(img, 48 * ss, 40 * ss, ss)
        assert np.array_equal(g.view(np.uint32), o.view(np.uint32))
    ones = np.ones((4, 27), np.float32)  # baker_test.cpp:67-88
    assert (baker.downsample(ones, 1, 1, 2) == 1.0).all()


# 28 = SH + visibility, 32 the largest count staged in LDS at once, 40 the
# one-channel-at-a-time kernel
@pytest.mark.parametrize("ch,passes", [(27, 1), (27, 4), (1, 2), (1, 16), (28, 3), (32, 2),
                                       (40, 2)])
def test_dilate_bit_exact(baker, ch, passes):
    rng = np.random.default_rng(passes * 31 + ch)
    w, h = 67, 45
    mask = (rng.random(w * h) < 0.08).astype(np.uint8)
    data = rng.normal(size=(w * h, ch)).astype(np.float32)
    gd, gm = baker.dilate(data, mask, w, h, passes)
    od, om = orc.dilate(data, mask, w, h, passes)
    np.testing.assert_array_equal(gm, om)
    assert np.array_equal(gd.view(np.uint32), od.view(np.uint32))


def test_dilate_kats(baker):
    # dilation_test.cpp:11-33 SimpleFill
    d = np.zeros((9, 27), np.float32)
    m = np.zeros(9, np.uint8)
    m[4] = 1
    d[4, 0] = 1.0
    gd, gm = baker.dilate(d, m, 3, 3, 1)
    assert gm.all() and abs(gd[1, 0] - 1) < 1e-3 and abs(gd[0, 0] - 1) < 1e-3
    # dilation_test.cpp:35-58 AverageNeighbors
    d = np.zeros((9, 27), np.float32)
    m = np.zeros(9, np.uint8)
    m[3] = m[5] = 1
    d[3, 0] = 1.0
    d[5, 2] = 1.0
    gd, gm = baker.dilate(d, m, 3, 3, 1)
    assert gm[4] == 1 and abs(gd[4, 0] - 0.5) < 1e-3 and abs(gd[4, 2] - 0.5) < 1e-3
