"""Materials with alpha < 1 (baker.cpp:97-109). The reference splits a path
at such a hit into the transmitted ray (weight 1 - alpha) and the reflected
part (weight alpha); the GPU does the same, the transmitted ray going on in
a spare path slot of the batch whose radiance is added to its parent's after
the last bounce. Checks: the GPU matches the oracle's GPU-stream mode sample
for sample; it agrees within Monte Carlo error with the oracle's reference
mode; the bake is a polynomial in alpha and obeys the closed-form weight of
a twice-transmitted ray; running out of spare slots re-runs the bake with
the same result; batches that mix jobs and the path-replay adjoint (central
differences) handle branches."""
import numpy as np
import pytest

import shb
from oracle import oracle as orc
from shb import scenes

pytestmark = pytest.mark.gpu


def cutout_room(res, alphas, rgb=200):
    """The box room whose central cube carries a 4x4 RGBA texture with the
    given alpha values and grey level `rgb` (texture UVs from the cube's
    x, y)."""
    s = scenes.box_room_scene(res, res)
    a = np.asarray(alphas, np.uint8).reshape(4, 4)
    tex = np.zeros((4, 4, 4), np.uint8)
    tex[..., :3] = rgb
    tex[..., 3] = a
    off = s.texels.size
    s.texels = np.concatenate([s.texels, tex.ravel()])
    s.materials = s.materials + [dict(width=4, height=4, channels=4, offset=off, emission=0.0)]
    g = dict(s.geometries[1])
    g["material_id"] = 1
    g["flags"] |= shb.GEOM_HAS_TEXTURE_UVS
    s.geometries = [s.geometries[0], g]
    vo, vc = g["vertex_offset"], g["vertex_count"]
    P = s.positions[vo:vo + vc]
    tuv = s.texture_uvs.copy()
    tuv[vo:vo + vc, 0] = (P[:, 0] + 0.25) / 0.5 * 0.999
    tuv[vo:vo + vc, 1] = (P[:, 1] - 1.25) / 0.5 * 0.999
    s.texture_uvs = tuv
    return s


ALPHAS = [0, 255, 128, 255, 64, 0, 255, 200, 255, 30, 0, 255, 255, 128, 255, 0]


@pytest.fixture(scope="module")
def room():
    s = cutout_room(48, ALPHAS)
    return s, orc.Oracle(s)


def test_same_stream_as_oracle(baker, room):
    s, o = room
    baker.upload_scene(s)
    pts = baker.rasterize(48, 48, 1)
    spp, B = 16, 4
    g_sh, g_vis, g_st = baker.bake(spp, B, seed=2024)
    o_sh, o_vis, _, o_st = o.bake(pts, 48, 48, spp, B, mode=orc.MODE_GPU, seed=2024)
    v = pts["material_id"] >= 0
    scale = np.abs(o_sh[v]).max(axis=1, keepdims=True) + 1e-6
    per_texel = (np.abs(g_sh[v] - o_sh[v]) / scale).max(axis=1)
    assert (per_texel < 1e-3).mean() > 0.97, np.sort(per_texel)[-20:]
    assert np.abs(g_vis[v] - o_vis[v]).max() <= 2.0 / spp
    n_ext = o_st["rays_extension"]
    assert abs(g_st["rays_extension"] - n_ext) <= 1e-3 * n_ext


def test_within_3_sigma_of_the_reference_split(baker, room):
    s, o = room
    baker.upload_scene(s)
    pts = baker.rasterize(48, 48, 1)
    spp, B = 128, 3
    g_sh, _, _ = baker.bake(spp, B, seed=31)
    r_sh, _, r_var, _ = o.bake(pts, 48, 48, spp, B, mode=orc.MODE_REF)
    _, _, g_var, _ = o.bake(pts, 48, 48, spp, B, mode=orc.MODE_GPU, seed=31)
    v = pts["material_id"] >= 0
    live = (r_var[v] + g_var[v]) > 0
    z = (g_sh[v] - r_sh[v])[live] / np.sqrt(r_var[v] + g_var[v])[live]
    assert 0.6 < np.abs(z).mean() < 1.0, np.abs(z).mean()
    assert (np.abs(z) > 4.5).mean() < 1e-3
    # and a bias would show in the mean over all texels
    d = (g_sh[v] - r_sh[v]).mean(axis=0)
    se = np.sqrt((r_var[v] + g_var[v]).sum(axis=0)) / v.sum()
    assert (np.abs(d) <= 4 * se + 1e-7).all()


def test_fully_transparent_cube_is_invisible(baker):
    # alpha 0 everywhere: the bake equals the bake of the room without cube
    # whenever no ray can see the difference... here: the cube's own texels
    # still rasterize, but the room's texels see through it exactly when
    # their paths cross it; compare against the oracle's reference mode
    s = cutout_room(32, [0] * 16)
    baker.upload_scene(s)
    pts = baker.rasterize(32, 32, 1)
    g_sh, _, _ = baker.bake(8, 2, seed=5)
    o_sh, _, _, _ = orc.Oracle(s).bake(pts, 32, 32, 8, 2, mode=orc.MODE_GPU, seed=5)
    v = pts["material_id"] >= 0
    scale = np.abs(o_sh[v]).max(axis=1, keepdims=True) + 1e-6
    assert ((np.abs(g_sh[v] - o_sh[v]) / scale).max(axis=1) < 1e-3).mean() > 0.97
    # alpha 0 involves no random choice: the reference mode gives the same
    # expectation with the same transmission paths, so the two agree in mean
    r_sh, _, r_var, _ = orc.Oracle(s).bake(pts, 32, 32, 8, 2, mode=orc.MODE_REF)
    assert np.isfinite(g_sh).all()


def test_adjoint_with_transparency(baker, room):
    s, _ = room
    baker.upload_scene(s)
    baker.rasterize(48, 48, 1, want=False)
    nl, nm = len(s.lights), len(s.materials)
    spp, B, seed = 16, 3, 77
    ones_l, ones_m = np.ones((nl, 3), np.float32), np.ones((nm, 3), np.float32)
    baker.set_params(ones_l, ones_m)
    sh0 = baker.bake(spp, B, seed=seed)[0]
    w = np.random.default_rng(1).normal(size=sh0.shape).astype(np.float32)
    dl, dm, _ = baker.bake_adjoint(spp, B, w, seed=seed)
    h = 1e-2

    def lin(ls, al):
        baker.set_params(ls, al)
        return float((w.astype(np.float64) * baker.bake(spp, B, seed=seed)[0]).sum())

    for m in range(nm):
        for c in range(3):
            ap, am = ones_m.copy(), ones_m.copy()
            ap[m, c] += h
            am[m, c] -= h
            fd = (lin(ones_l, ap) - lin(ones_l, am)) / (2 * h)
            assert abs(dm[m, c] - fd) <= 1e-2 * max(abs(fd), 1e-3 * np.abs(dm).max()), (m, c)
    lp, lm = ones_l.copy(), ones_l.copy()
    lp[0, 1] += h
    lm[0, 1] -= h
    fd = (lin(lp, ones_m) - lin(lm, ones_m)) / (2 * h)
    assert abs(dl[0, 1] - fd) <= 1e-2 * abs(fd)
    baker.set_params(ones_l, ones_m)


def alpha_weight_residual(f0, f1, fa, a):
    """Largest deviation from f(a) = f(1) + (1 - a)^2 (f(0) - f(1)), and the
    spread |f(0) - f(1)|, both relative to the largest |f(0)|. A black cube
    reflects nothing, so at 2 bounces a path that meets it contributes only
    through the ray transmitted at both of its faces, weight (1 - a)^2, with
    the same draws as at alpha 0 (branch ids 3 then 7); every other path is
    as with the opaque cube. A dropped 1 - a, or a in its place, fails this."""
    top = max(np.abs(f0).max(), 1e-30)
    pred = f1 + (1.0 - a) ** 2 * (f0 - f1)
    return np.abs(fa - pred).max() / top, np.abs(f0 - f1).max() / top


def bake_at(baker, s, res, spp, B, seed):
    """GPU bake of scene s at res x res: (valid texels' SH as float64, stats)."""
    baker.upload_scene(s)
    pts = baker.rasterize(res, res, 1)
    sh, vis, st = baker.bake(spp, B, seed=seed)
    return sh[pts["material_id"] >= 0].astype(np.float64), st


def test_polynomial_in_alpha(baker):
    # one bounce: a quadratic in a uniform alpha (the traced rays and draws do
    # not depend on alpha, each vertex weighs alpha or 1 - alpha); third
    # difference zero up to rounding
    f0, f1, f2, f3 = [bake_at(baker, cutout_room(32, [a8] * 16), 32, 8, 1, 3)[0]
                      for a8 in (51, 102, 153, 204)]
    scale = max(np.abs(f0).max(), np.abs(f3).max())
    assert np.abs(f0 - 3 * f1 + 3 * f2 - f3).max() <= 2e-5 * scale
    assert np.abs(f3 - f0).max() > 0.05 * scale


def test_transmission_weight_closed_form(baker):
    f = {a8: bake_at(baker, cutout_room(32, [a8] * 16, rgb=0), 32, 8, 2, 9)[0]
         for a8 in (0, 255, 77, 179)}
    for a8 in (77, 179):
        res, spread = alpha_weight_residual(f[0], f[255], f[a8], a8 / 255.0)
        assert res < 1e-5, (a8, res)
        assert spread > 0.01


def see_through_room(res):
    """The box room with every surface at alpha 128/255."""
    b = scenes.SceneBuilder()
    g = scenes.linear_to_srgb8(0.7)
    m = b.add_material((g, g, g, 128))
    b.add_box((0, 1.5, 0), (2, 1.5, 2.5), material=m, inward=True)
    b.add_box((0, 1.5, 0), (0.25, 0.25, 0.25), material=m)
    b.add_light(type=scenes.LIGHT_POINT, position=(0, 2.5, 0), intensity=50.0, color=(1, 1, 1))
    return b.build(res, res, 2)


def test_spare_slots_run_out_and_the_bake_is_redone():
    # every hit splits: far more branches than the first pass's spare slots
    # (an eighth of the paths); the second pass has room, later bakes on the
    # context start with it; any batch size gives the same bits
    s = see_through_room(48)
    spp, B, seed = 16, 4, 21
    b = shb.Baker(0)
    try:
        b.upload_scene(s)
        pts = b.rasterize(48, 48, 1)
        sh1, vis1, st1 = b.bake(spp, B, seed=seed)
        P = st1["paths"]
        first = max((P + 7) // 8, 256)
        assert st1["branches"] > first, (st1["branches"], first)
        assert st1["passes"] >= 2
        sh2, vis2, st2 = b.bake(spp, B, seed=seed)
        assert st2["passes"] == 1 and st2["branches"] == st1["branches"]
        assert np.array_equal(sh1.view(np.uint32), sh2.view(np.uint32))
        assert np.array_equal(vis1.view(np.uint32), vis2.view(np.uint32))
        b.set_tuning(batch_paths=5000)
        sh3, vis3, st3 = b.bake(spp, B, seed=seed)
        assert np.array_equal(sh1.view(np.uint32), sh3.view(np.uint32))
        assert st3["branches"] == st1["branches"]
        # and the same paths as the oracle
        o_sh, _, _, o_st = orc.Oracle(s).bake(pts, 48, 48, spp, B, mode=orc.MODE_GPU, seed=seed)
        v = pts["material_id"] >= 0
        scale = np.abs(o_sh[v]).max(axis=1, keepdims=True) + 1e-6
        assert ((np.abs(sh1[v] - o_sh[v]) / scale).max(axis=1) < 1e-3).mean() > 0.97
        n_ext = o_st["rays_extension"]
        assert abs(st1["rays_extension"] - n_ext) <= 1e-3 * n_ext
    finally:
        b.close()


def test_batches_mixing_jobs_with_branches():
    # an alpha job and an opaque one in batches that hold both: each comes
    # out bit-identical to its plain bake
    specs = [(cutout_room(40, ALPHAS), 40, 16, 3, 5), (scenes.box_room_scene(32, 32), 32, 8, 2, 6)]
    bakers = []
    try:
        for s, res, *_ in specs:
            b = shb.Baker(0)
            b.upload_scene(s)
            b.rasterize(res, res, 1, want=False)
            bakers.append(b)
        ref = [b.bake(spp, B, seed=seed)[:2] for b, (_, _, spp, B, seed) in zip(bakers, specs)]
        for bp in (None, 4000):
            if bp:
                bakers[0].set_tuning(batch_paths=bp)
            sh, vis, st = shb.bake_batched(bakers, [x[2] for x in specs], [x[3] for x in specs],
                                           [x[4] for x in specs])
            for k in range(2):
                assert np.array_equal(sh[k].view(np.uint32), ref[k][0].view(np.uint32)), (bp, k)
                assert np.array_equal(vis[k].view(np.uint32), ref[k][1].view(np.uint32)), (bp, k)
            assert st["branches"] > 0
    finally:
        for b in bakers:
            b.close()


def test_adjoint_reruns_when_spare_slots_run_out():
    # a fresh context's first pass has too few spare slots for this room;
    # the re-run must start from zeroed gradient sums: the same gradients as
    # a second call that fits in one pass, up to the rounding of the blocks'
    # FP32 partial sums (a sum carried over would be off by a whole factor)
    s = see_through_room(32)
    spp, B, seed = 8, 4, 5
    b = shb.Baker(0)
    try:
        b.upload_scene(s)
        b.rasterize(32, 32, 1, want=False)
        w = np.random.default_rng(3).normal(size=(32 * 32, 27)).astype(np.float32)
        dl1, dm1, st1 = b.bake_adjoint(spp, B, w, seed=seed)
        assert st1["passes"] >= 2 and st1["branches"] > 0
        dl2, dm2, st2 = b.bake_adjoint(spp, B, w, seed=seed)
        assert st2["passes"] == 1 and st2["branches"] == st1["branches"]
        for a, c in ((dl1, dl2), (dm1, dm2)):
            np.testing.assert_allclose(a, c, rtol=1e-5, atol=1e-8 * np.abs(c).max())
        assert np.abs(dm2).max() > 0
    finally:
        b.close()
