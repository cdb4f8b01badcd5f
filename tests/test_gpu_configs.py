"""The benchmark configurations at their full geometry and lightmap size, at
low sample counts, checked on the GPU (BakeSHLightMap, src/baker.cpp:186-304,
and the post-bake steps of src/main.cpp:138-163):

  C1  box room + cube, 256^2, 64 spp, 3 bounces: same stream as the oracle
      and within the Monte Carlo error of its reference stream (every 4th
      texel row), the post-bake pipeline bit exact;
  C2  262k-triangle room, 1024^2 x ss3 (3072^2 texels), 3 bounces at 2 spp:
      same stream as the oracle on every 12th texel row;
  C3  the 16-job mixed batch (1k-500k triangles, 64^2-2048^2, 1-5 bounces)
      at 1/32 of its sample counts: every job of the batched bake bit-
      identical to its plain bake;
  C4  the adjoint on the 262k-triangle room at 512^2, 3 bounces, 8 spp:
      gradients against central finite differences under common random
      numbers;
  C5  4.2M triangles, 256 lights, 4096^2 x ss2 (8192^2 texels), 5 bounces
      at 2 spp: same stream as the oracle on every 64th texel row, and the
      bake as 64 tile jobs bit-identical to the plain bake.
The full sample counts run in bench.py."""
import numpy as np
import pytest

import bench
import shb
from shb import scenes
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def same_stream(g_sh, g_vis, o_sh, o_vis, pts, row_stride, width, spp):
    """Per-texel agreement of the GPU with the oracle's GPU-stream mode on
    the rows the oracle baked."""
    rows = (np.arange(pts.size) // width) % row_stride == 0
    v = (pts["material_id"] >= 0) & rows
    scale = np.abs(o_sh[v]).max(axis=1, keepdims=True) + 1e-6
    per_texel = (np.abs(g_sh[v] - o_sh[v]) / scale).max(axis=1)
    share = float((per_texel < 1e-3).mean())
    print(f"  {int(v.sum())} texels compared, share within 1e-3: {share:.4f}", flush=True)
    assert v.sum() > 500
    assert share > 0.97, np.sort(per_texel)[-20:]
    assert np.abs(g_vis[v] - o_vis[v]).max() <= 2.0 / spp


def test_c1_full_config(baker):
    print("C1: box room 256^2, 64 spp, 3 bounces", flush=True)
    s = scenes.box_room_scene(256, 256)
    o = orc.Oracle(s)
    baker.upload_scene(s)
    pts = baker.rasterize(256, 256, 1)
    g_sh, g_vis, st = baker.bake(64, 3, seed=12345)
    o_sh, o_vis, _, _ = o.bake(pts, 256, 256, 64, 3, mode=orc.MODE_GPU, seed=12345,
                               row_stride=4)
    same_stream(g_sh, g_vis, o_sh, o_vis, pts, 4, 256, 64)
    r_sh, _, r_var, _ = o.bake(pts, 256, 256, 64, 3, mode=orc.MODE_REF, row_stride=4)
    _, _, g_var, _ = o.bake(pts, 256, 256, 64, 3, mode=orc.MODE_GPU, seed=12345, row_stride=4)
    rows = (np.arange(pts.size) // 256) % 4 == 0
    v = (pts["material_id"] >= 0) & rows
    live = (r_var[v] + g_var[v]) > 0
    z = (g_sh[v] - r_sh[v])[live] / np.sqrt(r_var[v] + g_var[v])[live]
    print(f"  mean |z| {np.abs(z).mean():.3f}, share |z| > 3 {(np.abs(z) > 3).mean():.4f}")
    assert 0.6 < np.abs(z).mean() < 1.0
    assert (np.abs(z) > 4.5).mean() < 2e-3
    # the configuration's post-bake steps (ss 1, dilation 0): the mask only
    sh, vis, mask, _ = baker.bake_lightmap(64, 3, 0, seed=12345)
    assert np.array_equal(sh.view(np.uint32), g_sh.view(np.uint32))
    np.testing.assert_array_equal(mask, orc.validity_mask(pts))


def test_c2_full_geometry_and_size(baker):
    print("C2: 262k-triangle room, 1024^2 ss3, 3 bounces, 2 spp", flush=True)
    s = scenes.synthetic_room_scene(1024, 1024)
    baker.upload_scene(s)
    pts = baker.rasterize(1024, 1024, 3)
    g_sh, g_vis, st = baker.bake(2, 3, seed=12345)
    assert st["valid_texels"] == int((pts["material_id"] >= 0).sum())
    o_sh, o_vis, _, _ = orc.Oracle(s).bake(pts, 3072, 3072, 2, 3, mode=orc.MODE_GPU,
                                           seed=12345, row_stride=12)
    same_stream(g_sh, g_vis, o_sh, o_vis, pts, 12, 3072, 2)


def test_c3_batched_bit_identical_at_full_sizes():
    specs = bench.c3_jobs()
    print("C3: " + ", ".join(f"{j['tris']}t/{j['res']}^2/B{j['bounces']}" for j in specs),
          flush=True)
    bakers = []
    try:
        for sp in specs:
            b = shb.Baker(0)
            b.upload_scene(scenes.scaled_synthetic_scene(sp["tris"], sp["res"], sp["seed"]))
            b.rasterize(sp["res"], sp["res"], 1, want=False)
            bakers.append(b)
        spp = [max(1, sp["spp"] // 32) for sp in specs]
        bnc = [sp["bounces"] for sp in specs]
        seeds = [sp["seed"] for sp in specs]
        sh, vis, st = shb.bake_batched(bakers, spp, bnc, seeds)
        rays = 0
        for k, b in enumerate(bakers):
            p_sh, p_vis, p_st = b.bake(spp[k], bnc[k], seed=seeds[k])
            rays += p_st["rays_extension"] + p_st["rays_shadow"]
            assert np.array_equal(sh[k].view(np.uint32), p_sh.view(np.uint32)), k
            assert np.array_equal(vis[k].view(np.uint32), p_vis.view(np.uint32)), k
        assert st["rays_extension"] + st["rays_shadow"] == rays
    finally:
        for b in bakers:
            b.close()


def test_c4_adjoint_against_finite_differences(baker):
    print("C4: adjoint, 262k-triangle room 512^2, 3 bounces, 8 spp", flush=True)
    s = scenes.synthetic_room_scene(512, 512)
    baker.upload_scene(s)
    baker.rasterize(512, 512, 1, want=False)
    nl, nm = len(s.lights), len(s.materials)
    spp, B, seed = 8, 3, 4242
    ones_l, ones_a = np.ones((nl, 3), np.float32), np.ones((nm, 3), np.float32)
    target = baker.bake(spp, B, seed=seed + 1)[0].astype(np.float64)  # another draw
    baker.set_params(ones_l, ones_a)

    def loss(ls, al):
        baker.set_params(ls, al)
        d = baker.bake(spp, B, seed=seed)[0].astype(np.float64) - target
        return 0.5 * float((d * d).sum())

    sh = baker.bake(spp, B, seed=seed)[0].astype(np.float64)
    g = (sh - target).astype(np.float32)
    dl, dm, _ = baker.bake_adjoint(spp, B, g, seed=seed)
    try:
        for kind, grad in (("light", dl), ("albedo", dm)):
            for flat in np.argsort(-np.abs(grad).ravel())[:2]:
                i, c = int(flat // 3), int(flat % 3)
                h = 1e-2
                lp, lm, ap, am = ones_l.copy(), ones_l.copy(), ones_a.copy(), ones_a.copy()
                if kind == "light":
                    lp[i, c], lm[i, c] = 1 + h, 1 - h
                    fd = (loss(lp, ones_a) - loss(lm, ones_a)) / (2 * h)
                else:
                    ap[i, c], am[i, c] = 1 + h, 1 - h
                    fd = (loss(ones_l, ap) - loss(ones_l, am)) / (2 * h)
                rel = abs(float(grad[i, c]) - fd) / abs(fd)
                print(f"  {kind}[{i}][{c}] adjoint {grad[i, c]:.6e} fd {fd:.6e} rel {rel:.2e}")
                assert rel < 1e-2
    finally:
        baker.set_params(ones_l, ones_a)


@pytest.mark.timeout(600)
def test_c5_full_geometry_and_size():
    print("C5: 4.2M triangles, 256 lights, 4096^2 ss2, 5 bounces, 2 spp", flush=True)
    s = scenes.stress_scene(4096)
    print(f"  scene: {s.num_triangles} triangles, {len(s.lights)} lights", flush=True)
    b = shb.Baker(0)
    try:
        b.upload_scene(s)
        pts = b.rasterize(4096, 4096, 2)
        g_sh, g_vis, st = b.bake(2, 5, seed=12345)
        print(f"  GPU: {st['valid_texels']} texels, {st['rays_extension'] + st['rays_shadow']} "
              f"rays, {st['seconds_total']:.2f}s", flush=True)
        # the configuration's form: 64 tile jobs (8 x 8 tiles of 1024^2) in one
        # batched run, bit-identical to the plain bake
        tiles = [(x, y, 1024, 1024) for y in range(0, 8192, 1024) for x in range(0, 8192, 1024)]
        t_sh, t_vis, tst = b.bake_tiled(2, 5, tiles, seed=12345)
        print(f"  64 tiles: {tst['seconds_total']:.2f}s", flush=True)
        assert np.array_equal(t_sh.view(np.uint32), g_sh.view(np.uint32))
        assert np.array_equal(t_vis.view(np.uint32), g_vis.view(np.uint32))
        del t_sh, t_vis
    finally:
        b.close()
    o_sh, o_vis, _, _ = orc.Oracle(s).bake(pts, 8192, 8192, 2, 5, mode=orc.MODE_GPU,
                                           seed=12345, row_stride=64)
    same_stream(g_sh, g_vis, o_sh, o_vis, pts, 64, 8192, 2)
