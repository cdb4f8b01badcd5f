"""Batched bake (shb_bake_batched), hipGraph replay and the autotuner.

The batched form bakes several lightmaps through shared wavefront batches
over their concatenated BVHs; the random stream is keyed by (seed, texel,
sample, depth), so every job must come out bit-identical to a plain bake of
the same context, whatever the batch packing, launch shapes or replay."""
import numpy as np
import pytest

import shb
from shb import scenes

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.fixture(scope="module")
def jobs():
    specs = [  # (scene, atlas, ss, spp, bounces, seed)
        (scenes.box_room_scene(48, 48), 48, 1, 16, 3, 7),
        (scenes.scaled_synthetic_scene(3000, 64, 2), 64, 1, 8, 1, 8),
        (scenes.box_room_scene(32, 32), 32, 2, 4, 5, 9),
        (scenes.scaled_synthetic_scene(12000, 96, 4), 96, 1, 32, 2, 10),
    ]
    out = []
    for s, res, ss, spp, B, seed in specs:
        b = shb.Baker(0)
        b.upload_scene(s)
        b.rasterize(res, res, ss, want=False)
        out.append((b, spp, B, seed))
    yield out
    for b, *_ in out:
        b.close()


def plain(jobs):
    return [b.bake(spp, B, seed=seed)[:2] for b, spp, B, seed in jobs]


def test_batched_bit_identical_to_plain(jobs):
    ref = plain(jobs)
    bakers = [j[0] for j in jobs]
    sh, vis, st = shb.bake_batched(bakers, [j[1] for j in jobs], [j[2] for j in jobs],
                                   [j[3] for j in jobs])
    for k in range(len(jobs)):
        assert np.array_equal(bits(sh[k]), bits(ref[k][0])), k
        assert np.array_equal(bits(vis[k]), bits(ref[k][1])), k
    assert st["valid_texels"] == sum(int((b.validity_mask() > 0).sum()) for b in bakers)


def test_batched_small_batches_mix_jobs_and_replay(jobs):
    # a tiny batch size splits jobs over many batches whose segments mix jobs;
    # the second call replays the captured graph
    ref = plain(jobs)
    bakers = [j[0] for j in jobs]
    old = bakers[0].tuning()
    bakers[0].set_tuning(batch_paths=3000)
    try:
        for _ in range(2):
            sh, vis, _ = shb.bake_batched(bakers, [j[1] for j in jobs], [j[2] for j in jobs],
                                          [j[3] for j in jobs])
            for k in range(len(jobs)):
                assert np.array_equal(bits(sh[k]), bits(ref[k][0])), k
                assert np.array_equal(bits(vis[k]), bits(ref[k][1])), k
        # a new seed changes the recorded launches: the graph is captured again
        sh2, _, _ = shb.bake_batched(bakers, [j[1] for j in jobs], [j[2] for j in jobs],
                                     [j[3] + 1 for j in jobs])
        assert not np.array_equal(bits(sh2[0]), bits(ref[0][0]))
        r2 = bakers[0].bake(jobs[0][1], jobs[0][2], seed=jobs[0][3] + 1)[0]
        assert np.array_equal(bits(sh2[0]), bits(r2))
    finally:
        bakers[0].set_tuning(**old)


def test_plain_graph_flag_replays_identically(jobs):
    b, spp, B, seed = jobs[0]
    ref = b.bake(spp, B, seed=seed)[0]
    for _ in range(3):
        g = b.bake(spp, B, seed=seed, flags=shb.BAKE_USE_GRAPH)[0]
        assert np.array_equal(bits(g), bits(ref))


def test_batched_errors(jobs):
    b = jobs[0][0]
    with pytest.raises(shb.ShbError) as e:
        shb.bake_batched([b, b], 4, 1)
    assert e.value.code == -1
    fresh = shb.Baker(0)
    with pytest.raises(shb.ShbError) as e:
        shb.bake_batched([b, fresh], 4, 1)
    assert "job 1" in str(e.value)
    fresh.close()


def test_autotune_keeps_results(jobs):
    b, spp, B, seed = jobs[3]
    ref = b.bake(spp, B, seed=seed)[0]
    old = b.tuning()
    try:
        t = b.autotune(spp, B, seed=seed)
        assert t["block_size"] in (64, 128, 256) and 2 <= t["bvh_width"] <= 8
        assert b.tuning() == t
        # the tuning changes launch shapes and the BVH, never the result
        assert np.array_equal(bits(b.bake(spp, B, seed=seed)[0]), bits(ref))
    finally:
        b.set_tuning(**old)


def test_tiled_bake_bit_identical_to_plain():
    """shb_bake_tiled: tiles of one surface as jobs of one batched run over
    the context's single BVH; every texel in a tile gets the plain bake's
    bits, texels outside every tile stay 0."""
    b = shb.Baker(0)
    try:
        b.upload_scene(scenes.scaled_synthetic_scene(6000, 48, 21))
        pts = b.rasterize(48, 48, 2)
        ref, rvis, rst = b.bake(8, 3, seed=77)
        grid = [(x, y, 24, 24) for y in range(0, 96, 24) for x in range(0, 96, 24)]
        sh, vis, st = b.bake_tiled(8, 3, grid, seed=77)
        assert np.array_equal(bits(sh), bits(ref)) and np.array_equal(bits(vis), bits(rvis))
        assert st["rays_extension"] == rst["rays_extension"]
        assert st["valid_texels"] == rst["valid_texels"]
        # uneven tiles covering part of the surface
        part = [(0, 0, 96, 10), (5, 30, 17, 41), (60, 20, 36, 76)]
        sh, vis, st = b.bake_tiled(8, 3, part, seed=77)
        inside = np.zeros((96, 96), bool)
        for x, y, w, h in part:
            inside[y:y + h, x:x + w] = True
        inside = inside.ravel()
        assert np.array_equal(bits(sh[inside]), bits(ref[inside]))
        assert (sh[~inside] == 0).all() and (vis[~inside] == 0).all()
        assert st["valid_texels"] == int(((pts["material_id"] >= 0) & inside).sum())
        # one tile: the whole surface
        sh, _, _ = b.bake_tiled(8, 3, [(0, 0, 96, 96)], seed=77)
        assert np.array_equal(bits(sh), bits(ref))
        # small batches: tiles that follow each other in the texel list
        # share segments, which now end and restart inside tiles
        old = b.tuning()
        try:
            for bp in (1024, 1736, 4096):
                b.set_tuning(batch_paths=bp)
                sh, vis, st = b.bake_tiled(8, 3, grid, seed=77)
                assert np.array_equal(bits(sh), bits(ref)) and np.array_equal(bits(vis), bits(rvis))
                assert st["rays_extension"] == rst["rays_extension"]
        finally:
            b.set_tuning(**old)
    finally:
        b.close()


def test_tiled_bake_rejects_bad_tiles():
    b = shb.Baker(0)
    try:
        b.upload_scene(scenes.box_room_scene(32, 32))
        b.rasterize(32, 32, 1, want=False)
        for tiles in ([(0, 0, 16, 16), (8, 8, 16, 16)],   # overlap
                      [(0, 0, 33, 4)],                     # outside
                      [(0, 0, 0, 4)],                      # empty
                      [(x, 0, 1, 1) for x in range(32)] * 3):  # more than 64
            with pytest.raises(shb.ShbError) as e:
                b.bake_tiled(4, 1, tiles)
            assert e.value.code in (-1, -5)
    finally:
        b.close()


def test_batched_jobs_with_different_light_sample_counts():
    """Jobs of one batch with 1, 3 and 2 light samples per NEE: per-pick
    arrays sized for the largest, each job bit-identical to its own plain
    bake (the one-sample job included)."""
    specs = [(scenes.scaled_synthetic_scene(4000, 48, 31), 48, 1),
             (scenes.scaled_synthetic_scene(6000, 64, 32), 64, 3),
             (scenes.box_room_scene(40, 40), 40, 2)]
    bakers = []
    try:
        for sc, res, _ in specs:
            b = shb.Baker(0)
            b.upload_scene(sc)
            b.rasterize(res, res, 1, want=False)
            bakers.append(b)
        nls = [n for _, _, n in specs]
        ref = [b.bake(8, 3, seed=5, num_light_samples=n)[:2] for b, n in zip(bakers, nls)]
        sh, vis, _ = shb.bake_batched(bakers, 8, 3, 5, num_light_samples=nls)
        for k in range(len(specs)):
            assert np.array_equal(bits(sh[k]), bits(ref[k][0])), k
            assert np.array_equal(bits(vis[k]), bits(ref[k][1])), k
        # the one-sample job alone is also the plain one-sample bake
        one = bakers[0].bake(8, 3, seed=5)[0]
        assert np.array_equal(bits(one), bits(ref[0][0]))
    finally:
        for b in bakers:
            b.close()


def emitter_floor(kinds, seed):
    """A 4 x 4 m floor under lights 0.6 m above it: kinds[i] triangles for
    area light i (a fan when more than one), 0 for a point light. Scenes
    with other kinds differ in light count, in each light's type and in
    where an area light's triangles start in the area-triangle table."""
    rng = np.random.default_rng(seed)
    b = scenes.SceneBuilder()
    mf = b.add_material((200, 200, 200, 255))
    b.add_quads([((-2.0, 0.0, -2.0), (0.0, 0.0, 4.0), (4.0, 0.0, 0.0))], mf, grid=[(4, 4)])
    for k in kinds:
        c = np.array([rng.uniform(-1.5, 1.5), 0.6, rng.uniform(-1.5, 1.5)])
        if k == 0:
            b.add_light(type=shb.LIGHT_POINT, position=tuple(c), intensity=2.0, color=(1, 1, 1))
            continue
        e = float(rng.uniform(1, 5))
        em = b.add_material((255, 255, 255, 255), emission=e)
        if k == 1:
            a = rng.uniform(0, 2 * np.pi)
            p = [c + 0.4 * np.array([np.cos(a + i * 2 * np.pi / 3), 0, np.sin(a + i * 2 * np.pi / 3)])
                 for i in range(3)]
            if np.cross(p[1] - p[0], p[2] - p[0])[1] > 0:  # face down
                p = [p[0], p[2], p[1]]
            g = b.add_triangle(p, em)
        else:
            g = b.add_emitter_fan(c, 0.4, k, em, rng)
        b.add_light(type=shb.LIGHT_AREA, geometry_index=g, material_index=em, intensity=e)
    return b.build(64, 64, 2)


@pytest.mark.parametrize("nls", [(1, 1, 1), (1, 3, 2)])
def test_batched_jobs_with_different_light_tables(nls):
    """Depth >= 1 waves hold paths of several jobs whose light records differ
    lane by lane (count, type, first area triangle): every job bit-identical
    to its plain bake, with the area lights lighting the floor."""
    # light counts 3, 5, 3 (the pair loop's odd tail at different lights);
    # light 1 is a one-triangle area light starting at area triangle 4 in
    # the second scene and 3 in the third
    kinds = [[1, 3, 0], [4, 1, 1, 0, 1], [3, 1, 5]]
    bakers = []
    try:
        for i, kd in enumerate(kinds):
            b = shb.Baker(0)
            b.upload_scene(emitter_floor(kd, 40 + i))
            b.rasterize(64, 64, 1, want=False)
            bakers.append(b)
        ref = [b.bake(8, 2, seed=11, num_light_samples=n)[:2] for b, n in zip(bakers, nls)]
        sh, vis, _ = shb.bake_batched(bakers, 8, 2, 11, num_light_samples=list(nls))
        for k in range(len(kinds)):
            assert (np.abs(ref[k][0]).max(axis=1) > 0).sum() > 500, k
            assert np.array_equal(bits(sh[k]), bits(ref[k][0])), k
            assert np.array_equal(bits(vis[k]), bits(ref[k][1])), k
    finally:
        for b in bakers:
            b.close()
