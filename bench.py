#!/usr/bin/env python3
"""Benchmark of the SH-lightmap bake on one MI355X.

Workload (default): the Sponza-scale synthetic room (about 262k triangles,
69 lights, 128 albedo materials), 1024x1024 RGB SH9 lightmap, supersample
scale 3 (3072x3072 traced texels), 128 spp, 3 bounces, 2 dilation passes.
A step = one full bake of that lightmap (BakeSHLightMap + downsample +
dilation, inputs resident in HBM). `value` = Mrays/s over the timed steps
(extension + shadow ray queries / wall seconds).

Also measured in the same run: the traversal kernels' HIP-event time and
algorithmic bytes (roofline), and the CPU restatement of the reference
bake (oracle, all host cores) on a row-stratified sample of the same
workload.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sh-baker_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
NODE_BYTES, TRI_BYTES, RAY_BYTES, HIT_BYTES = 80, 48, 32, 16


def log(*a):
    print(*a, file=sys.stderr, flush=True)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_k_trace_synthetic_room_1024_ss3_spp128.md")


VALU_FILE = os.path.join(ROOT, "profiles",
                         "valu_issue_ceiling_probe_and_synthetic_room_1024_ss3_spp128.md")


def read_profile_values(path, prefix):
    """`<prefix> <key>: <number>` lines of a committed profile summary."""
    out = {}
    if os.path.exists(path):
        for line in open(path):
            if line.startswith(prefix + " "):
                k, v = line[len(prefix) + 1:].split(":")
                out[k.strip()] = float(v)
    return out


def read_traffic():
    """HBM bytes per launch of each traversal kernel kind, measured with
    rocprofv3 FETCH_SIZE + WRITE_SIZE (scaled by a copy kernel of known size)
    and recorded as `traffic_bytes_per_launch <kind>: <bytes>` lines."""
    return {k: int(v) for k, v in read_profile_values(TRAFFIC_FILE,
                                                      "traffic_bytes_per_launch").items()}


def copy_probe(mib):
    """A device-to-device copy of known size (the calibration kernel for the
    FETCH_SIZE / WRITE_SIZE counters): copies `mib` MiB three times."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    c = torch.empty_like(a)
    for _ in range(3):
        c.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_each_way": n * 4, "copies": 3}), flush=True)


VALU_PROBE_KINDS = ("fp32_fma", "int_add_rotate_xor", "mixed_4fma_6int")


def probe_lib():
    """libshb_probe.so: the VALU issue probe and a CU count, on the same HIP
    runtime as libshb_hip.so (torch's CUDA init reported no GPUs once
    libshb_hip.so had opened the device)."""
    import ctypes
    lib = ctypes.CDLL(os.path.join(ROOT, "sh-baker_amd", "shb", "libshb_probe.so"))
    lib.shb_probe_cus.restype = ctypes.c_int
    lib.shb_probe_cus.argtypes = [ctypes.POINTER(ctypes.c_int)]
    return lib


def device_cus():
    import ctypes
    n = ctypes.c_int(0)
    if probe_lib().shb_probe_cus(ctypes.byref(n)) != 0 or n.value <= 0:
        raise RuntimeError("shb_probe_cus failed")
    return n.value


def valu_probe(iters=20000, blocks_per_cu=8, reps=3):
    """VALU issue ceiling (sh-baker_amd/csrc/probe/valu_probe.hip): wave64
    VALU instructions issued per SIMD per shader clock for independent
    chains of FP32 fma, of integer add/rotate/xor, and of both, with 8 waves
    per SIMD. The traversal and shading kernels are compared with it using
    their SQ_INSTS_VALU counts."""
    import ctypes
    lib = probe_lib()
    f = lib.shb_probe_valu
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int] * 5 + [ctypes.POINTER(ctypes.c_float),
                                       ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_double)]
    cus = device_cus()
    simds = 4 * cus
    res = []
    for kind, name in enumerate(VALU_PROBE_KINDS):
        ms, wi, mhz = ctypes.c_float(), ctypes.c_double(), ctypes.c_double()
        rc = f(kind, cus, blocks_per_cu, iters, reps, ctypes.byref(ms), ctypes.byref(wi),
               ctypes.byref(mhz))
        if rc != 0:
            raise RuntimeError(f"shb_probe_valu({name}) failed: {rc}")
        rate = wi.value / (ms.value * 1e-3) / simds
        res.append({"kind": name, "ms": round(ms.value, 4), "wave_instr": wi.value,
                    "shader_mhz": round(mhz.value, 1),
                    "valu_per_simd_per_clk": round(rate / (mhz.value * 1e6), 4),
                    "valu_per_simd_per_clk_at_2400mhz": round(rate / 2.4e9, 4)})
    print(json.dumps({"valu_probe": res, "cus": cus, "waves_per_simd": blocks_per_cu,
                      "iters": iters}), flush=True)


C3_SIZES = (64, 128, 256, 512, 1024, 2048)
C3_SPP = (32, 64, 128, 256)


def c3_jobs(n_jobs=16, seed=0x5EED0003):
    """The 16-job mixed batch: seeded synthetic rooms of 1k-500k triangles
    (log-uniform), lightmaps of 64^2..2048^2, 32..256 spp, 1..5 bounces,
    supersample scale 1."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_jobs):
        tris = int(round(math.exp(rng.uniform(math.log(1e3), math.log(5e5)))))
        res = int(rng.choice(C3_SIZES))
        spp = int(rng.choice(C3_SPP))
        bounces = int(rng.integers(1, 6))
        out.append(dict(tris=tris, res=res, spp=spp, bounces=bounces, seed=int(seed + k)))
    return out


def small_jobs(n_jobs=48, seed=0x5EED0013):
    """Many small lightmaps: 64^2..256^2, 1k-20k triangles, 16..64 spp,
    1..3 bounces (the regime a batched launch is for)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_jobs):
        out.append(dict(tris=int(round(math.exp(rng.uniform(math.log(1e3), math.log(2e4))))),
                        res=int(rng.choice((64, 128, 256))), spp=int(rng.choice((16, 32, 64))),
                        bounces=int(rng.integers(1, 4)), seed=int(seed + k)))
    return out


def bench_batched(shb, scenes, specs, reps=3):
    """Batched vs plain on a set of jobs, both with results left in device
    memory, after autotuning the largest job's shape for all jobs."""
    bakers = []
    t0 = time.time()
    for sp in specs:
        s = scenes.scaled_synthetic_scene(sp["tris"], sp["res"], sp["seed"])
        b = shb.Baker(0)
        b.upload_scene(s)
        b.rasterize(sp["res"], sp["res"], 1, want=False)
        sp["valid_texels"] = int(b.validity_mask().sum())
        sp["tris"] = s.num_triangles
        bakers.append(b)
    log(f"batched: {len(specs)} scenes built and rasterized in {time.time() - t0:.1f}s")
    big = max(range(len(specs)), key=lambda k: specs[k]["valid_texels"] * specs[k]["spp"])
    tuned = bakers[big].autotune(specs[big]["spp"], specs[big]["bounces"])
    for b in bakers:
        b.set_tuning(**tuned)
    spp = [sp["spp"] for sp in specs]
    bnc = [sp["bounces"] for sp in specs]
    seeds = [sp["seed"] for sp in specs]

    per_job = [[] for _ in specs]

    def plain_once():
        rays = 0
        for k, (b, sp) in enumerate(zip(bakers, specs)):
            ts = time.perf_counter()
            _, _, st = b.bake(sp["spp"], sp["bounces"], seed=sp["seed"], want=False)
            per_job[k].append(time.perf_counter() - ts)
            rays += st["rays_extension"] + st["rays_shadow"]
        return rays

    def batched_once():
        _, _, st = shb.bake_batched(bakers, spp, bnc, seeds, want=False)
        return st["rays_extension"] + st["rays_shadow"]

    plain_once()
    batched_once()
    tp, tb = [], []
    for _ in range(reps):
        ts = time.perf_counter()
        rays_p = plain_once()
        tp.append(time.perf_counter() - ts)
        ts = time.perf_counter()
        rays_b = batched_once()
        tb.append(time.perf_counter() - ts)
    tp, tb = float(np.median(tp)), float(np.median(tb))
    for b in bakers:
        b.close()
    job_s = [float(np.median(v)) for v in per_job]
    for sp, s in zip(specs, job_s):
        sp["plain_seconds"] = round(s, 4)
    # batching can only recover time outside the jobs that fill the GPU on
    # their own: the bound if every job but the two largest cost nothing
    top2 = sum(sorted(job_s)[-2:])
    res = {"jobs": len(specs), "plain_seconds": round(tp, 4), "batched_seconds": round(tb, 4),
           "speedup": round(tp / tb, 3),
           "speedup_bound_if_all_but_two_largest_free": round(tp / top2, 3) if len(specs) > 2
           else None, "rays": rays_b, "rays_equal": rays_b == rays_p,
           "batched_mrays_s": round(rays_b / tb / 1e6, 1), "tuning": tuned,
           "job_specs": [{k: sp[k] for k in ("tris", "res", "spp", "bounces", "valid_texels",
                                             "plain_seconds")}
                         for sp in specs]}
    log(f"batched: {json.dumps({k: v for k, v in res.items() if k != 'job_specs'})}")
    return res


def bench_env(shb, w=8192, h=4096, reps=10, cpu=True):
    """Environment SH of a w x h RGB float LatLong map (HBM-bound reduction):
    the kernels' time from HIP events with the map resident, against one
    read of the map (12 B per texel)."""
    rng = np.random.default_rng(9)
    th = ((np.arange(h) + 0.5) * np.pi / h).astype(np.float32)[:, None]
    ph = ((np.arange(w) + 0.5) * 2 * np.pi / w).astype(np.float32)[None, :]
    sky = (1.0 + np.cos(th) * (1.5 + 0.5 * np.sin(3 * ph))).astype(np.float32)
    px = np.empty((h, w, 3), np.float32)
    for c, k in enumerate((0.6, 0.8, 1.0)):
        px[..., c] = k * sky
    px *= rng.uniform(0.9, 1.1, (h, w, 1)).astype(np.float32)
    b = shb.Baker(0)
    sh, _ = b.env_sh_latlong(px)
    secs = []
    for _ in range(reps):
        sh2, s = b.env_sh_latlong(None, shape=(h, w, 3))
        secs.append(s)
    b.close()
    sec = float(np.median(secs))
    nbytes = w * h * 3 * 4
    gbs = nbytes / sec / 1e9
    res = {"map": f"{w}x{h} RGB f32", "alg_bytes_per_call": nbytes,
           "kernel_seconds": round(sec, 6), "achieved_GBs": round(gbs, 1),
           "roofline_frac": round(gbs / HBM_PEAK_GBS, 4), "deterministic": bool(np.array_equal(sh, sh2)),
           "kernels": "k_env_latlong + k_env_finish"}
    if cpu:
        from oracle import oracle as orc
        ts = time.perf_counter()
        ref, s64 = orc.env_latlong(px)
        res["cpu_seconds_serial_float"] = round(time.perf_counter() - ts, 3)
        res["max_abs_err_gpu_vs_fp64"] = float(np.abs(sh - s64).max())
        res["max_abs_err_serial_float_vs_fp64"] = float(np.abs(ref - s64).max())
    log(f"environment SH: {json.dumps(res)}")
    return res


def bench_stress(shb, scenes, spp=512):
    """The large stress configuration (4.2M triangles, 256 lights, 4096^2
    lightmap, supersample 2, 5 bounces, 4 dilation passes) at `spp` samples
    per texel instead of 512: the same scene, surface and path depth, 512 /
    spp times fewer independent samples. Reports throughput and the bake
    time scaled to 512 spp."""
    t0 = time.time()
    s = scenes.stress_scene(4096)
    gen_s = time.time() - t0
    log(f"stress: {s.num_triangles} tris, {len(s.lights)} lights ({gen_s:.1f}s)")
    b = shb.Baker(0)
    t0 = time.perf_counter()
    b.upload_scene(s)
    up_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    b.rasterize(4096, 4096, 2, want=False)
    ra_s = time.perf_counter() - t0
    valid = int(b.validity_mask().sum())
    log(f"stress: upload + BVH {up_s:.2f}s, raster 8192^2 {ra_s:.2f}s, {valid} valid texels")
    t0 = time.perf_counter()
    _, _, _, st = b.bake_lightmap(spp, 5, 4, seed=12345, to_host=False)
    bake_s = time.perf_counter() - t0
    # the configuration's batched form: 64 tile jobs (8 x 8 tiles of 1024^2
    # supersampled texels) in one batched run, against the plain raw bake
    # (both leave the result on the device)
    tiles = [(x, y, 1024, 1024) for y in range(0, 8192, 1024) for x in range(0, 8192, 1024)]
    t0 = time.perf_counter()
    _, _, pst = b.bake(spp, 5, seed=12345, want=False)
    plain_raw_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, _, tst = b.bake_tiled(spp, 5, tiles, seed=12345, want=False)
    tiled_s = time.perf_counter() - t0
    b.close()
    rays = st["rays_extension"] + st["rays_shadow"]
    res = {"config": "stress: 4.2M-triangle room, 256 lights, 4096^2 ss2, 5 bounces, dilation 4",
           "tris": s.num_triangles, "lights": len(s.lights), "valid_texels": valid,
           "spp_run": spp, "spp_config": 512, "rays": rays,
           "mrays_s": round(rays / bake_s / 1e6, 1), "bake_seconds": round(bake_s, 3),
           "bake_seconds_scaled_to_512spp": round(bake_s * 512 / spp, 1),
           "upload_bvh_seconds": round(up_s, 2), "raster_seconds": round(ra_s, 3),
           "scene_generation_seconds": round(gen_s, 1),
           "tiled_64_jobs": {"plain_bake_seconds": round(plain_raw_s, 3),
                             "tiled_bake_seconds": round(tiled_s, 3),
                             "speedup": round(plain_raw_s / tiled_s, 3),
                             "rays_equal": pst["rays_extension"] + pst["rays_shadow"] ==
                             tst["rays_extension"] + tst["rays_shadow"],
                             "mrays_s": round((tst["rays_extension"] + tst["rays_shadow"]) /
                                              tiled_s / 1e6, 1)}}
    log(f"stress: {json.dumps(res)}")
    return res


def bench_adjoint(shb, scenes, res=512, spp=64, bounces=3, iters=50, reps=3):
    """Differentiation form: the 262k-triangle room at res^2 (ss 1), spp,
    bounces; adjoint vs forward time, gradient vs central finite differences
    under common random numbers, and an Adam fit of 69x3 light and 192x3
    albedo scales to a target baked with seeded lognormal(0, 0.3) factors."""
    from shb import fit
    s = scenes.synthetic_room_scene(res, res)
    b = shb.Baker(0)
    b.upload_scene(s)
    b.rasterize(res, res, 1, want=False)
    nl, nm = len(s.lights), len(s.materials)
    seed = 4242
    true_l, true_a = fit.lognormal_factors(nl, nm, 0.3, seed=7)
    b.set_params(true_l, true_a)
    target = b.bake(spp, bounces, seed=seed)[0]
    ones_l, ones_a = np.ones((nl, 3), np.float32), np.ones((nm, 3), np.float32)
    b.set_params(ones_l, ones_a)
    sh = b.bake(spp, bounces, seed=seed)[0]
    g = (sh.astype(np.float64) - target).astype(np.float32) / sh.shape[0]
    b.bake_adjoint(spp, bounces, g, seed=seed)  # warm-up
    tf, ta = [], []
    for _ in range(reps):
        ts = time.perf_counter()
        b.bake(spp, bounces, seed=seed, want=False)
        tf.append(time.perf_counter() - ts)
        ts = time.perf_counter()
        dl, dm, _ = b.bake_adjoint(spp, bounces, g, seed=seed)
        ta.append(time.perf_counter() - ts)
    tf, ta = float(np.median(tf)), float(np.median(ta))

    def loss(ls, al):
        b.set_params(ls, al)
        d = b.bake(spp, bounces, seed=seed)[0].astype(np.float64) - target
        return 0.5 * float((d * d).sum()) / sh.shape[0]

    fd_checks = []
    for kind, grad in (("light", dl), ("albedo", dm)):
        for flat in np.argsort(-np.abs(grad).ravel())[:3]:
            i, c = int(flat // 3), int(flat % 3)
            for h in (1e-3, 1e-2):
                lp, lm_, ap, am = ones_l.copy(), ones_l.copy(), ones_a.copy(), ones_a.copy()
                if kind == "light":
                    lp[i, c], lm_[i, c] = 1 + h, 1 - h
                    fdv = (loss(lp, ones_a) - loss(lm_, ones_a)) / (2 * h)
                else:
                    ap[i, c], am[i, c] = 1 + h, 1 - h
                    fdv = (loss(ones_l, ap) - loss(ones_l, am)) / (2 * h)
                fd_checks.append({"param": f"{kind}[{i}][{c}]", "h": h, "adjoint": float(grad[i, c]),
                                  "fd": fdv, "rel_err": abs(float(grad[i, c]) - fdv) / max(abs(fdv), 1e-30)})
    b.set_params(ones_l, ones_a)
    t0 = time.perf_counter()
    ls, al, losses = fit.fit(b, target, spp, bounces, iters=iters, lr=0.05, seed=seed)
    t_fit = time.perf_counter() - t0
    err_l = float(np.abs(np.log(ls / true_l)).mean())
    err_a = float(np.abs(np.log(al / true_a)).mean())
    b.close()
    res_d = {"lightmap": f"{res}x{res}", "spp": spp, "bounces": bounces, "params": 3 * (nl + nm),
             "forward_seconds": round(tf, 4), "adjoint_seconds": round(ta, 4),
             "adjoint_forward_ratio": round(ta / tf, 3),
             "fd_max_rel_err_h1e-3": max(f["rel_err"] for f in fd_checks if f["h"] == 1e-3),
             "fd_max_rel_err_h1e-2": max(f["rel_err"] for f in fd_checks if f["h"] == 1e-2),
             "fd_checks": fd_checks,
             "adam_iters": iters, "loss_first": losses[0], "loss_last": losses[-1],
             "fit_seconds": round(t_fit, 2),
             "mean_abs_log_error_lights": round(err_l, 4),
             "mean_abs_log_error_albedos": round(err_a, 4)}
    log(f"adjoint: {json.dumps({k: v for k, v in res_d.items() if k != 'fd_checks'})}")
    return res_d


def cpu_share():
    """(threads, description) of the host cores this job may use: the CPU
    share the environment declares for it (OMP_NUM_THREADS, as a one-GPU
    job's share of a shared host is set), else every core of the process's
    affinity mask; nproc and the affinity are reported beside it."""
    aff = len(os.sched_getaffinity(0))
    declared = os.environ.get("OMP_NUM_THREADS")
    threads = int(declared) if declared and declared.isdigit() and int(declared) > 0 else aff
    threads = min(threads, aff)
    return threads, {"nproc": os.cpu_count(), "affinity_cores": aff,
                     "omp_num_threads": declared, "threads_used": threads}


def per_texel_agreement(g_sh, r_sh, r_var, g_var, pts, width, stride):
    """z = (GPU - reference stream) / sqrt(var_ref + var_gpu) per texel and
    coefficient on the valid texels of every `stride`-th row, each var the
    stream's own sample variance of its per-texel mean. Both are needed: a
    texel lit by rare, large terms (an emitter seen by few primary rays) can
    have a sample variance of 0 in one stream and not in the other.
    Coefficients with zero variance in both (zero or deterministic terms)
    are left out and counted."""
    rows = (np.arange(pts.size) // width) % stride == 0
    v = (pts["material_id"] >= 0) & rows
    var = r_var[v].astype(np.float64) + g_var[v].astype(np.float64)
    live = var > 0
    z = (g_sh[v].astype(np.float64) - r_sh[v])[live] / np.sqrt(var[live])
    az = np.abs(z)
    dead = ~live
    gd, rd = g_sh[v][dead], r_sh[v][dead]
    uneq = ~np.isclose(gd, rd, rtol=1e-5, atol=0)
    return {"texels": int(v.sum()), "coefficients": int(live.sum()),
            "coefficients_zero_variance": int(dead.sum()),
            "zero_variance_unequal": int(uneq.sum()),
            "zero_variance_max_abs_diff": float(np.abs(gd - rd).max()) if gd.size else 0.0,
            "texels_black_in_both": int(((np.abs(g_sh[v]).max(1) == 0) &
                                         (np.abs(r_sh[v]).max(1) == 0)).sum()),
            "share_within_3_sigma": round(float((az <= 3.0).mean()), 5),
            "mean_abs_z": round(float(az.mean()), 4),
            "share_above_4_5_sigma": round(float((az > 4.5).mean()), 6),
            "gaussian_expectation": {"share_within_3_sigma": 0.9973, "mean_abs_z": 0.7979},
            "reference": "oracle, std::mt19937 reference stream, same texels and spp"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--res", type=int, default=1024)
    ap.add_argument("--ss", type=int, default=3)
    ap.add_argument("--spp", type=int, default=128)
    ap.add_argument("--bounces", type=int, default=3)
    ap.add_argument("--dilation", type=int, default=2)
    ap.add_argument("--tris", type=int, default=262144)
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target CPU-baseline sample duration (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-count", action="store_true")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the 16-job batched-vs-plain measurement")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the adjoint / fitting measurement")
    ap.add_argument("--no-env", action="store_true",
                    help="skip the environment-SH measurement")
    ap.add_argument("--env-only", action="store_true",
                    help="only the environment-SH measurement (its own JSON line)")
    ap.add_argument("--no-stress", action="store_true",
                    help="skip the 4.2M-triangle / 256-light stress bake")
    ap.add_argument("--stress", action="store_true",
                    help="only the 4.2M-triangle / 256-light stress bake (its own JSON line)")
    ap.add_argument("--stress-spp", type=int, default=512)
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only the adjoint / fitting measurement (its own JSON line)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a device copy of this size (counter calibration)")
    ap.add_argument("--valu-probe", action="store_true",
                    help="only measure the VALU issue ceiling (its own JSON line)")
    args = ap.parse_args()
    if args.valu_probe:
        valu_probe()
        return
    if args.copy_probe_mib:
        copy_probe(args.copy_probe_mib)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); "
              "this build has no multi-GPU path", file=sys.stderr)
        sys.exit(2)

    import shb
    from shb import scenes

    if args.env_only:
        print(json.dumps({"environment_sh": bench_env(shb, cpu=args.cpu_seconds > 0)}), flush=True)
        return
    if args.stress:
        print(json.dumps({"stress": bench_stress(shb, scenes, args.stress_spp)}), flush=True)
        return
    if args.adjoint_only:
        print(json.dumps({"adjoint_512": bench_adjoint(shb, scenes)}), flush=True)
        return

    t0 = time.time()
    if args.tris == 262144:
        scene = scenes.synthetic_room_scene(args.res, args.res)
    else:
        scene = scenes.scaled_synthetic_scene(args.tris, args.res, scenes.SYNTH_SEED)
    log(f"scene: {scene.num_triangles} tris, {len(scene.lights)} lights, "
        f"{len(scene.materials)} materials ({time.time() - t0:.1f}s)")
    b = shb.Baker(0)
    t0 = time.perf_counter()
    b.upload_scene(scene)
    upload_s = time.perf_counter() - t0
    log(f"upload + BVH build {upload_s:.3f}s")
    pts = b.rasterize(args.res, args.res, args.ss)
    valid = int((pts["material_id"] >= 0).sum())
    log(f"raster {args.res}x{args.res} ss{args.ss}: {valid} valid texels "
        f"({valid / pts.size:.3f})")

    def step(to_host=False):
        # the result stays in HBM (read back with read_lightmap); the PCIe
        # transfer of the host-output form is timed separately below
        return b.bake_lightmap(args.spp, args.bounces, args.dilation, seed=12345,
                               to_host=to_host)

    for i in range(args.warmup):
        _, _, _, st = step()
        log(f"warmup {i}: {st['seconds_total']:.3f}s rays {st['rays_extension'] + st['rays_shadow']}")
    # every entry point returns after its stream has drained, so the timed
    # region is bracketed by device synchronisation on both sides
    times, rays, trace_s, trace_sh_s, launches = [], 0, 0.0, 0.0, [0, 0]
    rays_split = {"extension": 0, "shadow_queries": 0, "shadow_traced": 0,
                  "shadow_resolved_by_emitter_test": 0}
    t_start = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        _, _, _, st = step()
        times.append(time.perf_counter() - ts)
        rays += st["rays_extension"] + st["rays_shadow"]
        rays_split["extension"] += st["rays_extension"]
        rays_split["shadow_queries"] += st["rays_shadow"]
        rays_split["shadow_traced"] += st["rays_shadow"] - st["rays_shadow_resolved"]
        rays_split["shadow_resolved_by_emitter_test"] += st["rays_shadow_resolved"]
        trace_s += st["seconds_trace"]
        trace_sh_s += st["seconds_trace_shadow"]
        launches[0] += st["launches_extension"]
        launches[1] += st["launches_shadow"]
        log(f"step {i}: {times[-1]:.3f}s  device {st['seconds_device']:.3f}s  "
            f"trace {st['seconds_trace']:.3f}s (shadow {st['seconds_trace_shadow']:.3f}s)  "
            f"ext {st['rays_extension']}  shadow {st['rays_shadow']}")
    elapsed = time.perf_counter() - t_start
    mrays = rays / elapsed / 1e6
    ms_per_step = elapsed / args.steps * 1e3
    bake_s = float(np.median(times))
    ts = time.perf_counter()
    step(to_host=True)
    host_s = time.perf_counter() - ts
    log(f"with outputs copied to host memory: {host_s:.3f}s")

    roofline, kernels, valu_roofline = None, None, None
    if not args.no_count:
        _, _, stc = b.bake(args.spp, args.bounces, seed=12345, flags=shb.BAKE_COUNT_VISITS)
        # the shadow traversal's rays: the queries minus those its own
        # emitter's triangle test answered in the shading kernels
        n_sh = stc["rays_shadow"] - stc["rays_shadow_resolved"]
        n_ext = stc["rays_extension"]
        nodes_sh, tris_sh = stc["node_visits_shadow"], stc["tri_tests_shadow"]
        nodes_ext, tris_ext = stc["node_visits"] - nodes_sh, stc["tri_tests"] - tris_sh
        traffic = read_traffic()
        kernels = {}
        for name, n, nodes, tris, secs, nl in (
                ("k_trace<closest, HitEpilogue> (extension rays)", n_ext, nodes_ext, tris_ext,
                 trace_s - trace_sh_s, launches[0]),
                ("k_trace<any, BakeShadowEpilogue> (shadow rays)", n_sh, nodes_sh, tris_sh,
                 trace_sh_s, launches[1])):
            alg = n * (RAY_BYTES + HIT_BYTES) + nodes * NODE_BYTES + tris * TRI_BYTES  # per bake
            per_launch = alg / max(nl / args.steps, 1)
            avg_launch_s = secs / max(nl, 1)
            achieved = per_launch / avg_launch_s / 1e9 if avg_launch_s > 0 else 0.0
            kernels[name] = {
                "kind": "shadow" if "shadow" in name else "extension",
                "rays_per_bake": n, "nodes_per_ray": round(nodes / max(n, 1), 2),
                "tris_per_ray": round(tris / max(n, 1), 2),
                "alg_bytes_per_launch": int(per_launch),
                "avg_launch_ms": round(avg_launch_s * 1e3, 4),
                "launches_per_bake": nl // args.steps,
                "achieved_GBs": round(achieved, 1), "seconds_per_bake": round(secs / args.steps, 4)}
        dom = max(kernels, key=lambda k: kernels[k]["seconds_per_bake"])
        kd = kernels[dom]
        roofline = {"bound": "hbm", "achieved": kd["achieved_GBs"], "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(kd["achieved_GBs"] / HBM_PEAK_GBS, 4),
                    "traffic": traffic.get(kd["kind"]),
                    "kernel": dom}
        # VALU issue roofline of the same kernel: SQ_INSTS_VALU per launch
        # (counted on this workload) over the live launch time, against the
        # probe's mixed-stream ceiling (bench.py --valu-probe)
        vi = read_profile_values(VALU_FILE, "valu_instr_per_launch")
        vc = read_profile_values(VALU_FILE, "valu_issue_ceiling")
        default_workload = (args.res, args.ss, args.spp, args.bounces, args.tris) == (
            1024, 3, 128, 3, 262144)
        if default_workload and kd["kind"] in vi and "mixed_4fma_6int" in vc:
            simds = 4 * device_cus()
            rate = vi[kd["kind"]] / (kd["avg_launch_ms"] * 1e-3 * simds * 2.4e9)
            ceil = vc["mixed_4fma_6int"]
            valu_roofline = {"bound": "valu-issue", "achieved": round(rate, 4), "peak": ceil,
                             "unit": "wave64 VALU instructions / SIMD / clock at 2400 MHz",
                             "frac": round(rate / ceil, 4),
                             "valu_instr_per_launch": vi[kd["kind"]],
                             "peak_source": "mixed FP32+integer probe stream (measured); "
                                            f"FP32-only {vc.get('fp32_fma')}, integer-only "
                                            f"{vc.get('int_add_rotate_xor')}",
                             "kernel": dom}
        log(f"counted bake: {json.dumps(kernels)}")

    adjoint = None if args.no_adjoint else bench_adjoint(shb, scenes)
    env = None if args.no_env else bench_env(shb, cpu=args.cpu_seconds > 0)
    stress = None if args.no_stress else bench_stress(shb, scenes, args.stress_spp)

    batched = small = None
    if not args.no_batched:
        batched = bench_batched(shb, scenes, c3_jobs())
        small = bench_batched(shb, scenes, small_jobs())

    cpu = None
    agreement = None
    if args.cpu_seconds > 0:
        from oracle import oracle as orc
        threads, share = cpu_share()
        threads = args.cpu_threads or threads
        o = orc.Oracle(scene)
        W = args.res * args.ss
        stride = 1024
        while True:
            r_sh, _, r_var, cst = o.bake(pts, W, W, args.spp, args.bounces,
                                         mode=orc.MODE_REF, threads=threads, row_stride=stride)
            if cst["seconds"] > args.cpu_seconds / 8 or stride == 1:
                break
            stride = max(1, stride // 4)
        # scale to about cpu_seconds
        want = max(1, int(stride * cst["seconds"] / args.cpu_seconds))
        if want < stride:
            stride = want
            r_sh, _, r_var, cst = o.bake(pts, W, W, args.spp, args.bounces,
                                         mode=orc.MODE_REF, threads=threads, row_stride=stride)
        crays = cst["rays_extension"] + cst["rays_shadow"]
        cpu_mrays = crays / cst["seconds"] / 1e6
        frac_texels = cst["valid_texels"] / max(valid, 1)
        cpu = {"value": round(cpu_mrays, 3), "unit": "Mrays/s", "cores": threads,
               "kind": "port",
               "sample": f"every {stride}th texel row of the same {W}x{W} bake "
                         f"({cst['valid_texels']} texels, {frac_texels:.4f} of the valid ones), "
                         f"{cst['seconds']:.1f}s; std::mt19937 reference stream, host BVH2 "
                         f"(Embree unavailable)",
               "bake_seconds_extrapolated": round(cst["seconds"] / max(frac_texels, 1e-12), 1),
               "host": share}
        log(f"cpu baseline: {cpu}")
        # per-texel agreement of the GPU bake with the reference-stream bake
        # on the texels the CPU leg baked (its last, timed run); the GPU's
        # per-texel variance comes from the oracle's GPU-stream mode, which
        # draws the GPU's samples
        g_sh = b.bake(args.spp, args.bounces, seed=12345)[0]
        ts = time.perf_counter()
        _, _, g_var, _ = o.bake(pts, W, W, args.spp, args.bounces, mode=orc.MODE_GPU,
                                seed=12345, threads=threads, row_stride=stride)
        agreement = per_texel_agreement(g_sh, r_sh, r_var, g_var, pts, W, stride)
        log(f"agreement ({time.perf_counter() - ts:.1f}s for the GPU-stream variances): "
            f"{agreement}")

    out = {
        "metric": "Mrays/s + bake seconds @1024^2, 128spp, 3 bounces; adjoint/forward time ratio",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 1),
        "bake_seconds": round(bake_s, 4),
        # the same rate over the rays the traversal kernels traced (shadow
        # queries answered by the emitter's triangle test left out)
        "mrays_s_traced_only": round((rays_split["extension"] + rays_split["shadow_traced"]) /
                                     elapsed / 1e6, 2),
        "bake_seconds_incl_bvh_build": round(bake_s + upload_s, 4),
        "bake_seconds_outputs_to_host": round(host_s, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": "synthetic-room-262k bake", "tris": scene.num_triangles,
                   "lights": len(scene.lights), "lightmap": f"{args.res}x{args.res}",
                   "supersample_scale": args.ss, "spp": args.spp, "bounces": args.bounces,
                   "dilation": args.dilation, "valid_texels": valid,
                   "rays_per_step": rays // args.steps,
                   "rays_per_step_by_kind": {k: v // args.steps for k, v in rays_split.items()}},
        "roofline": roofline,
        "roofline_kernels": kernels,
        "valu_roofline": valu_roofline,
        "adjoint_forward_ratio": adjoint["adjoint_forward_ratio"] if adjoint else None,
        "adjoint_512": adjoint,
        "batched_16_jobs": batched,
        "batched_48_small_jobs": small,
        "environment_sh": env,
        "stress_4m_tris_256_lights": stress,
        "cpu_baseline": cpu,
        "agreement": agreement,
    }
    if cpu:
        out["speedup_vs_cpu"] = round(mrays / cpu["value"], 1)
        out["speedup_bake_seconds_vs_cpu"] = round(cpu["bake_seconds_extrapolated"] / bake_s, 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
