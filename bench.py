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
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); "
              "this build has no multi-GPU path", file=sys.stderr)
        sys.exit(2)

    import shb
    from shb import scenes

    t0 = time.time()
    if args.tris == 262144:
        scene = scenes.synthetic_room_scene(args.res, args.res)
    else:
        scene = scenes.scaled_synthetic_scene(args.tris, args.res, scenes.SYNTH_SEED)
    log(f"scene: {scene.num_triangles} tris, {len(scene.lights)} lights, "
        f"{len(scene.materials)} materials ({time.time() - t0:.1f}s)")
    b = shb.Baker(0)
    t0 = time.time()
    b.upload_scene(scene)
    log(f"upload + BVH build {time.time() - t0:.2f}s")
    pts = b.rasterize(args.res, args.res, args.ss)
    valid = int((pts["material_id"] >= 0).sum())
    log(f"raster {args.res}x{args.res} ss{args.ss}: {valid} valid texels "
        f"({valid / pts.size:.3f})")

    def step():
        return b.bake_lightmap(args.spp, args.bounces, args.dilation, seed=12345)

    for i in range(args.warmup):
        _, _, _, st = step()
        log(f"warmup {i}: {st['seconds_total']:.3f}s rays {st['rays_extension'] + st['rays_shadow']}")
    import torch  # noqa: F401  (the harness's timing convention: synchronize around the region)
    times, rays, trace_s, dev_s = [], 0, 0.0, 0.0
    t_start = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        _, _, _, st = step()
        times.append(time.perf_counter() - ts)
        rays += st["rays_extension"] + st["rays_shadow"]
        trace_s += st["seconds_trace"]
        dev_s += st["seconds_device"]
        log(f"step {i}: {times[-1]:.3f}s  device {st['seconds_device']:.3f}s  "
            f"trace {st['seconds_trace']:.3f}s  ext {st['rays_extension']}  "
            f"shadow {st['rays_shadow']}")
    elapsed = time.perf_counter() - t_start
    mrays = rays / elapsed / 1e6
    ms_per_step = elapsed / args.steps * 1e3

    roofline = None
    if not args.no_count:
        _, _, stc = b.bake(args.spp, args.bounces, seed=12345, flags=shb.BAKE_COUNT_VISITS)
        nrays = stc["rays_extension"] + stc["rays_shadow"]
        alg_bytes = (nrays * (RAY_BYTES + HIT_BYTES) + stc["node_visits"] * NODE_BYTES
                     + stc["tri_tests"] * TRI_BYTES)
        per_step_bytes = alg_bytes  # same work as one timed step
        achieved = per_step_bytes / (trace_s / args.steps) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                    "kernel": "k_trace (closest + any hit)",
                    "nodes_per_ray": round(stc["node_visits"] / nrays, 2),
                    "tris_per_ray": round(stc["tri_tests"] / nrays, 2)}
        log(f"counted bake: nodes/ray {stc['node_visits'] / nrays:.2f} "
            f"tris/ray {stc['tri_tests'] / nrays:.2f} alg bytes {alg_bytes / 1e9:.1f} GB")

    cpu = None
    if args.cpu_seconds > 0:
        from oracle import oracle as orc
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        o = orc.Oracle(scene)
        W = args.res * args.ss
        stride = 1024
        while True:
            _, _, _, cst = o.bake(pts, W, W, args.spp, args.bounces, mode=orc.MODE_REF,
                                  threads=threads, row_stride=stride)
            if cst["seconds"] > args.cpu_seconds / 8 or stride == 1:
                break
            stride = max(1, stride // 4)
        # scale to about cpu_seconds
        want = max(1, int(stride * cst["seconds"] / args.cpu_seconds))
        if want < stride:
            stride = want
            _, _, _, cst = o.bake(pts, W, W, args.spp, args.bounces, mode=orc.MODE_REF,
                                  threads=threads, row_stride=stride)
        crays = cst["rays_extension"] + cst["rays_shadow"]
        cpu_mrays = crays / cst["seconds"] / 1e6
        frac_texels = cst["valid_texels"] / max(valid, 1)
        cpu = {"value": round(cpu_mrays, 3), "unit": "Mrays/s", "cores": threads,
               "kind": "port",
               "sample": f"every {stride}th texel row of the same {W}x{W} bake "
                         f"({cst['valid_texels']} texels, {frac_texels:.4f} of the valid ones), "
                         f"{cst['seconds']:.1f}s; std::mt19937 reference stream, host BVH2 "
                         f"(Embree unavailable)",
               "bake_seconds_extrapolated": round(cst["seconds"] / max(frac_texels, 1e-12), 1)}
        log(f"cpu baseline: {cpu}")

    out = {
        "metric": "Mrays/s + bake seconds @1024^2, 128spp, 3 bounces; adjoint/forward time ratio",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 1),
        "bake_seconds": round(ms_per_step / 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": "synthetic-room-262k bake", "tris": scene.num_triangles,
                   "lights": len(scene.lights), "lightmap": f"{args.res}x{args.res}",
                   "supersample_scale": args.ss, "spp": args.spp, "bounces": args.bounces,
                   "dilation": args.dilation, "valid_texels": valid,
                   "rays_per_step": rays // args.steps},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["speedup_vs_cpu"] = round(mrays / cpu["value"], 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
