"""Benchmark of the MI355X path tracer: Mrays/s on the icospheres workload.

Workload (one "step" = one full-frame render_tile pass): 200 displaced
icospheres = 1,024,000 triangles, Principled materials, 1 point + 1 sun + 4
area lights, 1024 x 1024 pixels, 64 samples per pixel, max depth 12, MIS.
The scene is packed and uploaded before the timed region; the output stays in
device memory (lucid_render_tile_device).  Mrays/s = (closest-hit + shadow
queries) / wall time of the timed steps.

Also reported:
* roofline of the dominant kernel (BVH closest-hit traversal) from one
  instrumented pass: algorithmic bytes per launch (ray in 32 B + hit out 16 B
  per query, 64 B per BVH node and 48 B per triangle fetched, counted on the
  device) over the HIP-event average launch duration, against 8 TB/s;
* cpu_baseline: the oracle's -O3 -ffast-math build of the reference
  algorithm on the host cores, on a bounded strip of the same frame.

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "blender-lucid-renderer_amd")
ORACLE = os.path.join(ROOT, "oracle")
for _p in (PKG, ORACLE):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "Mrays/s + time-to-relMSE 0.01 (guided vs plain), 1080p, 1×MI355X"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(scene, cam, W, H, spp, depth, seconds_target=15.0):
    """Times the oracle (fast build) on a strip of rows of the same frame."""
    import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
    o = oracle.OracleScene(scene, fast=True)
    o.set_camera(cam["pos"], cam["dir"], cam["up"], cam["fov"])
    # calibrate on a small strip, then size the strip for ~seconds_target
    rows, samples = 16, 2
    t0 = time.perf_counter()
    _, ex = o.render_tile("mis", 0, H // 2, W, rows, W, H, samples=samples, max_depth=depth, threads=threads)
    dt = time.perf_counter() - t0
    rays = int(ex["counts"].sum())
    scale = max(1.0, min(64.0, seconds_target / max(dt, 1e-3)))
    rows2 = int(min(H, rows * scale))
    y0 = max(0, H // 2 - rows2 // 2)
    t0 = time.perf_counter()
    _, ex = o.render_tile("mis", 0, y0, W, rows2, W, H, samples=samples, max_depth=depth, threads=threads)
    dt = time.perf_counter() - t0
    rays = int(ex["counts"].sum())
    return {"value": round(rays / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"{W}x{rows2} rows {y0}..{y0 + rows2 - 1} of the {W}x{H} frame, {samples} spp, depth {depth}, "
                      f"mis; {rays} queries in {dt:.1f} s (oracle -O3 -ffast-math, OpenMP)",
            "spp_full_frame_s_est": round(dt * (H / rows2) * (spp / samples), 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="icospheres")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--wave-paths", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import ctypes as C

    import numpy as np
    import torch

    import lucid_capi
    import workloads

    wl = workloads.WORKLOADS[args.workload]
    W, H, spp, depth, algo = wl["width"], wl["height"], wl["spp"], wl["max_depth"], wl["algorithm"]
    cam = wl["camera"]
    t0 = time.perf_counter()
    scene = wl["make"]()
    log(f"scene generated in {time.perf_counter() - t0:.1f} s")

    ctx = lucid_capi.Context(0)
    L = ctx.L
    t0 = time.perf_counter()
    sc = ctx.scene(scene)
    ntri = L.lucid_scene_num_triangles(sc)
    log(f"scene packed+uploaded: {ntri} triangles in {time.perf_counter() - t0:.1f} s")
    if args.wave_paths:
        ctx.check(L.lucid_set_wave_paths(ctx.h, args.wave_paths))
    camd = lucid_capi.camera(cam)
    prm = lucid_capi.params(0, 0, W, H, W, H, spp, 0, depth, algo)
    out = torch.empty(W * H * 4, dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream

    def step():
        ctx.check(L.lucid_render_tile_device(ctx.h, sc, C.byref(camd), C.byref(prm), C.c_void_p(out.data_ptr()),
                                             C.c_void_p(stream)))
        ctx.check(L.lucid_synchronize(ctx.h))
        return ctx.stats()

    for i in range(args.warmup):
        s = step()
        log(f"warmup {i}: {s['ms_render']:.1f} ms, {s['rays_closest'] + s['rays_shadow']} rays")
    torch.cuda.synchronize()
    rays = 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        s = step()
        rays += s["rays_closest"] + s["rays_shadow"]
        log(f"step {i}: {s['ms_render']:.1f} ms")
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    mrays = rays / dt / 1e6
    ms_step = dt / args.steps * 1e3
    img = out.view(H, W, 4)
    assert torch.isfinite(img).all().item(), "non-finite pixels"

    # instrumented pass: node / triangle fetch counts + per-kernel event timing
    ctx.check(L.lucid_set_profiling(ctx.h, lucid_capi.PROFILE_WORK | lucid_capi.PROFILE_KERNELS))
    p = step()
    ctx.check(L.lucid_set_profiling(ctx.h, 0))
    q = p["rays_closest"]
    bytes_closest = q * (32 + 16) + p["nodes_closest"] * 64 + p["tris_closest"] * 48
    n_launch = max(1, p["launches_intersect"])
    avg_ms = p["ms_intersect"] / n_launch
    per_launch = bytes_closest / n_launch
    achieved = per_launch / (avg_ms / 1e3) / 1e9
    log(f"intersect: {n_launch} launches, avg {avg_ms:.3f} ms, nodes/query {p['nodes_closest'] / max(q, 1):.1f}, "
        f"tris/query {p['tris_closest'] / max(q, 1):.1f}; shade {p['ms_shade']:.1f} ms, shadow {p['ms_shadow']:.1f} ms")

    result = {
        "metric": METRIC,
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (procedural scene, seeded)",
        "config": {"workload": f"{args.workload}: {ntri} triangles, {W}x{H}, {spp} spp, depth {depth}, {algo}, "
                               f"1 point + 1 sun + 4 area lights",
                   "rays_per_step": rays // max(args.steps, 1), "paths_per_step": W * H * spp,
                   "msamples_per_s": round(W * H * spp * args.steps / dt / 1e6, 2)},
        "roofline": {"bound": "hbm", "kernel": "k_intersect (closest-hit BVH traversal)",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_launch": int(per_launch), "avg_launch_ms": round(avg_ms, 4),
                     "launches": n_launch,
                     "per_query": {"nodes": round(p["nodes_closest"] / max(q, 1), 2),
                                   "tris": round(p["tris_closest"] / max(q, 1), 2),
                                   "node_bytes": 64, "tri_bytes": 48, "ray_bytes": 32, "hit_bytes": 16}},
    }
    if not args.no_cpu:
        log("cpu baseline (oracle, fast build) ...")
        result["cpu_baseline"] = cpu_baseline(scene, cam, W, H, spp, depth)
    print(json.dumps(result), flush=True)
    L.lucid_scene_release(sc)
    ctx.close()


if __name__ == "__main__":
    main()
