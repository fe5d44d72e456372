"""Benchmark of the MI355X path tracer on the metric's workload: Mrays/s and
time-to-relMSE 0.01 (guided vs plain) at 1080p.

Workload (default ``occluded_room``): a 295,648-triangle glossy room
(roughness U[0.1, 0.4]) lit only through a 5 cm slit by a hidden area light,
1920 x 1080, traceMIS at depth 12.  One "step" = one progressive pass of
16 samples per pixel over the full frame (lucid_render_tile_device; the scene
is packed and uploaded before the timed region, the output stays in device
memory).  Mrays/s = (closest-hit + shadow queries) / wall time of the timed
plain passes.

time-to-relMSE: plain MIS and guided MIS (online-trained BF16 MFMA guiding
network, training time included) each render progressive 16-spp passes for
a fixed time budget.  After every pass the running mean is compared with
the FP64-accumulated oracle image on a stratified subset of 16,384 pixels
(tests/golden/occluded_room_reference.npz, written by
tests/golden/make_occluded_room_reference.py):
relMSE = mean over pixels and channels of (x - r)^2 / (r^2 + 0.01), minus the
reference's own estimated relMSE (sample variance / spp), which the
independent noise of the reference adds on average.  The time at which the
corrected relMSE crosses 0.01 is reported as measured when a run gets there
within its budget, else extrapolated from relMSE * t over the last half of
the run (relMSE proportional to 1/t), and labelled as such.

Also reported: the roofline of the dominant kernel (BVH closest-hit
traversal) from one instrumented pass -- algorithmic bytes per launch (ray
32 B + hit 16 B per query, 64 B per BVH node and 48 B per triangle fetched,
counted on the device) over the HIP-event average launch duration, against
8 TB/s -- and cpu_baseline: the oracle's -O3 -ffast-math build of the same
algorithm on the host cores, on a bounded strip of the same frame.

sessions: 32 progressive sessions of different scenes and sizes (5.7M
pixels, 5M triangles in all), Mrays/s of batched 16-spp passes over the
library's four stream lanes against the same passes one session at a time.

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1] [--budget SECONDS]
       python bench.py --workload icospheres      (1M triangles, 1024^2, 64 spp: Mrays/s only)
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
REFERENCE = os.path.join(ROOT, "tests", "golden", "occluded_room_reference.npz")
# rocprofv3 FETCH_SIZE + WRITE_SIZE of the closest-hit traversal on the default workload,
# scaled by a 1 GiB device copy (profiles/ holds the commands and the table)
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_trace_refill_occluded_room_1080p_16spp.md")
TARGET = 0.01


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _host_info():
    """CPU model and core counts of the host the baseline runs on."""
    import subprocess

    info = {"os_cpu_count": os.cpu_count()}
    try:
        info["nproc"] = int(subprocess.run(["nproc"], capture_output=True, text=True, timeout=10).stdout.strip())
    except Exception:
        info["nproc"] = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                info["model"] = line.split(":", 1)[1].strip()
            elif line.startswith("Socket(s):"):
                info["sockets"] = line.split(":", 1)[1].strip()
    except Exception:
        pass
    return info


def _cpu_rate(o, W, H, depth, threads, seconds_target, algo="mis"):
    """Mrays/s of the oracle on a strip of rows of the frame sized to about
    `seconds_target` seconds; returns (Mrays/s, sample text)."""
    rows, samples = 8, 2
    t0 = time.perf_counter()
    o.render_tile(algo, 0, H // 2, W, rows, W, H, samples=samples, max_depth=depth, threads=threads)
    dt = time.perf_counter() - t0
    total = seconds_target / max(dt / (rows * samples), 1e-9)  # row-samples that fit the target
    rows2 = int(max(1, min(H, total / samples)))
    samples = int(max(samples, total / rows2))  # a full frame that is quick gets more samples
    y0 = max(0, H // 2 - rows2 // 2)
    t0 = time.perf_counter()
    _, ex = o.render_tile(algo, 0, y0, W, rows2, W, H, samples=samples, max_depth=depth, threads=threads)
    dt = time.perf_counter() - t0
    rays = int(ex["counts"].sum())
    return rays / dt / 1e6, (f"{W}x{rows2} rows {y0}..{y0 + rows2 - 1} of the {W}x{H} frame, {samples} spp, "
                             f"depth {depth}, {algo}; {rays} queries in {dt:.1f} s")


def cpu_configs(threads, seconds_target=5.0):
    """The oracle (-O3 -ffast-math) on bounded strips of the other configs at
    reduced spp: the 1M-triangle icospheres (1024 x 1024, depth 12, mis) and
    the largest view of the 4M-triangle smoke scene (256^3 grid, depth 8,
    volume_mis).  Per-sample rates: the GPU figures of the same configs are
    whole-frame rates of the same algorithm."""
    import tempfile

    import oracle
    import workloads

    out = {}
    ico = workloads.WORKLOADS["icospheres"]
    o = oracle.OracleScene(ico["make"](), fast=True)
    c = ico["camera"]
    o.set_camera(c["pos"], c["dir"], c["up"], c["fov"])
    v, smp = _cpu_rate(o, ico["width"], ico["height"], ico["max_depth"], threads, seconds_target)
    out["icospheres_1m"] = {"value": round(v, 3), "unit": "Mrays/s", "sample": smp}
    del o
    with tempfile.TemporaryDirectory() as td:
        scene = workloads.volume_scene(os.path.join(td, "smoke256.bin"))
        cam, w, h = max(workloads.volume_sessions(), key=lambda t: t[1] * t[2])
        o = oracle.OracleScene(scene, fast=True)
        o.set_camera(cam["pos"], cam["dir"], cam["up"], cam["fov"])
        v, smp = _cpu_rate(o, w, h, 8, threads, seconds_target, "volume_mis")
        out["volume_4m_256grid"] = {"value": round(v, 3), "unit": "Mrays/s", "sample": smp}
        del o
    return out


def cpu_baseline(scene, cam, W, H, depth, seconds_target=15.0):
    """The oracle (the CPU restatement of the reference tracer) on the host:
    the -O3 -ffast-math build (the reference module's flags) on a strip of
    the measured frame, the IEEE parity build on a shorter strip, and the
    34-triangle closed box (256 x 256, 16 spp, depth 8, mis) in full.
    Threads: OMP_NUM_THREADS, which the GPU host sets to the CPU share of one
    GPU (16); the host's own core counts and CPU model are recorded beside."""
    import oracle
    import workloads

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
    o = oracle.OracleScene(scene, fast=True)
    o.set_camera(cam["pos"], cam["dir"], cam["up"], cam["fov"])
    fast, sample = _cpu_rate(o, W, H, depth, threads, seconds_target)
    oi = oracle.OracleScene(scene, fast=False)
    oi.set_camera(cam["pos"], cam["dir"], cam["up"], cam["fov"])
    ieee, sample_i = _cpu_rate(oi, W, H, depth, threads, seconds_target / 3)
    box = workloads.closed_box()
    ob = oracle.OracleScene(box, fast=True)
    bc = workloads.CLOSED_BOX_CAMERA
    ob.set_camera(bc["pos"], bc["dir"], bc["up"], bc["fov"])
    t0 = time.perf_counter()
    _, ex = ob.render_tile("mis", 0, 0, 256, 256, 256, 256, samples=16, max_depth=8, threads=threads)
    t_box = time.perf_counter() - t0
    box_rays = int(ex["counts"].sum())
    return {"value": round(fast, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": sample + " (oracle -O3 -ffast-math, OpenMP)",
            "ieee_build": {"value": round(ieee, 3), "unit": "Mrays/s", "sample": sample_i + " (oracle -O2, IEEE)"},
            "closed_box_full": {"seconds": round(t_box, 3), "mrays_s": round(box_rays / t_box / 1e6, 3),
                                "queries": box_rays, "config": "34 triangles, 256x256, 16 spp, depth 8, mis, full"},
            "other_configs": cpu_configs(threads),
            "host": _host_info()}


def heldout_leg(ctx, sc, args, spp, depth, W=480, H=270, batches=16, batch_spp=8192):
    """Time to relMSE 0.01 on a view of the room the guiding defaults were not
    chosen on.  The reference r is plain MIS at batches x batch_spp samples per
    pixel from seeds neither tracer uses; its own expected error (the
    variance of its batch means) is subtracted from each relMSE, as for the
    oracle fixture.  Both tracers then render 16-spp passes until their
    running estimate crosses the target (every pixel counts)."""
    import ctypes as C

    import numpy as np
    import torch

    import lucid_capi
    import workloads

    L = ctx.L
    camd = lucid_capi.camera(workloads.OCCLUDED_ROOM_HELDOUT_CAMERA)
    out = torch.empty(W * H * 4, dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream

    def render(algorithm, n_spp, offset):
        prm = lucid_capi.params(0, 0, W, H, W, H, n_spp, offset, depth, algorithm)
        ctx.check(L.lucid_render_tile_device(ctx.h, sc, C.byref(camd), C.byref(prm), C.c_void_p(out.data_ptr()),
                                             C.c_void_p(stream)))
        ctx.check(L.lucid_synchronize(ctx.h))
        return out.view(H, W, 4)[..., :3].double()  # sums of n_spp samples

    t0 = time.perf_counter()
    means = torch.stack([render("mis", batch_spp, 400_000_000 + k * batch_spp) / batch_spp for k in range(batches)])
    R = means.mean(0)
    ref_err = float((means.var(0, unbiased=True) / batches / (R * R + 0.01)).mean())
    t_ref = time.perf_counter() - t0
    log(f"held-out reference: {batches * batch_spp} spp in {t_ref:.1f} s, own relMSE {ref_err:.5f}")

    def crossing(algorithm):
        acc = torch.zeros_like(R)
        t_run, n, i = 0.0, 0, 0
        while t_run < args.max_seconds:
            torch.cuda.synchronize()
            ts = time.perf_counter()
            x = render(algorithm, spp, args.seed * 10_000_000 + i * spp)
            torch.cuda.synchronize()
            t_run += time.perf_counter() - ts
            acc += x
            n += spp
            i += 1
            e = float(((acc / n - R) ** 2 / (R * R + 0.01)).mean()) - ref_err
            if e <= TARGET:
                return t_run, n, "measured"
        return t_run, n, "not reached"

    tp, np_, mp = crossing("mis")
    gc = lucid_capi.GuidingConfig(args.alpha, args.train_steps, args.batch, args.lr, args.record_stride,
                                  7 + args.seed, args.guide_depth, args.learn_selection)
    ctx.check(L.lucid_set_guiding(ctx.h, C.byref(gc)))
    ctx.check(L.lucid_guiding_reset(ctx.h))
    tg, ng, mg = crossing("guided")
    log(f"held-out view: plain {tp:.1f} s ({np_} spp), guided {tg:.1f} s ({ng} spp)")
    return {"metric": METRIC, "mode": "heldout", "view": workloads.OCCLUDED_ROOM_HELDOUT_CAMERA,
            "width": W, "height": H, "seed": args.seed, "target": TARGET, "pass_spp": spp,
            "reference": {"spp": batches * batch_spp, "kind": "plain MIS on the GPU, held-out seeds",
                          "self_relmse": round(ref_err, 5), "seconds": round(t_ref, 1)},
            "plain_s": round(tp, 2), "plain_spp": np_, "plain_method": mp,
            "guided_s": round(tg, 2), "guided_spp": ng, "guided_method": mg,
            "speedup": round(tp / tg, 3) if mp == mg == "measured" else None}


def time_to_target(curve):
    """curve: [(seconds, relMSE)].  Measured crossing (linear interpolation
    between checkpoints) or extrapolation from relMSE * t over the last half."""
    for i, (t, e) in enumerate(curve):
        if e <= TARGET:
            if i == 0:
                return t, "measured"
            t0, e0 = curve[i - 1]
            return t0 + (t - t0) * (e0 - TARGET) / max(e0 - e, 1e-30), "measured"
    tail = curve[len(curve) // 2:]
    import statistics
    c = statistics.median([t * e for t, e in tail])
    return c / TARGET, "extrapolated (relMSE ~ 1/t over the last half of the run)"


def sessions_leg(ctx, passes, spp=16, count=32, seed=7):
    """Batched sessions: `count` sessions (edges log-uniform on [128, 1024],
    scenes log-uniform on [10k, 500k] triangles, Principled and Emission
    materials, 1-8 lights), one batched pass of `spp` samples per session
    per step: fused (one launch per kernel and bounce over all sessions, the
    library default), spread over the four stream lanes, and issued one
    session at a time."""
    import lucid_capi
    import workloads

    t0 = time.perf_counter()
    specs = workloads.batched_sessions(seed=seed, count=count)
    scenes = [ctx.scene(sc) for sc, _, _, _ in specs]
    sessions = [lucid_capi.Session(ctx, h, lucid_capi.camera(cam), w, hh, 8, "mis")
                for h, (_, cam, w, hh) in zip(scenes, specs)]
    tris = sum(int(ctx.L.lucid_scene_num_triangles(h)) for h in scenes)
    pixels = sum(w * hh for _, _, w, hh in specs)
    log(f"sessions: {count} scenes, {tris} triangles, {pixels} pixels, set up in {time.perf_counter() - t0:.1f} s")
    # warm-up: allocates every lane's buffers, and runs each pass shape twice
    # (once launch by launch, once captured), so the timed passes replay graphs
    for _ in range(2):
        lucid_capi.render_sessions(ctx, sessions, spp)

    def timed(batched):
        rays, t = 0, 0.0
        for _ in range(passes):
            ts = time.perf_counter()
            if batched:
                lucid_capi.render_sessions(ctx, sessions, spp)
                st = ctx.stats()
                rays += st["rays_closest"] + st["rays_shadow"]
            else:
                for s in sessions:
                    lucid_capi.render_sessions(ctx, [s], spp)
                    st = ctx.stats()
                    rays += st["rays_closest"] + st["rays_shadow"]
            t += time.perf_counter() - ts
        return rays / t / 1e6, t / passes * 1e3

    b_mrays, b_ms = timed(True)
    ctx.check(ctx.L.lucid_set_session_fusion(ctx.h, 0))
    for _ in range(2):  # the lane passes' warm-up (allocation, graph capture)
        lucid_capi.render_sessions(ctx, sessions, spp)
    l_mrays, l_ms = timed(True)
    for _ in range(2):  # the one-at-a-time passes' own warm-up
        for s in sessions:
            lucid_capi.render_sessions(ctx, [s], spp)
    s_mrays, s_ms = timed(False)
    ctx.check(ctx.L.lucid_set_session_fusion(ctx.h, 1))
    log(f"sessions: fused {b_mrays:.0f} Mrays/s ({b_ms:.0f} ms/pass), four lanes {l_mrays:.0f} Mrays/s "
        f"({l_ms:.0f} ms/pass), one at a time {s_mrays:.0f} Mrays/s")
    for s in sessions:
        s.close()
    for h in scenes:
        ctx.L.lucid_scene_release(h)
    return {"count": count, "seed": seed, "triangles": tris, "pixels": pixels, "pass_spp": spp, "max_depth": 8,
            "algorithm": "mis", "passes": passes, "batched": "fused",
            "batched_mrays_s": round(b_mrays, 1), "batched_ms_per_pass": round(b_ms, 1),
            "lanes_mrays_s": round(l_mrays, 1), "lanes_ms_per_pass": round(l_ms, 1),
            "serial_mrays_s": round(s_mrays, 1), "serial_ms_per_pass": round(s_ms, 1),
            "speedup": round(b_mrays / max(s_mrays, 1e-9), 3),
            "fused_over_lanes": round(b_mrays / max(l_mrays, 1e-9), 3)}


def volumes_leg(ctx, passes, spp=16):
    """Participating media at scale: 16 sessions (256^2 .. 1080p) of one
    3,993,616-triangle scene with a 256^3 fBm smoke grid, traceVolumeMIS at
    depth 8, one batched pass of `spp` samples per step; FP16 grid storage
    against FP32."""
    import tempfile

    import lucid_capi
    import workloads

    views = workloads.volume_sessions()
    res = {"sessions": len(views), "pixels": sum(w * h for _, w, h in views), "pass_spp": spp,
           "max_depth": 8, "algorithm": "volume_mis"}
    with tempfile.TemporaryDirectory() as td:
        t0 = time.perf_counter()
        scene = workloads.volume_scene(os.path.join(td, "smoke256.bin"))
        log(f"volumes: scene generated in {time.perf_counter() - t0:.1f} s")
        for bits in (16, 32):
            ctx.check(ctx.L.lucid_set_grid_format(ctx.h, bits))
            t0 = time.perf_counter()
            h = ctx.scene(scene)
            pack_s = time.perf_counter() - t0
            ss = [lucid_capi.Session(ctx, h, lucid_capi.camera(cam), w, hh, 8, "volume_mis") for cam, w, hh in views]
            lucid_capi.render_sessions(ctx, ss, spp)  # warm-up
            rays, t = 0, 0.0
            for _ in range(passes):
                ts = time.perf_counter()
                lucid_capi.render_sessions(ctx, ss, spp)
                t += time.perf_counter() - ts
                st = ctx.stats()
                rays += st["rays_closest"] + st["rays_shadow"]
            mrays = rays / t / 1e6
            log(f"volumes: FP{bits} grid ({ctx.L.lucid_scene_grid_bytes(h) >> 20} MiB): {mrays:.0f} Mrays/s, "
                f"{t / passes * 1e3:.0f} ms per pass (packed in {pack_s:.1f} s)")
            res[f"fp{bits}"] = {"mrays_s": round(mrays, 1), "ms_per_pass": round(t / passes * 1e3, 1),
                                "grid_bytes": int(ctx.L.lucid_scene_grid_bytes(h))}
            res["triangles"] = int(ctx.L.lucid_scene_num_triangles(h))
            for x in ss:
                x.close()
            ctx.L.lucid_scene_release(h)
        ctx.check(ctx.L.lucid_set_grid_format(ctx.h, 32))
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="occluded_room")
    ap.add_argument("--budget", type=float, default=45.0, help="seconds of progressive rendering per tracer")
    ap.add_argument("--pass-spp", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-relmse", action="store_true", help="skip the time-to-relMSE leg")
    ap.add_argument("--no-sessions", action="store_true", help="skip the 32-session batched leg")
    ap.add_argument("--volumes", action="store_true",
                    help="also measure 16 volume sessions of the 4M-triangle smoke scene (FP16 vs FP32 grid)")
    ap.add_argument("--session-passes", type=int, default=2)
    ap.add_argument("--no-autotune", action="store_true", help="keep the library's default launch shape")
    ap.add_argument("--shade-order", choices=("auto", "0", "1"), default="auto",
                    help="shading order after tuning: auto (the autotuner's), 0 compaction, 1 by material class")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="run one device copy of this many MiB first (calibrates profiler byte counters)")
    ap.add_argument("--alpha", type=float, default=0.5)
    ap.add_argument("--train-steps", type=int, default=16)
    ap.add_argument("--batch", type=int, default=32768)
    ap.add_argument("--lr", type=float, default=1e-2)
    ap.add_argument("--record-stride", type=int, default=32)
    ap.add_argument("--guide-depth", type=int, default=0, help="guided bounces: depth < N (0: all)")
    ap.add_argument("--learn-selection", type=int, default=1, choices=(0, 1),
                    help="1: guide-sampling probability learned per surface point (starting at --alpha)")
    ap.add_argument("--crossing", action="store_true",
                    help="time-to-relMSE only: render plain, then guided, until each crosses the target (measured)")
    ap.add_argument("--seed", type=int, default=0, help="--crossing: sample-index and network seed set")
    ap.add_argument("--max-seconds", type=float, default=540.0, help="--crossing: cap per tracer")
    ap.add_argument("--guided-only", action="store_true",
                    help="time-to-relMSE leg: the guided tracer only (parameter sweeps)")
    ap.add_argument("--heldout", action="store_true",
                    help="time to relMSE 0.01 on a second view of the room (480x270) against a plain MIS reference "
                         "from held-out seeds, both crossings measured")
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
    W, H, depth, algo = wl["width"], wl["height"], wl["max_depth"], wl["algorithm"]
    relmse_leg = args.workload == "occluded_room" and not args.no_relmse
    spp = args.pass_spp if args.workload == "occluded_room" else wl["spp"]
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
    camd = lucid_capi.camera(cam)
    out = torch.empty(W * H * 4, dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream

    def render(algorithm, n_spp, offset):
        prm = lucid_capi.params(0, 0, W, H, W, H, n_spp, offset, depth, algorithm)
        ctx.check(L.lucid_render_tile_device(ctx.h, sc, C.byref(camd), C.byref(prm), C.c_void_p(out.data_ptr()),
                                             C.c_void_p(stream)))
        ctx.check(L.lucid_synchronize(ctx.h))
        return ctx.stats()

    if args.copy_probe_mib > 0:
        a = torch.empty(args.copy_probe_mib << 18, dtype=torch.float32, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
        log(f"copy probe: {args.copy_probe_mib} MiB read + written")

    shape = None
    if not args.no_autotune:  # launch shapes do not change results, only speed
        ls = lucid_capi.LaunchShape()
        t0 = time.perf_counter()
        tune_pr = lucid_capi.params(0, 0, W, H, W, H, spp, 0, depth, algo)  # one pass of the measured size
        ctx.check(L.lucid_autotune(ctx.h, sc, C.byref(camd), C.byref(tune_pr), C.byref(ls)))
        shape = {"traverse_groups_per_cu": ls.traverse_groups, "refill_idle_lanes": ls.refill,
                 "shade_groups_per_cu": ls.shade_groups, "ray_order": ls.ray_order, "shade_order": ls.shade_order,
                 "wave_paths": int(ls.wave_paths), "walk_min": int(ls.walk_min), "leaf_tris": int(ls.leaf_tris)}
        log(f"autotune ({time.perf_counter() - t0:.1f} s): {shape}, {ls.ms:.1f} ms per pass")
    if args.shade_order != "auto":
        ctx.check(L.lucid_set_shade_order(ctx.h, int(args.shade_order)))
        shape = dict(shape or {}, shade_order=int(args.shade_order))

    # ---- timed plain passes: Mrays/s ----
    for i in range(args.warmup):
        s = render(algo, spp, 1000000 + i * spp)
        log(f"warmup {i}: {s['ms_render']:.1f} ms, {s['rays_closest'] + s['rays_shadow']} rays")
    torch.cuda.synchronize()
    rays = 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        s = render(algo, spp, i * spp)
        rays += s["rays_closest"] + s["rays_shadow"]
        log(f"step {i}: {s['ms_render']:.1f} ms")
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    mrays = rays / dt / 1e6
    ms_step = dt / args.steps * 1e3
    assert torch.isfinite(out).all().item(), "non-finite pixels"

    # ---- instrumented pass: node / triangle fetch counts + per-kernel event timing ----
    ctx.check(L.lucid_set_profiling(ctx.h, lucid_capi.PROFILE_WORK | lucid_capi.PROFILE_KERNELS))
    p = render(algo, spp, 2000000)
    ctx.check(L.lucid_set_profiling(ctx.h, 0))
    q = p["rays_closest"]
    bytes_closest = q * (32 + 16) + p["nodes_closest"] * 64 + p["tris_closest"] * 48
    n_launch = max(1, p["launches_intersect"])
    avg_ms = p["ms_intersect"] / n_launch
    per_launch = bytes_closest / n_launch
    achieved = per_launch / (avg_ms / 1e3) / 1e9
    log(f"intersect: {n_launch} launches, avg {avg_ms:.3f} ms, nodes/query {p['nodes_closest'] / max(q, 1):.1f}, "
        f"tris/query {p['tris_closest'] / max(q, 1):.1f}; shade {p['ms_shade']:.1f} ms, shadow {p['ms_shadow']:.1f} ms, "
        f"ray binning {p['ms_ray_binning']:.1f} ms")

    # the other kernels of the pass, against the same 8 TB/s: shading reads and
    # writes the 172-byte path state, the hit's material (64 B) and one light
    # record (144 B) per shaded path; the shadow query reads its 32-byte ray,
    # writes a 4-byte result and fetches N'_node nodes and N'_tri triangles
    def _rate(nbytes, ms):
        return nbytes / max(ms, 1e-9) / 1e6  # GB/s

    bytes_shade = q * (2 * 172 + 64 + 144)
    qs = p["rays_shadow"]
    bytes_shadow = qs * (32 + 4) + p["nodes_shadow"] * 64 + p["tris_shadow"] * 48
    bytes_pass = bytes_closest + bytes_shade + bytes_shadow
    kernels = {
        "k_shade": {"bound": "hbm", "limiter": "latency: dependent gathers (queue entry, path state, triangle, "
                                                "material, light record)", "bytes_per_path_bounce": 2 * 172 + 64 + 144,
                    "path_bounces": int(q), "launches": int(p["launches_shade"]), "ms": round(p["ms_shade"], 2),
                    "achieved": round(_rate(bytes_shade, p["ms_shade"]), 1),
                    "frac": round(_rate(bytes_shade, p["ms_shade"]) / HBM_PEAK_GBS, 4)},
        "k_trace_refill<1>": {"bound": "hbm", "limiter": "latency (dependent node fetches); its node bytes are "
                                                         "mostly L2 hits, so the algorithmic rate can pass the HBM peak",
                              "queries": int(qs), "launches": int(p["launches_shadow"]),
                              "per_query": {"nodes": round(p["nodes_shadow"] / max(qs, 1), 2),
                                            "tris": round(p["tris_shadow"] / max(qs, 1), 2),
                                            "ray_bytes": 32, "result_bytes": 4},
                              "ms": round(p["ms_shadow"], 2),
                              "achieved": round(_rate(bytes_shadow, p["ms_shadow"]), 1),
                              "frac": round(_rate(bytes_shadow, p["ms_shadow"]) / HBM_PEAK_GBS, 4)},
        "end_to_end": {"bytes": int(bytes_pass), "ms": round(p["ms_render"], 2),
                       "achieved": round(_rate(bytes_pass, p["ms_render"]), 1),
                       "frac": round(_rate(bytes_pass, p["ms_render"]) / HBM_PEAK_GBS, 4),
                       "note": "closest hit + shade + shadow algorithmic bytes over the whole pass"},
    }

    # measured HBM traffic of the closest-hit launch for the ray order this
    # run used (the autotuner's choice; the library default bins this wave;
    # order 3 bins the shadow rays only)
    traffic = None
    ray_order = (shape or {}).get("ray_order", 1)
    if args.workload == "occluded_room" and spp == 16 and os.path.exists(TRAFFIC_PROFILE):
        key = f"traffic_bytes_per_launch_ray_order_{1 if ray_order == 1 else 0}:"
        for line in open(TRAFFIC_PROFILE):
            if line.startswith(key):
                traffic = int(line.split(":")[1])
    lights = {"occluded_room": "1 area light behind a 5 cm slit",
              "icospheres": "1 point + 1 sun + 4 area lights"}.get(args.workload, "")
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
        "config": {"workload": f"{args.workload}: {ntri} triangles, {W}x{H}, {spp} spp per step, depth {depth}, "
                               f"{algo}, {lights}",
                   "rays_per_step": rays // max(args.steps, 1), "paths_per_step": W * H * spp,
                   "launch_shape": shape,
                   "pass_ms": {"intersect": round(p["ms_intersect"], 1), "shade": round(p["ms_shade"], 1),
                               "shadow": round(p["ms_shadow"], 1), "ray_binning": round(p["ms_ray_binning"], 1)},
                   "msamples_per_s": round(W * H * spp * args.steps / dt / 1e6, 2)},
        "roofline": {"bound": "hbm", "kernel": "k_trace_refill<0> (closest-hit BVH traversal, persistent with ray refill)",
                     "limiter": "latency: dependent node fetches, mostly L2 hits (traffic beyond L2 is a "
                                "fraction of the algorithmic bytes)",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": "profiles/k_trace_refill_occluded_room_1080p_16spp.md (rocprofv3 PMC)"
                     if traffic else None,
                     "bytes_per_launch": int(per_launch), "avg_launch_ms": round(avg_ms, 4),
                     "launches": n_launch,
                     "per_query": {"nodes": round(p["nodes_closest"] / max(q, 1), 2),
                                   "tris": round(p["tris_closest"] / max(q, 1), 2),
                                   "node_bytes": 64, "tri_bytes": 48, "ray_bytes": 32, "hit_bytes": 16},
                     "other_kernels": kernels},
    }

    # ---- time-to-relMSE 0.01, guided vs plain ----
    if relmse_leg:
        d = np.load(REFERENCE)
        xs, ys, rspp = d["xs"], d["ys"], int(d["spp"])
        R = d["sum"] / rspp
        rvar = np.maximum(d["sumsq"] / rspp - R * R, 0.0) * rspp / max(rspp - 1, 1)
        ref_err = float(np.mean(rvar / rspp / (R * R + 0.01)))
        idx = torch.tensor(((H - 1 - ys) * W + xs).astype(np.int64), device="cuda")  # rows are bottom-first
        Rt = torch.tensor(R, device="cuda", dtype=torch.float64)

        def progressive(algorithm, seed_base, stop_at_target=False, budget=None):
            acc = torch.zeros(len(xs), 3, dtype=torch.float64, device="cuda")
            t_run, n, curve, i = 0.0, 0, [], 0
            budget = args.budget if budget is None else budget
            t_log = 0.0
            while t_run < budget:
                torch.cuda.synchronize()
                ts = time.perf_counter()
                render(algorithm, spp, seed_base + i * spp)
                torch.cuda.synchronize()
                t_run += time.perf_counter() - ts
                n += spp
                i += 1
                acc += out.view(-1, 4)[idx, :3].double()
                x = acc / n
                e = torch.mean((x - Rt) ** 2 / (Rt * Rt + 0.01)).item() - ref_err
                curve.append((t_run, e))
                if t_run - t_log >= 15.0:
                    log(f"{algorithm}: {t_run:.0f} s, {n} spp, relMSE {e:.4f}")
                    t_log = t_run
                if stop_at_target and e <= TARGET:
                    break
            log(f"{algorithm}: {n} spp in {t_run:.1f} s, relMSE {curve[-1][1]:.4f}")
            return curve, n

        if args.crossing:
            # both crossings measured: each tracer renders until its running
            # estimate reaches the target (checkpoints every pass)
            base = args.seed * 10_000_000
            plain_curve, plain_spp = progressive("mis", base, True, args.max_seconds)
            gc = lucid_capi.GuidingConfig(args.alpha, args.train_steps, args.batch, args.lr, args.record_stride,
                                          7 + args.seed, args.guide_depth, args.learn_selection)
            ctx.check(L.lucid_set_guiding(ctx.h, C.byref(gc)))
            ctx.check(L.lucid_guiding_reset(ctx.h))
            guided_curve, guided_spp = progressive("guided", base, True, args.max_seconds)
            tp, mp = time_to_target(plain_curve)
            tg, mg = time_to_target(guided_curve)
            print(json.dumps({
                "metric": METRIC, "mode": "crossing", "workload": result["config"]["workload"], "seed": args.seed,
                "sample_offset": base, "network_seed": 7 + args.seed, "target": TARGET, "pixels": int(len(xs)),
                "pass_spp": spp, "plain_s": round(tp, 2), "plain_method": mp, "plain_spp": plain_spp,
                "guided_s": round(tg, 2), "guided_method": mg, "guided_spp": guided_spp,
                "speedup": round(tp / tg, 3), "mrays_s_plain_passes": result["value"],
                "reference": {"spp": rspp, "self_relmse": round(ref_err, 5)},
                "guiding": {"alpha": args.alpha, "learn_selection": args.learn_selection,
                            "train_steps_per_pass": args.train_steps, "batch": args.batch,
                            "lr": args.lr, "record_stride": args.record_stride},
                "plain_curve": [(round(t, 2), round(e, 5)) for t, e in plain_curve[::8]],
                "guided_curve": [(round(t, 2), round(e, 5)) for t, e in guided_curve[::8]]}), flush=True)
            L.lucid_scene_release(sc)
            ctx.close()
            return

        if args.heldout:
            print(json.dumps(heldout_leg(ctx, sc, args, spp, depth)), flush=True)
            L.lucid_scene_release(sc)
            ctx.close()
            return

        if args.guided_only:
            gc = lucid_capi.GuidingConfig(args.alpha, args.train_steps, args.batch, args.lr, args.record_stride,
                                          7 + args.seed, args.guide_depth, args.learn_selection)
            ctx.check(L.lucid_set_guiding(ctx.h, C.byref(gc)))
            guided_curve, guided_spp = progressive("guided", args.seed * 10_000_000)
            print(json.dumps({"mode": "guided-only", "budget_s": args.budget, "seed": args.seed, "alpha": args.alpha,
                              "learn_selection": args.learn_selection,
                              "train_steps": args.train_steps, "batch": args.batch, "lr": args.lr,
                              "record_stride": args.record_stride, "guided_spp": guided_spp,
                              "relmse_end": round(guided_curve[-1][1], 5),
                              "relmse_x_t": round(guided_curve[-1][1] * guided_curve[-1][0], 4)}), flush=True)
            L.lucid_scene_release(sc)
            ctx.close()
            return
        plain_curve, plain_spp = progressive("mis", 0)
        gc = lucid_capi.GuidingConfig(args.alpha, args.train_steps, args.batch, args.lr, args.record_stride, 7,
                                      args.guide_depth, args.learn_selection)
        ctx.check(L.lucid_set_guiding(ctx.h, C.byref(gc)))
        guided_curve, guided_spp = progressive("guided", 0)
        gst = lucid_capi.GuidingStats()
        ctx.check(L.lucid_guiding_get_stats(ctx.h, C.byref(gst)))
        # one instrumented guided pass: MFMA throughput of the guiding network
        # (algorithmic FLOPs: 2 (64*64*3 + 64*40) per inference, one inference
        # per closest-hit query; training: forward + backward over 3x that per sample)
        ctx.check(L.lucid_set_profiling(ctx.h, lucid_capi.PROFILE_KERNELS))
        gp = render("guided", spp, 3000000)
        ctx.check(L.lucid_set_profiling(ctx.h, 0))
        flop_inf = 2 * (64 * 64 * 3 + 64 * 40)
        inf_tflops = gp["rays_closest"] * flop_inf / max(gp["ms_guide_infer"], 1e-9) / 1e9
        trn_tflops = args.train_steps * args.batch * 3 * flop_inf / max(gp["ms_guide_train"], 1e-9) / 1e9
        tp, mp = time_to_target(plain_curve)
        tg, mg = time_to_target(guided_curve)
        # equal-time comparison at the end of the shorter run
        t_eq = min(plain_curve[-1][0], guided_curve[-1][0])
        at = lambda cv: min(cv, key=lambda te: abs(te[0] - t_eq))[1]
        result["time_to_relmse"] = {
            "target": TARGET, "pixels": int(len(xs)), "pass_spp": spp, "budget_s": args.budget,
            "plain_s": round(tp, 2), "plain_method": mp, "guided_s": round(tg, 2), "guided_method": mg,
            "speedup": round(tp / tg, 3),
            "plain_relmse_end": round(plain_curve[-1][1], 5), "plain_spp": plain_spp,
            "guided_relmse_end": round(guided_curve[-1][1], 5), "guided_spp": guided_spp,
            "equal_time_relmse_ratio": round(at(plain_curve) / max(at(guided_curve), 1e-12), 3),
            "reference": {"spp": rspp, "self_relmse": round(ref_err, 5), "kind": "oracle FP64, stratified subset"},
            "guided_pass_ms": {"render": round(gp["ms_render"], 1), "intersect": round(gp["ms_intersect"], 1),
                               "shade": round(gp["ms_shade"], 1), "shadow": round(gp["ms_shadow"], 1),
                               "ray_binning": round(gp["ms_ray_binning"], 1),
                               "guide_infer": round(gp["ms_guide_infer"], 1),
                               "guide_train": round(gp["ms_guide_train"], 1)},
            "mfma": {"bound": "mfma", "peak": 2500.0, "unit": "TFLOP/s",
                     "infer_achieved": round(inf_tflops, 1), "infer_frac": round(inf_tflops / 2500.0, 4),
                     "train_achieved": round(trn_tflops, 1), "train_frac": round(trn_tflops / 2500.0, 4),
                     "flop_per_inference": flop_inf},
            "guiding": {"alpha": args.alpha, "learn_selection": args.learn_selection,
                            "train_steps_per_pass": args.train_steps, "batch": args.batch,
                        "lr": args.lr, "record_stride": args.record_stride, "guided_depth": args.guide_depth,
                        "samples": int(gst.samples_total),
                        "adam_steps": int(gst.train_steps), "last_loss": round(float(gst.last_loss), 4)},
        }
    # the session legs are other workloads than the one the autotuner ran on:
    # they use the library's automatic ray order (regrouping waves of >= 2^24 paths)
    ctx.check(L.lucid_set_ray_order(ctx.h, 2))
    if not args.no_sessions:
        result["sessions"] = sessions_leg(ctx, args.session_passes)
    if args.volumes:
        result["volumes"] = volumes_leg(ctx, args.session_passes)
    if not args.no_cpu:
        log("cpu baseline (oracle, fast build) ...")
        result["cpu_baseline"] = cpu_baseline(scene, cam, W, H, depth)
        if "time_to_relmse" in result:  # same algorithm and spp: scaled by the throughput ratio
            tr = result["time_to_relmse"]
            tr["cpu_plain_s_est"] = round(tr["plain_s"] * mrays / result["cpu_baseline"]["value"], 1)
    print(json.dumps(result), flush=True)
    L.lucid_scene_release(sc)
    ctx.close()


if __name__ == "__main__":
    main()
