"""Writes occluded_room_reference.npz: the FP64-accumulated high-spp oracle
image of the occluded-room workload on the stratified 16,384-pixel subset
that the time-to-relMSE measurement compares against.

The oracle (the CPU restatement, parity build) renders traceMIS at depth 12
with render_tile's jittered camera rays; seeds sample_offset + s with
offsets from 2^24 on, disjoint from the sample indices the GPU runs use.
Rounds of --chunk samples are accumulated until --minutes have passed; the
file is rewritten after every round with the sums, the sample count and the
reference's own estimated relMSE (mean over pixels and channels of
Var/n / (r^2 + 0.01)).

Usage: python tests/golden/make_occluded_room_reference.py [--minutes 90] [--threads 6]
"""

import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "blender-lucid-renderer_amd")]

import oracle  # noqa: E402
import workloads  # noqa: E402

OUT = os.path.join(HERE, "occluded_room_reference.npz")
OFFSET0 = 1 << 24


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--minutes", type=float, default=90.0)
    ap.add_argument("--threads", type=int, default=6)
    ap.add_argument("--chunk", type=int, default=64)
    args = ap.parse_args()
    wl = workloads.WORKLOADS["occluded_room"]
    cam = wl["camera"]
    W, H, depth = wl["width"], wl["height"], wl["max_depth"]
    xs, ys = workloads.relmse_pixels(W, H)
    o = oracle.OracleScene(wl["make"]())
    o.set_camera(cam["pos"], cam["dir"], cam["up"], cam["fov"])
    sm, sq, spp = np.zeros((len(xs), 3)), np.zeros((len(xs), 3)), 0
    if os.path.exists(OUT):
        d = np.load(OUT)
        if np.array_equal(d["xs"], xs) and np.array_equal(d["ys"], ys):
            sm, sq, spp = d["sum"].copy(), d["sumsq"].copy(), int(d["spp"])
    t_end = time.time() + args.minutes * 60
    while time.time() < t_end:
        t0 = time.time()
        o.render_pixels("mis", W, H, xs, ys, args.chunk, 4, OFFSET0 + spp, depth, args.threads, acc=(sm, sq))
        spp += args.chunk
        r = sm / spp
        var = np.maximum(sq / spp - r * r, 0.0) * spp / max(spp - 1, 1)
        self_relmse = float(np.mean(var / spp / (r * r + 0.01)))
        tmp = OUT[:-4] + ".partial.npz"  # written whole, then renamed over the fixture
        np.savez(tmp, xs=xs, ys=ys, sum=sm, sumsq=sq, spp=spp, offset0=OFFSET0, algorithm="mis", max_depth=depth,
                 width=W, height=H, self_relmse=self_relmse)
        os.replace(tmp, OUT)
        print(f"spp {spp}: round {time.time() - t0:.1f} s, reference relMSE {self_relmse:.5f}", flush=True)


if __name__ == "__main__":
    main()
