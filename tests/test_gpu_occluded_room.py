"""The plain MI355X tracer on the 1080p occluded room stays within 3 sigma of
the CPU reference: the FP64-accumulated oracle image of the stratified
16,384-pixel subset (tests/golden/occluded_room_reference.npz, 169,216 spp per
pixel, seeds from 2^24 on) against 1,024 spp of render_tile on the GPU (seeds
0..1023, so the two estimates are independent).

Per pixel channel z = (m_gpu - m_ref) / sqrt(v (1/n_gpu + 1/n_ref)), with v
the reference's own per-sample variance (known to 0.3 %), so z has mean 0 and
variance 1 under the null hypothesis whatever the shape of the pixel's
sample distribution.  This scene's samples are heavy-tailed (light arrives
through a 5 cm slit; per-sample relative variance up to ~170), so a single
pixel's z is far from Gaussian at 1,024 spp and the bars are placed where
the central limit theorem holds:

* the image-mean z over all 16,384 pixels (mean of independent unit-variance
  terms): within 4 / sqrt(16,384);
* the z of the whole subset's sum (about 5e7 samples): within 4;
* the z of each of 64 blocks of 256 pixels (2 subset rows each): the
  fraction beyond 3 within the binomial bound of tests/sigma.py.
"""

import math
import os

import numpy as np
import pytest

import workloads
from conftest import make_renderer

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "occluded_room_reference.npz")


@pytest.mark.timeout(600)
def test_plain_tracer_within_3sigma_of_cpu_reference(lucid):
    d = np.load(FIXTURE)
    W, H, depth = int(d["width"]), int(d["height"]), int(d["max_depth"])
    n_ref = int(d["spp"])
    m_ref = d["sum"] / n_ref
    v_ref = np.maximum(d["sumsq"] / n_ref - m_ref * m_ref, 0.0) * n_ref / (n_ref - 1)
    xs, ys = d["xs"], d["ys"]
    wl = workloads.WORKLOADS["occluded_room"]
    r = make_renderer(lucid, wl["make"](), wl["camera"], "mis")
    n_gpu, chunk = 1024, 64
    acc = np.zeros((len(xs), 3))
    for s0 in range(0, n_gpu, chunk):
        img = r.render_tile_numpy(0, 0, W, H, W, H, chunk, s0, depth)  # rows bottom first
        acc += img[H - 1 - ys, xs, :3].astype(np.float64)
    m_gpu = acc / n_gpu
    se = np.sqrt(v_ref * (1.0 / n_gpu + 1.0 / n_ref))
    live = se > 0
    assert live.all(axis=1).mean() > 0.95, "most subset pixels see light"
    z = np.where(live, (m_gpu - m_ref) / np.where(live, se, 1.0), 0.0)
    m_pix = len(xs)
    # image-mean z: channels of a pixel are correlated, so the pixel count is the trial count
    zp = z.mean(axis=1)
    assert abs(zp.mean()) < 4.0 / math.sqrt(m_pix) * max(1.0, zp.std()), f"image-mean z {zp.mean():+.4f}"
    # the whole subset's sum
    var_sum = (v_ref * (1.0 / n_gpu + 1.0 / n_ref)).sum()
    z_all = (m_gpu - m_ref).sum() / math.sqrt(var_sum)
    assert abs(z_all) < 4.0, f"subset-sum z {z_all:+.3f}"
    # 64 blocks of 256 pixels
    blk = np.arange(m_pix) // 256
    zb = np.array([(m_gpu - m_ref)[blk == b].sum() / math.sqrt((se[blk == b] ** 2).sum()) for b in range(64)])
    p0 = 2.0 * (1.0 - 0.5 * (1.0 + math.erf(3.0 / math.sqrt(2.0))))
    bound = p0 + 4.0 * math.sqrt(p0 * (1.0 - p0) / 64) + 1.0 / 64
    frac = float(np.mean(np.abs(zb) > 3.0))
    assert frac <= bound, f"{frac:.3f} of blocks beyond 3 sigma (bound {bound:.3f}); block z {np.round(zb, 2)}"
    # relMSE of the 1,024-spp estimate against the FP64 oracle (stated, not a bar)
    rel = workloads.relmse(m_gpu, m_ref)
    print(f"occluded room 1024 spp: image-mean z {zp.mean():+.4f}, subset-sum z {z_all:+.3f}, "
          f"blocks beyond 3 sigma {frac:.3f}, relMSE {rel:.4f} (oracle's own {float(d['self_relmse']):.5f})")
