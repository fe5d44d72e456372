"""The plain MI355X tracer on the 1080p occluded room stays within 3 sigma of
the CPU reference: the FP64-accumulated oracle image of the stratified
16,384-pixel subset (tests/golden/occluded_room_reference.npz, 169,216 spp per
pixel, seeds from 2^24 on) against 1,024 spp of render_tile on the GPU (seeds
0..1023, so the two estimates are independent).

Per pixel channel z = (m_gpu - m_ref) / sqrt(v (1/n_gpu + 1/n_ref)), with v
the reference's own per-sample variance (known to 0.3 %), so z has mean 0 and
variance 1 under the null hypothesis whatever the shape of the pixel's
sample distribution.  This scene's samples are heavy-tailed (light arrives
through a 5 cm slit; per-sample relative variance up to ~170): a rare path
through the slit carries a large value, so a pixel's z has a long upper tail
at 1,024 spp (one such firefly can put a 256-pixel block at z = +5), while
the lower tail of a mean of non-negative samples is no heavier than the
Gaussian one.  The bars use what holds whatever the skew:

* the image-mean z over all 16,384 pixels (mean of independent unit-variance
  terms): within 3 / sqrt(16,384) -- a bias of 0.03 sigma per pixel fails it;
* the z of the whole subset's sum (about 5e7 samples): within 4;
* the z of each of 8 regions of 2,048 pixels (16 subset rows each): every
  one within 4.5 (Bonferroni over 8 at about 5e-4);
* per pixel, the lower tail: the fraction of channels with z < -3 within
  four binomial standard errors of the one-sided Gaussian rate 0.00135 -- a
  tracer that loses energy anywhere fails it; the fraction beyond +3 is
  reported.
"""

import math
import os

import numpy as np
import pytest

import workloads
from conftest import make_renderer

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "occluded_room_reference.npz")


def _reference():
    d = np.load(FIXTURE)
    n_ref = int(d["spp"])
    m_ref = d["sum"] / n_ref
    v_ref = np.maximum(d["sumsq"] / n_ref - m_ref * m_ref, 0.0) * n_ref / (n_ref - 1)
    return d, n_ref, m_ref, v_ref


def _render_subset(r, d, n_gpu, seed0=0, chunk=64):
    W, H, depth = int(d["width"]), int(d["height"]), int(d["max_depth"])
    xs, ys = d["xs"], d["ys"]
    acc = np.zeros((len(xs), 3))
    for s0 in range(seed0, seed0 + n_gpu, chunk):
        img = r.render_tile_numpy(0, 0, W, H, W, H, chunk, s0, depth)  # rows bottom first
        acc += img[H - 1 - ys, xs, :3].astype(np.float64)
    return acc / n_gpu


def _check_against_reference(m_gpu, n_gpu, what):
    d, n_ref, m_ref, v_ref = _reference()
    se = np.sqrt(v_ref * (1.0 / n_gpu + 1.0 / n_ref))
    live = se > 0
    assert live.all(axis=1).mean() > 0.95, "most subset pixels see light"
    z = np.where(live, (m_gpu - m_ref) / np.where(live, se, 1.0), 0.0)
    m_pix = len(d["xs"])
    zp = z.mean(axis=1)  # channels of a pixel are correlated: the pixel count is the trial count
    var_sum = (v_ref * (1.0 / n_gpu + 1.0 / n_ref)).sum()
    z_all = (m_gpu - m_ref).sum() / math.sqrt(var_sum)
    reg = np.arange(m_pix) // 2048
    zr = np.array([(m_gpu - m_ref)[reg == b].sum() / math.sqrt((se[reg == b] ** 2).sum()) for b in range(8)])
    lo, hi = float(np.mean(z[live] < -3.0)), float(np.mean(z[live] > 3.0))
    rel = workloads.relmse(m_gpu, m_ref)
    print(f"occluded room, {what}, {n_gpu} spp: image-mean z {zp.mean():+.4f}, subset-sum z {z_all:+.3f}, region z "
          f"{np.round(zr, 2)}, channels below -3 sigma {lo:.4f}, above +3 sigma {hi:.4f}, relMSE {rel:.4f} "
          f"(oracle's own {float(d['self_relmse']):.5f})")
    assert abs(zp.mean()) < 3.0 / math.sqrt(m_pix), f"{what}: image-mean z {zp.mean():+.4f}"
    assert abs(z_all) < 4.0, f"{what}: subset-sum z {z_all:+.3f}"
    assert np.abs(zr).max() < 4.5, f"{what}: region z {np.round(zr, 2)}"
    p1 = 0.00135
    assert lo <= p1 + 4.0 * math.sqrt(p1 * (1.0 - p1) / m_pix), f"{what}: {lo:.4f} of channels below -3 sigma"


@pytest.mark.timeout(600)
def test_plain_tracer_within_3sigma_of_cpu_reference(lucid):
    d = _reference()[0]
    wl = workloads.WORKLOADS["occluded_room"]
    r = make_renderer(lucid, wl["make"](), wl["camera"], "mis")
    _check_against_reference(_render_subset(r, d, 1024), 1024, "plain mis")


@pytest.mark.timeout(600)
def test_guided_tracer_within_3sigma_of_cpu_reference(lucid):
    """The guided estimator against the same reference: a network trained
    online for 32 passes (learned per-point selection probability, the bench
    default), then frozen (its weights copied into a guide that no longer
    trains), renders 1,024 spp of new seeds."""
    d = _reference()[0]
    wl = workloads.WORKLOADS["occluded_room"]
    W, H, depth = int(d["width"]), int(d["height"]), int(d["max_depth"])
    r = make_renderer(lucid, wl["make"](), wl["camera"], "guided")
    r.set_guiding(alpha=0.5, train_steps=16, batch=32768, learning_rate=1e-2, record_stride=32, seed=7,
                  learn_selection=True)
    for i in range(32):
        r.render_tile_numpy(0, 0, W, H, W, H, 16, 5_000_000 + 16 * i, depth)
    params = np.asarray(r.get_guiding_params(), np.float32)
    r.set_guiding(alpha=0.5, train_steps=0, batch=32768, learning_rate=1e-2, record_stride=32, seed=7,
                  learn_selection=True)
    r.set_guiding_params(params)
    _check_against_reference(_render_subset(r, d, 1024), 1024, "guided, frozen network")
