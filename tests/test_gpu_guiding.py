"""GPU checks of learned path guiding (BF16 MFMA network, one-sample MIS).

* The training kernels' forward outputs, loss and parameter gradient equal
  the CPU reference (oracle/guide_ref.py, BF16 mode: same roundings, float64
  sums) on the same parameters and samples; so do the outputs of the
  render-path inference kernel on a traced, shuffled queue.
* Online training lowers the loss of the guide on the closed box.
* The guided estimator is unbiased: per-pixel means agree with plain MIS
  within 3 sigma (independent seeds), with a fixed guide fraction and with
  the learned per-point selection probability.
* What the network gains: on the occluded room (light behind a 5 cm slit)
  the trained guided tracer has lower relMSE than plain MIS at equal spp.
"""

import ctypes as C

import math

import numpy as np
import pytest

import guide_ref as G
import lucid_capi
import workloads
import sigma
from conftest import make_renderer

pytestmark = pytest.mark.gpu

F = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))


def test_training_step_matches_reference():
    rng = np.random.default_rng(11)
    params = G.init_params(seed=12, scale_out=1.0).astype(np.float32)
    S = 200
    box = (np.float32([-2, -1, -3]), np.float32([2, 3, 1]))
    p = rng.uniform(box[0], box[1], (S, 3))
    dirs = rng.normal(size=(3, S, 3))
    dirs /= np.linalg.norm(dirs, axis=2, keepdims=True)
    smp = np.zeros((S, 16), np.float32)
    smp[:, 0:3], smp[:, 3:6], smp[:, 6:9], smp[:, 9:12] = p, dirs[0], dirs[1], dirs[2]
    smp[:, 12] = rng.uniform(0.05, 3.0, S)
    smp[:, 13] = rng.uniform(0.02, 2.0, S)  # the sampler's density: the selection term is on
    smp[::7, 13] = 0.0
    ctx = lucid_capi.Context(0)
    out = np.zeros((S, 40), np.float32)
    grad = np.zeros(G.N_PARAMS, np.float32)
    loss = np.zeros(1, np.float32)
    ctx.check(ctx.L.lucid_guiding_debug_step(ctx.h, F(params), F(smp), S, F(box[0]), F(box[1]), F(out), F(grad),
                                             F(loss)))
    o_ref, g_ref, l_ref = G.step(params.astype(np.float64), smp.astype(np.float64), box[0], box[1], bf=True)
    scale = np.max(np.abs(o_ref))
    assert np.max(np.abs(out - o_ref)) < 2e-2 * scale
    assert abs(loss[0] - l_ref) < 2e-3 * abs(l_ref)
    for (wo, bo), (i, j) in zip(G.OFFS, G.LAYERS):
        for a, b in ((wo, bo), (bo, bo + j)):
            e = np.linalg.norm(grad[a:b] - g_ref[a:b]) / max(np.linalg.norm(g_ref[a:b]), 1e-12)
            assert e < 3e-2, (a, b, e)
    # the selection column on its own (weights into output 40 and its bias)
    wo3, bo3 = G.OFFS[3]
    col = np.r_[wo3 + np.arange(64) * 48 + G.SEL, bo3 + G.SEL]
    assert np.linalg.norm(g_ref[col]) > 0
    e = np.linalg.norm(grad[col] - g_ref[col]) / np.linalg.norm(g_ref[col])
    assert e < 3e-2, e


def test_render_inference_matches_reference():
    rng = np.random.default_rng(21)
    params = G.init_params(seed=22, scale_out=1.0).astype(np.float32)
    S = 333  # not a multiple of 64: partial tiles
    box = (np.float32([-4, 0, -6]), np.float32([4, 3, 2]))
    dirs = rng.normal(size=(3, S, 3))
    dirs /= np.linalg.norm(dirs, axis=2, keepdims=True)
    smp = np.zeros((S, 16), np.float32)
    smp[:, 0:3] = rng.uniform(box[0], box[1], (S, 3))
    smp[:, 3:6], smp[:, 6:9], smp[:, 9:12] = dirs[0], dirs[1], dirs[2]
    ctx = lucid_capi.Context(0)
    out = np.zeros((S, 40), np.float32)
    ctx.check(ctx.L.lucid_guiding_debug_infer(ctx.h, F(params), F(smp), S, F(box[0]), F(box[1]), F(out)))
    x = G.features(smp[:, 0:3], smp[:, 3:6], smp[:, 6:9], box[0], box[1])
    _, o_ref = G.forward(params.astype(np.float64), x, bf=True)
    assert np.max(np.abs(out - o_ref[:, :40])) < 2e-2 * np.max(np.abs(o_ref))


def test_render_path_inference_kernel_matches_reference():
    """k_guide_infer itself (the kernel guided renders run, not the training
    forward pass): 300,007 rays traced in the closed box, the network run over
    a shuffled queue of 287,113 of them -- not a multiple of the 256-lane
    workgroup, more than one grid stride so every workgroup makes several
    pipelined trips, with misses among the entries (path 0 among them).
    Queued hits match the float64 BF16-mode reference on the device's own
    hit frames; every other output stays untouched (NaN)."""
    rng = np.random.default_rng(31)
    params = G.init_params(seed=32, scale_out=1.0)
    ws, bs = G.unpack(params)
    ws[3][:, G.SEL] = rng.uniform(-0.5, 0.5, 64)  # a selection logit that varies
    params = G.pack(ws, bs).astype(np.float32)
    n, m = 300007, 287113
    o = rng.uniform([-0.9, 0.1, -0.9], [0.9, 1.9, 0.9], (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d[: n // 10, 2] = np.abs(d[: n // 10, 2]) + 2.0  # out through the open front: misses
    d[: n // 10] /= np.linalg.norm(d[: n // 10], axis=1, keepdims=True)
    o[:50], d[:50] = [0.0, 1.0, 5.0], [0.0, 0.0, 1.0]  # from outside, away: misses, path 0 first
    rays = np.ascontiguousarray(np.concatenate([o, d], axis=1), np.float32)
    order = rng.permutation(n)[:m].astype(np.int32)
    order[0] = 0
    ctx = lucid_capi.Context(0)
    sh = ctx.scene(workloads.closed_box())
    out = np.zeros((n, 41), np.float32)
    frame = np.zeros((n, 6), np.float32)
    box = np.zeros(6, np.float32)
    ctx.check(ctx.L.lucid_guiding_debug_render_infer(ctx.h, sh, F(params), n, F(rays),
                                                     order.ctypes.data_as(C.POINTER(C.c_int32)), m, F(out), F(frame),
                                                     F(box)))
    ctx.L.lucid_scene_release(sh)
    ctx.close()
    hit = np.isfinite(frame[:, 0])
    assert not hit[0] and 0.05 < 1 - hit.mean() < 0.2
    queued = np.zeros(n, bool)
    queued[order] = True
    live = queued & hit
    assert np.isnan(out[~live]).all(), "outputs written for paths not queued or missed"
    assert np.isfinite(out[live]).all()
    x = G.features(frame[live, :3], frame[live, 3:6], -rays[live, 3:6], box[:3], box[3:])
    _, o_ref = G.forward(params.astype(np.float64), x, bf=True)
    err = np.abs(out[live] - o_ref[:, :41])
    assert err.max() < 2e-2 * np.max(np.abs(o_ref)), err.max()
    assert np.std(out[live, G.SEL]) > 0.05  # the selection logit is stored per path


def _box_renderer(lucid, algo):
    r = make_renderer(lucid, workloads.closed_box(), workloads.CLOSED_BOX_CAMERA, algo)
    return r


def test_online_training_lowers_loss(lucid):
    r = _box_renderer(lucid, "guided")
    r.set_guiding(alpha=0.5, train_steps=16, batch=8192, learning_rate=5e-3, record_stride=1, seed=3)
    losses = []
    for i in range(10):
        r.render_tile(0, 0, 64, 64, 64, 64, 8, 8 * i, 8)
        st = r.guiding_stats()
        losses.append(st["last_loss"])
    assert st["active"] and st["samples_total"] > 10000 and st["train_steps"] == 160
    assert np.mean(losses[-3:]) < losses[0] - 0.1, losses


@pytest.mark.parametrize("learn_selection,guide_depth", [(False, 0), (True, 0), (True, 2)])
def test_guided_is_unbiased(lucid, learn_selection, guide_depth):
    """guide_depth 2: bounces 0 and 1 guided, bounce 2 weighs its emitter hits
    against bounce 1's guide density, later bounces shade with the plain kernel."""
    W = H = 24
    n = 256
    g = _box_renderer(lucid, "guided")
    g.set_guiding(alpha=0.5, train_steps=16, batch=8192, learning_rate=5e-3, record_stride=1, seed=5,
                  learn_selection=learn_selection, max_depth=guide_depth)
    for i in range(6):
        g.render_tile(0, 0, W, H, W, H, 16, 100000 + 16 * i, 8)
    sg = g.debug_render_samples(0, 0, W, H, W, H, n, 0, 8).astype(np.float64)
    p = _box_renderer(lucid, "mis")
    sp = p.debug_render_samples(0, 0, W, H, W, H, n, 50000, 8).astype(np.float64)
    mg, mp = sg.mean(1), sp.mean(1)
    # two different estimators of one image: the fraction bound per pixel, and
    # the image mean within 3 standard errors of the plain one
    sigma.check(mg, sg.var(1, ddof=1), n, mp, sp.var(1, ddof=1), n, "guided vs plain", mean_z=False)
    se_img = np.sqrt(sg.sum(axis=(0, 2)).var(ddof=1) + sp.sum(axis=(0, 2)).var(ddof=1)) / math.sqrt(n)
    diff = sg.sum(axis=(0, 2)).mean() - sp.sum(axis=(0, 2)).mean()
    assert abs(diff) < 3.0 * se_img, (diff, se_img)


def test_guiding_gain_occluded_room(lucid):
    """What the trained network gains: on the occluded room (light behind a
    5 cm slit), guided MIS reaches a lower relMSE than plain MIS at equal
    spp.  relMSE is measured against a 32,768-spp plain MIS image of a
    96 x 54 frame; the guide trains on 24 passes of a 480 x 270 frame of the
    same view (its input is the scene position, so a finer frame only gives
    it more samples).  Two independent seeds, summed errors: measured gain
    about 2x; the bar is 1.3x."""
    W, H = 96, 54
    wl = workloads.WORKLOADS["occluded_room"]
    scene, cam = wl["make"](), wl["camera"]
    ref = make_renderer(lucid, scene, cam, "mis")
    R = np.zeros((H, W, 3))
    for k in range(8):
        R += np.asarray(ref.render_tile(0, 0, W, H, W, H, 4096, (1 << 22) + 4096 * k, 12), np.float64).reshape(H, W, 4)[..., :3]
    R /= 8 * 4096

    def relmse(X):
        return float(np.mean((X - R) ** 2 / (R * R + 0.01)))

    spp = 512
    e_plain = e_guided = 0.0
    for seed in range(2):
        p = make_renderer(lucid, scene, cam, "mis")
        X = np.asarray(p.render_tile(0, 0, W, H, W, H, spp, seed * 100000, 12), np.float64).reshape(H, W, 4)[..., :3]
        g = make_renderer(lucid, scene, cam, "guided")
        g.set_guiding(alpha=0.7, train_steps=32, batch=16384, learning_rate=5e-3, record_stride=1, seed=9 + seed)
        for i in range(24):
            g.render_tile(0, 0, 480, 270, 480, 270, 16, 500000 + 16 * i, 12)
        Y = np.asarray(g.render_tile(0, 0, W, H, W, H, spp, seed * 100000, 12), np.float64).reshape(H, W, 4)[..., :3]
        e_plain += relmse(X / spp)
        e_guided += relmse(Y / spp)
    print(f"relMSE at {spp} spp (2 seeds): plain {e_plain / 2:.4f}, guided {e_guided / 2:.4f}")
    assert e_plain > 1.3 * e_guided


@pytest.mark.parametrize("guide_depth", [0, 2])
def test_degenerate_guide_falls_back_to_plain(lucid, guide_depth):
    """A learned density that is not finite (here: weight logits beyond the
    FP16 range of the stored outputs) is not sampled: every guided vertex
    falls back to BSDF sampling and plain NEE weights, so the frozen guided
    tracer renders exactly what plain MIS renders on the same seeds.  With
    guide_depth 2 the bounces past the guided ones run the plain shading
    kernel: still the same image, bit for bit."""
    box = workloads.closed_box()
    cam = workloads.CLOSED_BOX_CAMERA
    g = make_renderer(lucid, box, cam, "guided")
    g.set_guiding(alpha=0.7, train_steps=0, batch=4096, learning_rate=1e-3, record_stride=1, seed=4,
                  max_depth=guide_depth)
    params = np.asarray(g.get_guiding_params(), np.float32).copy()
    _, b3 = G.OFFS[3]
    for k in range(G.N_LOBES):
        params[b3 + 5 * k + 4] = 1e6  # weight logits: +inf once stored as FP16
    g.set_guiding_params(params)
    a = g.debug_render_samples(0, 0, 32, 24, 32, 24, 4, 11, 8)
    p = make_renderer(lucid, box, cam, "mis").debug_render_samples(0, 0, 32, 24, 32, 24, 4, 11, 8)
    assert np.array_equal(a, p)
    assert g.guiding_stats()["active"]
