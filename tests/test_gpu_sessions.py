"""GPU checks of progressive sessions, batched session passes and the async
accumulator (render_start_async & co.).

A session that has taken N samples in any number of passes must hold
exactly the sum of one render_tile call of N samples (same seeds, jitter on
every sample, same summation order), so the checks are bit-exact against
render_tile through the same C ABI.
"""

import time

import numpy as np
import pytest

import lucid_capi
import workloads
from conftest import make_renderer

pytestmark = pytest.mark.gpu


def _render_tile(ctx, scene_h, cam, w, h, spp, depth, algo):
    out = np.zeros(w * h * 4, np.float32)
    pr = lucid_capi.params(0, 0, w, h, w, h, spp, 0, depth, algo)
    ctx.check(ctx.L.lucid_render_tile(ctx.h, scene_h, lucid_capi.C.byref(cam), lucid_capi.C.byref(pr),
                                      out.ctypes.data_as(lucid_capi.C.POINTER(lucid_capi.C.c_float))))
    return out


@pytest.fixture(scope="module")
def ctx():
    c = lucid_capi.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("algo", ["mis", "nee", "simple"])
def test_session_passes_equal_one_render_tile(ctx, algo):
    scene = workloads.closed_box()
    sh = ctx.scene(scene)
    cam = lucid_capi.camera(workloads.CLOSED_BOX_CAMERA)
    W, H, D = 48, 32, 8
    s = lucid_capi.Session(ctx, sh, cam, W, H, D, algo)
    for k in (4, 1, 3, 8):  # 16 samples in uneven passes
        lucid_capi.render_sessions(ctx, [s], k)
    assert s.samples == 16
    got = s.read(average=False)
    ref = _render_tile(ctx, sh, cam, W, H, 16, D, algo)
    assert np.array_equal(got, ref), np.abs(got - ref).max()
    avg = s.read(average=True)
    assert np.array_equal(avg[3::4], np.ones(W * H, np.float32))
    np.testing.assert_array_equal(avg[0::4], np.maximum((got[0::4] * np.float32(1 / 16)), 0))
    s.close()
    ctx.L.lucid_scene_release(sh)


@pytest.mark.parametrize("algo", ["mis", "volume_mis"])
def test_graph_replayed_passes_equal_render_tile(ctx, algo, tmp_path):
    """Equal passes run plain, are captured, then replay from the graph:
    after every pass, with eager renders in between, the session holds
    exactly render_tile of its sample count; with graphs off, the same."""
    scene = workloads.closed_box() if algo == "mis" else workloads.smoke_box(str(tmp_path / "d.bin"), grid_n=16)
    sh = ctx.scene(scene)
    cam = lucid_capi.camera(workloads.CLOSED_BOX_CAMERA)
    W, H, D = 40, 28, 6
    for graphs in (1, 0):
        ctx.check(ctx.L.lucid_set_session_graphs(ctx.h, graphs))
        s = lucid_capi.Session(ctx, sh, cam, W, H, D, algo)
        for k in range(6):
            lucid_capi.render_sessions(ctx, [s], 3)
            assert s.samples == 3 * (k + 1)
            ref = _render_tile(ctx, sh, cam, W, H, s.samples, D, algo)  # eager work between replays
            assert np.array_equal(s.read(average=False), ref), (graphs, k)
        s.close()
    ctx.check(ctx.L.lucid_set_session_graphs(ctx.h, 1))
    ctx.L.lucid_scene_release(sh)


def test_batched_sessions_equal_their_own_renders(ctx):
    """Eight sessions of different scenes, sizes, depths and algorithms in
    one batch (spread over the streams), two passes."""
    cases = []
    for i, (sc, cam, w, h) in enumerate(workloads.batched_sessions(seed=11, count=8, tri_range=(2e3, 4e4),
                                                                    edge_range=(17, 96))):
        algo = ("mis", "nee", "simple", "mis")[i % 4]
        depth = (8, 12, 5, 3)[i % 4]
        sh = ctx.scene(sc)
        c = lucid_capi.camera(cam)
        cases.append((sh, c, w, h, depth, algo, lucid_capi.Session(ctx, sh, c, w, h, depth, algo)))
    sessions = [x[-1] for x in cases]
    lucid_capi.render_sessions(ctx, sessions, 6)
    lucid_capi.render_sessions(ctx, sessions, 10)
    st = ctx.stats()
    assert st["paths"] == sum(w * h * 10 for _, _, w, h, _, _, _ in cases)
    assert st["rays_closest"] >= st["paths"] and st["stack_overflows"] == 0
    for _ in range(3):  # equal passes: plain, captured, replayed
        lucid_capi.render_sessions(ctx, sessions, 4)
        st = ctx.stats()
        assert st["paths"] == sum(w * h * 4 for _, _, w, h, _, _, _ in cases)
    for sh, c, w, h, depth, algo, s in cases:
        assert s.samples == 28
        got = s.read(average=False)
        ref = _render_tile(ctx, sh, c, w, h, 28, depth, algo)
        assert np.array_equal(got, ref), (w, h, algo, np.abs(got - ref).max())
        assert got[0::4].sum() > 0
    for sh, *_, s in cases:
        s.close()
        ctx.L.lucid_scene_release(sh)


@pytest.mark.parametrize("algo", ["mis", "nee", "simple"])
def test_fused_pass_equals_lanes_and_render_tile(ctx, algo):
    """Sessions of one algorithm on different scenes (one scene shared by two
    sessions), sizes and depths: a fused pass (one launch per kernel and
    bounce over all of them) and the stream-lane batch both leave every
    session with the sums of its own render_tile, bit for bit, over passes
    of several sizes (a pass of 7 samples spans two waves)."""
    cases = []
    scenes = {}
    for i, (sc, cam, w, h) in enumerate(workloads.batched_sessions(seed=5, count=6, tri_range=(2e3, 3e4),
                                                                    edge_range=(16, 80))):
        key = i if i != 5 else 1  # sessions 1 and 5 render the same scene
        if key not in scenes:
            scenes[key] = (ctx.scene(sc), lucid_capi.camera(cam))
        sh, c = scenes[key]
        depth = (8, 5, 12, 3, 8, 6)[i]
        cases.append((sh, c, w, h, depth))
    P = sum(w * h for _, _, w, h, _ in cases)
    for fused in (1, 0):
        ctx.check(ctx.L.lucid_set_session_fusion(ctx.h, fused))
        if fused:  # waves of 3 samples: the 7-sample pass takes three
            ctx.check(ctx.L.lucid_set_wave_paths(ctx.h, 3 * P))
        sessions = [lucid_capi.Session(ctx, sh, c, w, h, depth, algo) for sh, c, w, h, depth in cases]
        for spp in (4, 7, 1):
            lucid_capi.render_sessions(ctx, sessions, spp)
            st = ctx.stats()
            assert st["paths"] == P * spp and st["rays_closest"] >= P * spp
        ctx.check(ctx.L.lucid_set_wave_paths(ctx.h, 0))
        for (sh, c, w, h, depth), s in zip(cases, sessions):
            assert s.samples == 12
            got = s.read(average=False)
            ref = _render_tile(ctx, sh, c, w, h, 12, depth, algo)
            assert np.array_equal(got, ref), (fused, w, h, np.abs(got - ref).max())
            assert got[0::4].sum() > 0
        for s in sessions:
            s.close()
    ctx.check(ctx.L.lucid_set_session_fusion(ctx.h, 1))
    for sh, _ in scenes.values():
        ctx.L.lucid_scene_release(sh)


def test_session_set_camera_clears(ctx):
    scene = workloads.closed_box()
    sh = ctx.scene(scene)
    cam = lucid_capi.camera(workloads.CLOSED_BOX_CAMERA)
    s = lucid_capi.Session(ctx, sh, cam, 32, 24, 6, "mis")
    lucid_capi.render_sessions(ctx, [s], 5)
    cam2 = lucid_capi.camera(dict(workloads.CLOSED_BOX_CAMERA, pos=(0.2, 1.1, 3.4)))
    ctx.check(ctx.L.lucid_session_set_camera(s.h, lucid_capi.C.byref(cam2), 6))
    assert s.samples == 0
    lucid_capi.render_sessions(ctx, [s], 7)
    assert np.array_equal(s.read(), _render_tile(ctx, sh, cam2, 32, 24, 7, 6, "mis"))
    s.close()
    ctx.L.lucid_scene_release(sh)


def _wait(pred, seconds=30.0):
    t0 = time.time()
    while not pred():
        assert time.time() - t0 < seconds, "timed out waiting for the async session"
        time.sleep(0.005)


def test_async_session_accumulates_and_resets(ctx):
    scene = workloads.closed_box()
    sh = ctx.scene(scene)
    cam = lucid_capi.camera(workloads.CLOSED_BOX_CAMERA)
    W, H, D = 64, 48, 8
    s = lucid_capi.Session(ctx, sh, cam, W, H, D, "mis")
    ctx.check(ctx.L.lucid_session_start_async(s.h))
    assert ctx.L.lucid_session_async_running(s.h) == 1
    _wait(lambda: s.poll()[0] >= 64)
    n1, px1, rev1 = s.poll()
    assert px1 is not None and np.all(np.isfinite(px1)) and np.all(px1[3::4] == 1.0)
    # reset to another camera: the count restarts and old frames are not published
    cam2 = lucid_capi.camera(dict(workloads.CLOSED_BOX_CAMERA, dir=(0.1, 0.0, -1.0)))
    ctx.check(ctx.L.lucid_session_reset_async(s.h, lucid_capi.C.byref(cam2), W, H, D))
    _wait(lambda: s.poll()[2] > rev1)
    assert ctx.L.lucid_session_reset_async(s.h, lucid_capi.C.byref(cam2), W + 1, H, D) == 1  # other size
    _wait(lambda: s.poll()[0] >= 16)
    ctx.check(ctx.L.lucid_session_stop_async(s.h))
    assert ctx.L.lucid_session_async_running(s.h) == 0
    assert ctx.L.lucid_session_error(s.h) == b""
    n = s.samples
    assert n >= 16
    # the final accumulator is exactly render_tile(n) of the new camera,
    # the final snapshot its average
    assert np.array_equal(s.read(average=False), _render_tile(ctx, sh, cam2, W, H, n, D, "mis"))
    avg = s.read(average=True)
    ref_avg = _render_tile(ctx, sh, cam2, W, H, n, D, "mis")[0::4] * (np.float32(1.0) / np.float32(n))
    np.testing.assert_array_equal(avg[0::4], np.maximum(ref_avg, 0))
    s.close()
    ctx.L.lucid_scene_release(sh)


def test_renderer_async_api(lucid):
    r = make_renderer(lucid, workloads.closed_box(), workloads.CLOSED_BOX_CAMERA, "mis")
    r.render_start_async(80, 60, 8)
    assert r.is_render_async_running()
    _wait(lambda: r.render_poll_async()[0] > 0)
    rev0 = r.render_snapshot_revision_async()
    _wait(lambda: r.render_snapshot_revision_async() > rev0)
    n, px = r.render_poll_async()
    assert n > 0 and px.shape == (80 * 60 * 4,) and np.all(px[3::4] == 1.0) and np.all(px[0::4] >= 0)
    # the average converges to render_tile / samples: compare the image means
    ref = np.asarray(r.render_tile(0, 0, 80, 60, 80, 60, 256, 0, 8), np.float64).reshape(-1, 4)[:, :3] / 256
    _wait(lambda: r.render_poll_async()[0] >= 256)
    n, px = r.render_poll_async()
    assert abs(px.reshape(-1, 4)[:, :3].mean() / ref.mean() - 1.0) < 0.05
    r.render_reset_async(80, 60, 8)
    r.render_reset_async(81, 60, 8)  # size mismatch: logged, session keeps running
    assert r.is_render_async_running()
    r.render_stop_async()
    assert not r.is_render_async_running()
    assert r.render_poll_async()[0] == 0 and r.render_snapshot_revision_async() == 0
    # a render_tile after the async session is unaffected
    px2 = r.render_tile(0, 0, 16, 16, 16, 16, 4, 0, 8)
    assert len(px2) == 16 * 16 * 4


def test_batch_with_guided_session_and_empty_scene(ctx):
    """Guided sessions share the context's network and run on the first
    lane; a scene without geometry renders the environment; the plain
    sessions of the same batch stay bit-exact."""
    box = workloads.closed_box()
    sh_box = ctx.scene(box)
    env_only = {"meshes": [], "native_lights": [], "environment": {"color": [0.2, 0.3, 0.4], "strength": 2.0}}
    sh_env = ctx.scene(env_only)
    cam = lucid_capi.camera(workloads.CLOSED_BOX_CAMERA)
    gc = lucid_capi.GuidingConfig(0.5, 8, 4096, 5e-3, 1, 3, 0)
    ctx.check(ctx.L.lucid_set_guiding(ctx.h, lucid_capi.C.byref(gc)))
    guided = lucid_capi.Session(ctx, sh_box, cam, 40, 30, 8, "guided")
    plain = lucid_capi.Session(ctx, sh_box, cam, 40, 30, 8, "mis")
    empty = lucid_capi.Session(ctx, sh_env, cam, 16, 8, 8, "mis")
    for _ in range(3):
        lucid_capi.render_sessions(ctx, [guided, plain, empty], 4)
    assert guided.samples == plain.samples == empty.samples == 12
    g = guided.read(average=True)
    assert np.all(np.isfinite(g)) and g[0::4].mean() > 0
    assert np.array_equal(plain.read(), _render_tile(ctx, sh_box, cam, 40, 30, 12, 8, "mis"))
    e = empty.read(average=True).reshape(-1, 4)
    np.testing.assert_allclose(e[:, :3], np.tile(np.float32([0.4, 0.6, 0.8]), (16 * 8, 1)), rtol=1e-6)
    gs = lucid_capi.GuidingStats()
    ctx.check(ctx.L.lucid_guiding_get_stats(ctx.h, lucid_capi.C.byref(gs)))
    assert gs.train_steps == 3 * 8 and gs.samples_total > 0
    for s in (guided, plain, empty):
        s.close()
    ctx.L.lucid_scene_release(sh_box)
    ctx.L.lucid_scene_release(sh_env)


def test_async_guided_session(ctx):
    scene = workloads.closed_box()
    sh = ctx.scene(scene)
    gc = lucid_capi.GuidingConfig(0.5, 4, 4096, 5e-3, 1, 5, 0)
    ctx.check(ctx.L.lucid_set_guiding(ctx.h, lucid_capi.C.byref(gc)))
    s = lucid_capi.Session(ctx, sh, lucid_capi.camera(workloads.CLOSED_BOX_CAMERA), 48, 32, 6, "guided")
    ctx.check(ctx.L.lucid_session_start_async(s.h))
    _wait(lambda: s.poll()[0] >= 32)
    ctx.check(ctx.L.lucid_session_stop_async(s.h))
    assert ctx.L.lucid_session_error(s.h) == b""
    avg = s.read(average=True)
    assert np.all(np.isfinite(avg)) and avg[0::4].mean() > 0
    s.close()
    ctx.L.lucid_scene_release(sh)
