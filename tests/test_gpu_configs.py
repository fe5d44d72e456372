"""Every workload at its own size against the oracle (the smaller parity
scenes cover the branches; these run the real shapes):

* 1M-triangle icospheres (200 level-4 icospheres, 1024 x 1024, depth 12):
  per-sample radiance on a window of the frame (>= 99 % of samples within
  1e-3), window pixel means within 3 sigma at 64 spp, and the DiagnosticFilm
  of a 64 x 64 window at 64 spp (subsample 1) against the oracle's on the
  pixels whose paths are the same on both sides (their sums agree to 1e-4;
  elsewhere one path in 64 took another turn after a last-bit difference):
  record counts, group keys, depths, strategies and light types exact, group
  and per-strategy statistics to 1e-2 (depth-12 paths over glossy metals
  carry per-sample differences up to about 1e-3).
* 32 batched sessions (10k-500k triangles each, 128^2-1024^2): one batched
  16-spp pass gives every session exactly its own render_tile; two sessions'
  windows match the oracle per sample.
* 4M triangles around a 256^3 FP16 density grid (780 level-4 icospheres,
  smoke cube, emissive quad, spot light): a window of one volume session's
  view matches the oracle run on the grid rounded to FP16 (>= 97 % of
  samples within 1e-3, the volume bar: delta tracking turns last-bit
  differences of log() into other collision points now and then).
"""

import ctypes as C
import os

import numpy as np
import pytest

import filmcheck
import lucid_capi
import oracle
import sigma
import workloads
from conftest import make_oracle, make_renderer

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def spheres():
    return workloads.icospheres(seed=2, count=200, level=4)


def _samples(ctx, sh, camd, args, spp, offset, depth, algo):
    prm = lucid_capi.params(*args, spp, offset, depth, algo)
    out = np.zeros(args[2] * args[3] * spp * 3, np.float32)
    ctx.check(ctx.L.lucid_render_samples(ctx.h, sh, C.byref(camd), C.byref(prm),
                                         out.ctypes.data_as(C.POINTER(C.c_float))))
    return out.reshape(args[2] * args[3], spp, 3)


def _oracle_samples(o, algo, args, spp, offset, depth):
    tx, ty, tw, th, fw, fh = args
    return np.stack([o.trace_pixel(algo, tx + x, ty + y, fw, fh, spp, offset, spp, depth)
                     for y in range(th) for x in range(tw)])


def _agree(g, c, bar):
    close = (np.abs(g - c) <= 1e-3 * (1.0 + np.abs(c))).all(axis=2)
    assert close.mean() >= bar, f"only {close.mean():.4f} of samples agree"
    return close.mean()


@pytest.mark.timeout(900)
def test_icospheres_1m_window_per_sample_and_3sigma(lucid, spheres):
    cam = workloads.ICOSPHERES_CAMERA
    r = make_renderer(lucid, spheres, cam, "mis")
    o = make_oracle(spheres, cam)
    args = (480, 430, 32, 32, 1024, 1024)
    g = r.debug_render_samples(*args, 2, 0, 12)
    c = _oracle_samples(o, "mis", args, 2, 0, 12)
    _agree(g, c, 0.99)
    # independent seeds, equal counts: window means within 3 sigma
    args = (440, 400, 48, 48, 1024, 1024)
    n = 64
    g = r.debug_render_samples(*args, n, 200000, 12).astype(np.float64)
    _, ex = o.render_tile("mis", *args, samples=n, sample_offset=0, max_depth=12, stats=True)
    m_c, v_c = sigma.from_sums(ex["stats"][:, :3], ex["stats"][:, 3:], n)
    out = sigma.check(g.mean(axis=1), g.var(axis=1, ddof=1), n, m_c, v_c, n, "icospheres 1M window")
    assert out["n_pixels"] > 0.5 * 48 * 48


@pytest.mark.timeout(900)
def test_icospheres_1m_diagnostic_film_64x64_at_64spp(lucid, spheres):
    cam = workloads.ICOSPHERES_CAMERA
    r = make_renderer(lucid, spheres, cam, "mis")
    cfg = lucid.PathRecordingConfig()
    cfg.subsample_factor = 1
    r.enable_diagnostics(cfg)
    args = (448, 416, 64, 64, 1024, 1024)
    img = np.asarray(r.render_tile(*args, 64, 0, 12), np.float32)
    film = np.ascontiguousarray(r.get_diagnostic_film_bytes()).view(oracle.PIXEL_DT)[..., 0]
    strat = r.get_diagnostic_strategy_stats()
    o = make_oracle(spheres, cam)
    ref_img, ref_film, ref_strat, _ = o.render_tile_diag("mis", *args, samples=64, max_depth=12)
    assert film.shape == ref_film.shape == (64, 64) and strat.shape == ref_strat.shape
    rel = np.abs(img - ref_img).sum() / np.abs(ref_img).sum()
    assert rel < 2e-2, f"radiance rel L1 {rel:.3g}"
    # pixels whose 64 paths are the same on both sides: the sums agree to 1e-4
    a = img.reshape(64, 64, 4)[::-1, :, :3]
    b = ref_img.reshape(64, 64, 4)[::-1, :, :3]
    same = (np.abs(a - b) <= 1e-4 * (1.0 + np.abs(b))).all(axis=2)
    assert same.mean() >= 0.5, f"only {same.mean():.3f} of pixels have the same paths"
    np.testing.assert_array_equal(film["total_samples"][same], ref_film["total_samples"][same])
    # depth 12 over glossy metals: contributions differ by up to ~1e-3 relative
    filmcheck.compare(film, ref_film, strat, ref_strat, same, rtol=1e-2)
    # over the whole window: record totals within 1 %
    t_g, t_c = int(film["total_samples"].sum()), int(ref_film["total_samples"].sum())
    assert abs(t_g - t_c) <= 0.01 * t_c, (t_g, t_c)
    print(f"icospheres 64x64 film: {same.mean():.3f} of pixels compared, records {t_g} vs {t_c}")


@pytest.mark.timeout(900)
def test_32_batched_sessions_equal_their_render_tiles_and_the_oracle():
    specs = workloads.batched_sessions(seed=7, count=32)
    assert len(specs) == 32
    ctx = lucid_capi.Context(0)
    scenes, sessions, cams = [], [], []
    for sc, cam, w, h in specs:
        sh = ctx.scene(sc)
        camd = lucid_capi.camera(cam)
        scenes.append(sh)
        cams.append(camd)
        sessions.append(lucid_capi.Session(ctx, sh, camd, w, h, 8, "mis"))
    lucid_capi.render_sessions(ctx, sessions, 16)
    tris = []
    for (sc, cam, w, h), sh, camd, s in zip(specs, scenes, cams, sessions):
        ref = np.zeros(w * h * 4, np.float32)
        prm = lucid_capi.params(0, 0, w, h, w, h, 16, 0, 8, "mis")
        ctx.check(ctx.L.lucid_render_tile(ctx.h, sh, C.byref(camd), C.byref(prm),
                                          ref.ctypes.data_as(C.POINTER(C.c_float))))
        got = s.read(average=False)
        assert np.array_equal(got, ref), (w, h)
        tris.append(sum(len(m["triangles"]) for m in sc["meshes"]))
    assert min(tris) < 20000 and max(tris) > 300000
    # two sessions' windows against the oracle: the largest scene and the largest frame
    for i in (int(np.argmax(tris)), int(np.argmax([w * h for _, _, w, h in specs]))):
        sc, cam, w, h = specs[i]
        args = (w // 2 - 12, h // 2 - 12, 24, 24, w, h)
        g = _samples(ctx, scenes[i], cams[i], args, 2, 0, 8, "mis")
        o = make_oracle(sc, cam)
        c = _oracle_samples(o, "mis", args, 2, 0, 8)
        _agree(g, c, 0.99)
    for s in sessions:
        s.close()
    for sh in scenes:
        ctx.L.lucid_scene_release(sh)
    ctx.close()


@pytest.mark.timeout(900)
def test_volume_4m_triangles_256_grid_fp16_session_window(tmp_path):
    path = str(tmp_path / "smoke256.bin")
    scene = workloads.volume_scene(path)
    ntri = sum(len(m["triangles"]) for m in scene["meshes"])
    assert ntri > 3_900_000
    cam, w, h = workloads.volume_sessions()[3]
    ctx = lucid_capi.Context(0)
    ctx.check(ctx.L.lucid_set_grid_format(ctx.h, 16))
    sh = ctx.scene(scene)  # reads the FP32 sidecar, stores FP16
    assert ctx.L.lucid_scene_grid_bytes(sh) == 256 ** 3 * 2
    camd = lucid_capi.camera(cam)
    # the session's frame: a window near the smoke cube at the centre of the view
    args = (w // 2 - 12, h // 2 - 12, 24, 24, w, h)
    g = _samples(ctx, sh, camd, args, 2, 0, 8, "volume_mis")
    s = lucid_capi.Session(ctx, sh, camd, w, h, 8, "volume_mis")
    lucid_capi.render_sessions(ctx, [s], 2)
    got = s.read(average=False).reshape(h, w, 4)[::-1]
    s.close()
    ctx.L.lucid_scene_release(sh)
    ctx.close()
    # the session holds the same samples (the first two) as the window render
    tx, ty, tw, th = args[:4]
    np.testing.assert_array_equal(got[ty:ty + th, tx:tx + tw, :3].reshape(-1, 3),
                                  np.maximum((g[:, 0] + g[:, 1]), 0))
    g32 = np.fromfile(path, np.float32)
    g32.astype(np.float16).astype(np.float32).tofile(path)  # the oracle reads the FP16-rounded grid
    o = make_oracle(scene, cam)
    c = _oracle_samples(o, "volume_mis", args, 2, 0, 8)
    frac = _agree(g, c, 0.97)
    assert (c.sum(axis=(1, 2)) > 0).mean() > 0.5
    print(f"volume 4M + 256^3 FP16 window: {frac:.4f} of samples agree")
