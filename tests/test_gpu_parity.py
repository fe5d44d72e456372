"""GPU parity: the HIP path against the oracle on the same seeded inputs.

Bars (the tolerances are written in each test):
* camera-ray closest hits: bit-exact (mesh, triangle, u, v, t);
* per-sample radiance: the same seed stream gives the same path up to
  last-ulp differences of the device transcendentals (sin, cos, pow, log), so
  at least 99 % of samples agree to 1e-3 relative and the tile sums agree to
  1 % relative L1;
* per-pixel means over independent seeds agree within 3 sigma of the Monte
  Carlo error: the fraction of |z| > 3 within a binomial bound on the t
  rate, and the image-mean signed z within 4 standard errors of 0
  (tests/sigma.py).
"""

import numpy as np
import pytest

import workloads
import sigma
from conftest import make_oracle, make_renderer

pytestmark = pytest.mark.gpu

BOX = workloads.closed_box()
BOX_CAM = workloads.CLOSED_BOX_CAMERA


def _small_spheres():
    return workloads.icospheres(seed=5, count=24, level=3)


def _same_bits(a, b):
    return np.asarray(a).view(np.int32) == np.asarray(b).view(np.int32)


@pytest.mark.parametrize("spp", [1, 4])
def test_primary_hits_bit_exact_closed_box(lucid, spp):
    r = make_renderer(lucid, BOX, BOX_CAM)
    o = make_oracle(BOX, BOX_CAM)
    W = H = 256
    g = r.debug_primary_hits(0, 0, W, H, W, H, spp, 0)
    c = o.primary_hits(0, 0, W, H, W, H, samples=spp)
    for k in ("mesh", "tri", "u", "v", "t"):
        same = _same_bits(g[k], c[k])
        assert same.all(), f"{k}: {np.count_nonzero(~same)} of {same.size} rays differ"


def test_primary_hits_bit_exact_spheres_subtile(lucid):
    scene = workloads.icospheres(seed=2, count=200, level=4)
    cam = workloads.ICOSPHERES_CAMERA
    r = make_renderer(lucid, scene, cam)
    o = make_oracle(scene, cam)
    # a 256x192 window of the 1024^2 frame, 2 jittered samples per pixel
    g = r.debug_primary_hits(384, 400, 256, 192, 1024, 1024, 2, 7)
    c = o.primary_hits(384, 400, 256, 192, 1024, 1024, samples=2, sample_offset=7)
    assert (c["mesh"] >= 0).mean() > 0.2, "window should see geometry"
    for k in ("mesh", "tri", "u", "v", "t"):
        same = _same_bits(g[k], c[k])
        assert same.all(), f"{k}: {np.count_nonzero(~same)} of {same.size} rays differ"


@pytest.mark.parametrize("algorithm", ["mis", "nee", "simple"])
def test_per_sample_radiance_closed_box(lucid, algorithm):
    r = make_renderer(lucid, BOX, BOX_CAM, algorithm)
    o = make_oracle(BOX, BOX_CAM)
    W = H = 48
    spp = 4
    g = r.debug_render_samples(0, 0, W, H, W, H, spp, 3, 8)
    c = np.stack([o.trace_pixel(algorithm, x, y, W, H, spp, 3, spp, 8) for y in range(H) for x in range(W)])
    assert g.shape == c.shape
    tol = 1e-3 * (1.0 + np.abs(c))
    close = (np.abs(g - c) <= tol).all(axis=2)
    assert close.mean() >= 0.99, f"only {close.mean():.4f} of samples agree"


def test_render_tile_matches_oracle_closed_box(lucid):
    r = make_renderer(lucid, BOX, BOX_CAM, "mis")
    o = make_oracle(BOX, BOX_CAM)
    W = H = 128
    img = np.asarray(r.render_tile(0, 0, W, H, W, H, 16, 0, 8), np.float32)
    ref, _ = o.render_tile("mis", 0, 0, W, H, W, H, samples=16, max_depth=8)
    assert img.shape == ref.shape == (W * H * 4,)
    assert np.all(img[3::4] == 1.0)
    rel = np.abs(img - ref).sum() / np.abs(ref).sum()
    assert rel < 1e-2, f"relative L1 {rel:.3g}"


def test_render_tile_subtile_offsets(lucid):
    """tile_x/tile_y/sample_offset select the same pixels and seeds as the oracle."""
    r = make_renderer(lucid, BOX, BOX_CAM, "mis")
    o = make_oracle(BOX, BOX_CAM)
    args = (40, 72, 33, 17, 256, 256)
    img = np.asarray(r.render_tile(*args, 5, 11, 8), np.float32)
    ref, _ = o.render_tile("mis", *args, samples=5, sample_offset=11, max_depth=8)
    rel = np.abs(img - ref).sum() / np.abs(ref).sum()
    assert rel < 2e-2, f"relative L1 {rel:.3g}"


def test_spheres_statistics_3sigma(lucid):
    """Independent seeds: GPU and oracle pixel means agree within 3 sigma."""
    scene = _small_spheres()
    cam = workloads.ICOSPHERES_CAMERA
    r = make_renderer(lucid, scene, cam, "mis")
    o = make_oracle(scene, cam)
    W = H = 48
    n = 64
    g = r.debug_render_samples(0, 0, W, H, W, H, n, 1000, 12).astype(np.float64)
    _, ex = o.render_tile("mis", 0, 0, W, H, W, H, samples=n, sample_offset=0, max_depth=12, stats=True)
    st = ex["stats"]
    m_c, v_c = sigma.from_sums(st[:, :3], st[:, 3:], n)
    sigma.check(g.mean(axis=1), g.var(axis=1, ddof=1), n, m_c, v_c, n, "spheres mis")


def test_env_only_scene_exact(lucid):
    """Nothing to hit: every sample returns the environment colour x strength
    (test_path_tracer.cpp:132-152 / test_gpu_path_tracer.cpp env-only case)."""
    scene = {"meshes": [], "environment": {"color": [0.5, 0.6, 0.7], "strength": 1.0}}
    r = make_renderer(lucid, scene, BOX_CAM, "mis")
    img = np.asarray(r.render_tile(0, 0, 16, 16, 16, 16, 3, 0, 8), np.float32).reshape(-1, 4)
    want = np.broadcast_to(np.float32([0.5, 0.6, 0.7]) * np.float32(3), img[:, :3].shape)
    np.testing.assert_array_equal(img[:, :3], want)


def test_cancel_before_render_is_reset(lucid):
    r = make_renderer(lucid, BOX, BOX_CAM, "mis")
    r.cancel()
    assert r.is_cancelled()
    img = np.asarray(r.render_tile(0, 0, 8, 8, 8, 8, 2, 0, 4), np.float32)
    assert not r.is_cancelled()
    assert img[0::4].sum() > 0


def test_numpy_and_stats(lucid):
    r = make_renderer(lucid, BOX, BOX_CAM, "mis")
    a = r.render_tile_numpy(0, 0, 32, 16, 32, 16, 2, 0, 8)
    assert a.shape == (16, 32, 4)
    st = r.last_render_stats()
    assert st["paths"] == 32 * 16 * 2 and st["rays_closest"] >= st["paths"] and st["stack_overflows"] == 0


@pytest.mark.parametrize("case", ["box_mis", "box_nee", "spheres_mis", "smoke_volume_mis"])
def test_traversal_launch_shapes_agree(case, tmp_path):
    """The persistent refill traversal (default) and one lane per ray give
    the same per-sample radiance bit for bit, whatever the refill threshold,
    with or without regrouping the ray queues by bin and with or without
    regrouping the shading queue by material class, including
    volume-boundary re-queries (smoke box) and any-hit shadows."""
    import ctypes as C

    import lucid_capi

    if case.startswith("box"):
        scene, cam, W, H = workloads.closed_box(), workloads.CLOSED_BOX_CAMERA, 48, 40
    elif case.startswith("spheres"):
        scene, cam, W, H = _small_spheres(), workloads.ICOSPHERES_CAMERA, 64, 48
    else:
        scene = workloads.smoke_box(str(tmp_path / "d.bin"), grid_n=16)
        cam, W, H = workloads.SMOKE_BOX_CAMERA, 40, 32
    algo = case.split("_", 1)[1]
    ctx = lucid_capi.Context(0)
    sh = ctx.scene(scene)
    camd = lucid_capi.camera(cam)
    spp = 3
    prm = lucid_capi.params(0, 0, W, H, W, H, spp, 7, 8, algo)
    outs = []
    for shape, refill, order, shade in ((0, 48, 0, 0), (16, 48, 0, 0), (16, 1, 1, 1), (3, 64, 1, 0), (0, 48, 1, 1),
                                        (16, 48, 2, 1), (16, 48, 3, 0)):
        ctx.check(ctx.L.lucid_set_launch_shape(ctx.h, shape, 16))
        ctx.check(ctx.L.lucid_set_refill_threshold(ctx.h, refill))
        ctx.check(ctx.L.lucid_set_ray_order(ctx.h, order))
        ctx.check(ctx.L.lucid_set_shade_order(ctx.h, shade))
        out = np.zeros(W * H * spp * 3, np.float32)
        ctx.check(ctx.L.lucid_render_samples(ctx.h, sh, C.byref(camd), C.byref(prm),
                                             out.ctypes.data_as(C.POINTER(C.c_float))))
        st = ctx.stats()
        assert st["stack_overflows"] == 0
        outs.append((out, st["rays_closest"], st["rays_shadow"]))
    for out, rc, rs in outs[1:]:
        assert np.array_equal(out.view(np.int32), outs[0][0].view(np.int32))
        assert (rc, rs) == outs[0][1:]
    assert outs[0][0].sum() > 0
    ctx.L.lucid_scene_release(sh)
    ctx.close()


def test_queue_sizes_and_shapes_keep_samples():
    """Queues reach the persistent launches through 64 fetch ranges (chunks of
    256 positions dealt round-robin) and leave shading through 64 append
    segments (one per 128-position input chunk, modulo 64) joined by a
    concatenation launch.  On the room at 160 x 120 and 8 spp every bounce's
    queue has another length (Russian roulette), and waves of 1, 3 and 8
    samples per pixel and three launch shapes move those lengths across
    range and segment boundaries: every sample's radiance stays bit-identical,
    and so do the ray counts."""
    import ctypes as C
    import lucid_capi

    wl = workloads.WORKLOADS["occluded_room"]
    ctx = lucid_capi.Context(0)
    sh = ctx.scene(wl["make"]())
    camd = lucid_capi.camera(wl["camera"])
    W, H, spp = 160, 120, 8
    prm = lucid_capi.params(0, 0, W, H, W, H, spp, 11, 12, "mis")
    outs = []
    for wave_b, groups, refill in ((0, 16, 48), (1, 16, 48), (3, 8, 60), (8, 24, 1)):
        ctx.check(ctx.L.lucid_set_wave_paths(ctx.h, W * H * wave_b))
        ctx.check(ctx.L.lucid_set_launch_shape(ctx.h, groups, 16))
        ctx.check(ctx.L.lucid_set_refill_threshold(ctx.h, refill))
        out = np.zeros(W * H * spp * 3, np.float32)
        ctx.check(ctx.L.lucid_render_samples(ctx.h, sh, C.byref(camd), C.byref(prm),
                                             out.ctypes.data_as(C.POINTER(C.c_float))))
        st = ctx.stats()
        outs.append((out, st["rays_closest"], st["rays_shadow"]))
    for out, rc, rs in outs[1:]:
        assert np.array_equal(out.view(np.int32), outs[0][0].view(np.int32))
        assert (rc, rs) == outs[0][1:]
    assert outs[0][1] > W * H * spp * 4  # paths live for several bounces
    ctx.L.lucid_scene_release(sh)
    ctx.close()


def test_autotune_picks_a_shape_and_keeps_results(lucid):
    r = make_renderer(lucid, BOX, BOX_CAM, "mis")
    before = np.asarray(r.render_tile(0, 0, 64, 48, 64, 48, 8, 5, 8))
    shape = r.autotune(64, 48, 4, 8)
    assert shape["traverse_groups"] in (8, 12, 16, 24) and shape["refill"] in (32, 48, 60)
    assert shape["shade_groups"] in (4, 8, 16) and shape["ray_order"] in (0, 1, 3) and shape["ms"] > 0
    assert shape["shade_order"] in (0, 1) and shape["wave_paths"] == 0  # too few paths to split
    after = np.asarray(r.render_tile(0, 0, 64, 48, 64, 48, 8, 5, 8))
    assert np.array_equal(before, after)


def test_autotune_wave_size_keeps_results(lucid):
    """The wavefront batch is searched too (one wave, or the pass in 2 or 4
    waves of whole samples); the film adds samples in the same order either
    way, so the image is bit-identical."""
    r = make_renderer(lucid, BOX, BOX_CAM, "mis")
    W = H = 512
    before = np.asarray(r.render_tile(0, 0, W, H, W, H, 16, 3, 8))
    shape = r.autotune(W, H, 16, 8)
    assert shape["wave_paths"] in (0, W * H * 8, W * H * 4)
    after = np.asarray(r.render_tile(0, 0, W, H, W, H, 16, 3, 8))
    assert np.array_equal(before, after)
    r.set_wave_paths(W * H * 4)  # four waves of 4 samples
    assert np.array_equal(np.asarray(r.render_tile(0, 0, W, H, W, H, 16, 3, 8)), before)


def test_ray_order_arguments():
    """Ray orders 0-3 are accepted (3: shadow rays only), others refused."""
    import lucid_capi

    ctx = lucid_capi.Context(0)
    for order in (0, 1, 2, 3):
        assert ctx.L.lucid_set_ray_order(ctx.h, order) == 0
    for order in (-1, 4):
        assert ctx.L.lucid_set_ray_order(ctx.h, order) == 1
    ctx.close()
