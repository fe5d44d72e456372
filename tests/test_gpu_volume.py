"""GPU parity for participating media (traceVolumeMIS / traceVolumeSimple,
path_tracer.hpp:209-283, 586-835) on the smoke-box workload.

Bars: per-sample radiance on the same seed streams agrees to 1e-3 relative
for at least 97 % of samples (delta tracking amplifies last-ulp differences of
the device log() into occasional different collision points), tile sums to
2 % relative L1, and independent-seed pixel means within 3 sigma (the
fraction and image-mean bounds of tests/sigma.py)."""

import os

import numpy as np
import pytest

import workloads
import sigma
from conftest import make_oracle, make_renderer

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def smoke(tmp_path_factory):
    d = tmp_path_factory.mktemp("smoke")
    return workloads.smoke_box(os.path.join(str(d), "density.bin"), grid_n=24)


def _with_checker_floor(scene):
    """The smoke box with the textured box's checker floor (a node-tree
    material with triangle UVs) in place of its constant floor."""
    floor = workloads.textured_box()["meshes"][0]
    out = dict(scene, meshes=list(scene["meshes"]))
    out["meshes"][0] = dict(out["meshes"][0], material=floor["material"], triangle_uvs=floor["triangle_uvs"])
    return out


@pytest.mark.parametrize("algorithm, textured", [("volume_mis", False), ("volume_simple", False),
                                                 ("volume_mis", True)])
def test_per_sample_radiance_smoke(lucid, smoke, algorithm, textured):
    cam = workloads.SMOKE_BOX_CAMERA
    scene = _with_checker_floor(smoke) if textured else smoke
    r = make_renderer(lucid, scene, cam, algorithm)
    o = make_oracle(scene, cam)
    W = H = 40
    spp = 4
    g = r.debug_render_samples(0, 0, W, H, W, H, spp, 0, 8)
    c = np.stack([o.trace_pixel(algorithm, x, y, W, H, spp, 0, spp, 8) for y in range(H) for x in range(W)])
    close = (np.abs(g - c) <= 1e-3 * (1.0 + np.abs(c))).all(axis=2)
    assert close.mean() >= 0.97, f"only {close.mean():.4f} of samples agree"
    rel = np.abs(g.sum(1) - c.sum(1)).sum() / np.abs(c.sum(1)).sum()
    assert rel < 2e-2, f"relative L1 {rel:.3g}"


def test_smoke_statistics_3sigma(lucid, smoke):
    cam = workloads.SMOKE_BOX_CAMERA
    r = make_renderer(lucid, smoke, cam, "volume_mis")
    o = make_oracle(smoke, cam)
    W = H = 32
    n = 64
    g = r.debug_render_samples(0, 0, W, H, W, H, n, 5000, 8).astype(np.float64)
    _, ex = o.render_tile("volume_mis", 0, 0, W, H, W, H, samples=n, max_depth=8, stats=True)
    st = ex["stats"]
    m_c, v_c = sigma.from_sums(st[:, :3], st[:, 3:], n)
    sigma.check(g.mean(axis=1), g.var(axis=1, ddof=1), n, m_c, v_c, n, "smoke volume_mis")


def test_volume_algorithms_reduce_to_surface_ones_without_media(lucid):
    box = workloads.closed_box()
    cam = workloads.CLOSED_BOX_CAMERA
    a = make_renderer(lucid, box, cam, "volume_mis").debug_render_samples(0, 0, 16, 16, 16, 16, 2, 0, 6)
    b = make_renderer(lucid, box, cam, "mis").debug_render_samples(0, 0, 16, 16, 16, 16, 2, 0, 6)
    np.testing.assert_array_equal(a, b)


def _round_sidecar_to_fp16(scene):
    sg = next(m["material"]["volume"]["smoke_grid"] for m in scene["meshes"] if "volume" in m["material"])
    g = np.fromfile(sg["sidecar_path"], np.float32)
    g.astype(np.float16).astype(np.float32).tofile(sg["sidecar_path"])
    return g


def test_fp16_grid_is_exact_on_fp16_values(lucid, tmp_path):
    """FP16 storage widens exactly: on a grid whose values are FP16 numbers,
    every per-sample radiance equals the FP32-storage render bit for bit,
    and the grid takes half the device bytes."""
    scene = workloads.smoke_box(str(tmp_path / "d16.bin"), grid_n=24)
    _round_sidecar_to_fp16(scene)
    cam = workloads.SMOKE_BOX_CAMERA
    out = {}
    for bits in (32, 16):
        r = make_renderer(lucid, scene, cam, "volume_mis")
        r.set_grid_precision(bits)
        assert r.get_grid_precision() == bits
        out[bits] = r.debug_render_samples(0, 0, 32, 32, 32, 32, 4, 0, 8)
    assert np.array_equal(out[16], out[32])
    import lucid_capi
    ctx = lucid_capi.Context(0)
    nbytes = {}
    for bits in (32, 16):
        ctx.check(ctx.L.lucid_set_grid_format(ctx.h, bits))
        h = ctx.scene(scene)
        nbytes[bits] = ctx.L.lucid_scene_grid_bytes(h)
        ctx.L.lucid_scene_release(h)
    ctx.close()
    assert nbytes[32] == 24 ** 3 * 4 and nbytes[16] == 24 ** 3 * 2


def test_fp16_grid_matches_oracle_on_rounded_grid(lucid, tmp_path):
    """FP16 storage of an FP32 grid computes what the oracle computes on the
    grid rounded to nearest-even FP16 (same bar as the FP32 parity test)."""
    path = str(tmp_path / "d32.bin")
    scene = workloads.smoke_box(path, grid_n=24)
    cam = workloads.SMOKE_BOX_CAMERA
    r = make_renderer(lucid, scene, cam, "volume_mis")  # reads the FP32 sidecar now
    r.set_grid_precision(16)
    W = H = 40
    spp = 4
    g = r.debug_render_samples(0, 0, W, H, W, H, spp, 0, 8)
    g32 = _round_sidecar_to_fp16(scene)
    assert not np.array_equal(g32, np.fromfile(path, np.float32))  # the grid did change
    o = make_oracle(scene, cam)  # the oracle reads the rounded sidecar
    c = np.stack([o.trace_pixel("volume_mis", x, y, W, H, spp, 0, spp, 8) for y in range(H) for x in range(W)])
    close = (np.abs(g - c) <= 1e-3 * (1.0 + np.abs(c))).all(axis=2)
    assert close.mean() >= 0.97, f"only {close.mean():.4f} of samples agree"
