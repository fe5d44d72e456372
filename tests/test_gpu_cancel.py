"""Cancel during a render (PyRenderer::cancel, include/pybind_renderer.hpp:
225-245): a second thread cancels a running render_tile.

* The render returns early with LUCID_ERR_CANCELLED, having completed only
  some of its wavefront batches (the kernels poll the cancel word once per
  bounce, so a cancel lands within one bounce).
* The buffer holds exactly the batches completed before the cancel was
  seen, unflipped (top row first) as the reference's early return leaves it:
  it equals, bit for bit, the row-reversed output of an uncancelled render of
  that many samples (same batch size, hence the same summation order).
* Through the Renderer API the call returns a partial list, is_cancelled()
  stays set until the next render, and the next render is complete.
"""

import ctypes as C
import threading
import time

import numpy as np
import pytest

import lucid_capi
import workloads
from conftest import make_renderer

pytestmark = pytest.mark.gpu

W = H = 256
BATCH = 4        # samples per wavefront batch
TOTAL = 1024     # samples of the cancelled call


def _render(ctx, sh, camd, spp):
    prm = lucid_capi.params(0, 0, W, H, W, H, spp, 0, 8, "mis")
    out = np.zeros(W * H * 4, np.float32)
    rc = ctx.L.lucid_render_tile(ctx.h, sh, C.byref(camd), C.byref(prm), out.ctypes.data_as(C.POINTER(C.c_float)))
    return rc, out.reshape(H, W, 4)


def test_cancel_from_another_thread_returns_completed_batches_unflipped():
    ctx = lucid_capi.Context(0)
    sh = ctx.scene(workloads.closed_box())
    camd = lucid_capi.camera(workloads.CLOSED_BOX_CAMERA)
    ctx.check(ctx.L.lucid_set_wave_paths(ctx.h, W * H * BATCH))
    ctx.check(_render(ctx, sh, camd, BATCH)[0])  # allocations happen here, outside the timing
    t0 = time.perf_counter()
    rc, _ = _render(ctx, sh, camd, TOTAL)
    ctx.check(rc)
    t_full = time.perf_counter() - t0
    got = {}

    def worker():
        got["t0"] = time.perf_counter()
        got["rc"], got["img"] = _render(ctx, sh, camd, TOTAL)
        got["t"] = time.perf_counter() - got["t0"]

    th = threading.Thread(target=worker)
    th.start()
    time.sleep(0.3 * t_full)
    ctx.L.lucid_cancel(ctx.h)
    th.join(timeout=60)
    assert not th.is_alive()
    assert got["rc"] == 3, f"status {got['rc']}"  # LUCID_ERR_CANCELLED
    assert got["t"] < 0.8 * t_full, (got["t"], t_full)
    part = got["img"]
    assert np.all(part[..., 3] == 1.0)
    # find k: the number of batches in the buffer (sums grow with every batch);
    # the sums are taken over arrays of one layout (the partial buffer turned
    # to the row order of a full render), so that equal buffers give equal sums
    # (a float sum depends on the order of its terms)
    def image_sum(a):
        return float(np.ascontiguousarray(a[..., :3], dtype=np.float64).sum())

    total = image_sum(part[::-1])
    assert total > 0.0
    lo, hi = 1, TOTAL // BATCH - 1
    while lo < hi:
        k = (lo + hi) // 2
        rc, img = _render(ctx, sh, camd, BATCH * k)
        ctx.check(rc)
        if image_sum(img) < total:
            lo = k + 1
        else:
            hi = k
    rc, ref = _render(ctx, sh, camd, BATCH * lo)
    ctx.check(rc)
    assert np.array_equal(part, ref[::-1]), f"partial buffer != unflipped render of {BATCH * lo} samples"
    assert not np.array_equal(part, ref)  # the rows really are in the other order
    ctx.L.lucid_scene_release(sh)
    ctx.close()


def test_renderer_cancel_mid_render(lucid):
    scene = workloads.closed_box()
    cam = workloads.CLOSED_BOX_CAMERA
    r = make_renderer(lucid, scene, cam, "mis")
    r.render_tile(0, 0, 512, 512, 512, 512, 512, 0, 8)  # allocations happen here, outside the timing
    t0 = time.perf_counter()
    full = np.asarray(r.render_tile(0, 0, 512, 512, 512, 512, 512, 0, 8), np.float32)
    t_full = time.perf_counter() - t0
    out = {}

    def worker():
        ts = time.perf_counter()
        out["img"] = np.asarray(r.render_tile(0, 0, 512, 512, 512, 512, 512, 0, 8), np.float32)
        out["t"] = time.perf_counter() - ts

    th = threading.Thread(target=worker)
    th.start()
    time.sleep(0.3 * t_full)
    r.cancel()
    th.join(timeout=60)
    assert not th.is_alive() and out["t"] < 0.8 * t_full, (out.get("t"), t_full)
    assert r.is_cancelled()
    assert out["img"].shape == full.shape and out["img"][0::4].sum() < full[0::4].sum()
    again = np.asarray(r.render_tile(0, 0, 512, 512, 512, 512, 512, 0, 8), np.float32)
    assert not r.is_cancelled()
    np.testing.assert_array_equal(again, full)
