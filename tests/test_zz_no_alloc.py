"""The no-allocation render (lucid_render_tile_device_noalloc): its scratch
is a file-scope device array, so it calls no hipMalloc.  Its output equals
lucid_render_tile_device's on the same inputs bit for bit (the film adds
samples in sample order whatever the batch size), on two calls in a row, the
second with a smaller tile; the volume integrator with its extra per-path
factors too; tiles beyond the scratch are refused with a clear error."""

import ctypes as C
import os

import numpy as np
import pytest

import lucid_capi
import workloads

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime liblucid_hip.so itself links (the same runtime
    instance, by path), for the caller-side output buffers."""
    lucid_capi.Context  # the library, and with it this runtime, is loaded
    h = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return h


def _dev(h, nbytes):
    p = C.c_void_p()
    assert h.hipMalloc(C.byref(p), nbytes) == 0
    return p


def _both(ctx, sh, camd, args, spp, algo):
    """(render_tile_device, render_tile_device_noalloc) of the same call, on
    the library's stream, read back to the host."""
    h = _hip()
    prm = lucid_capi.params(*args, spp, 5, 8, algo)
    n = args[2] * args[3] * 4
    a, b = _dev(h, n * 4), _dev(h, n * 4)
    assert h.hipMemset(b, 0xFF, n * 4) == 0  # NaN: the no-allocation call must write every value
    assert h.hipDeviceSynchronize() == 0
    ctx.check(ctx.L.lucid_render_tile_device(ctx.h, sh, C.byref(camd), C.byref(prm), a, None))
    ctx.check(ctx.L.lucid_render_tile_device_noalloc(ctx.h, sh, C.byref(camd), C.byref(prm), b, None))
    ctx.check(ctx.L.lucid_synchronize(ctx.h))
    out = [np.zeros(n, np.float32), np.zeros(n, np.float32)]
    for host, dev in zip(out, (a, b)):
        assert h.hipMemcpy(host.ctypes.data_as(C.c_void_p), dev, n * 4, 2) == 0  # device to host
        h.hipFree(dev)
    return out


def test_noalloc_matches_render_tile_device(tmp_path):
    ctx = lucid_capi.Context(0)
    sh = ctx.scene(workloads.closed_box())
    camd = lucid_capi.camera(workloads.CLOSED_BOX_CAMERA)
    # 640 x 480 at 64 spp: more paths than the scratch holds, so several batches
    for args, spp in (((0, 0, 640, 480, 640, 480), 64), ((37, 21, 150, 90, 640, 480), 6)):
        a, b = _both(ctx, sh, camd, args, spp, "mis")
        assert a.sum() > 0 and np.array_equal(a, b), args
    # the volume integrator keeps three more float4 per path in the scratch
    smoke = workloads.smoke_box(os.path.join(str(tmp_path), "d.bin"), grid_n=16)
    sv = ctx.scene(smoke)
    camv = lucid_capi.camera(workloads.SMOKE_BOX_CAMERA)
    a, b = _both(ctx, sv, camv, (0, 0, 96, 64, 96, 64), 4, "volume_mis")
    assert a.sum() > 0 and np.array_equal(a, b)
    # a tile beyond the scratch is refused with the limit in the message
    prm = lucid_capi.params(0, 0, 2048, 2048, 2048, 2048, 1, 0, 8, "mis")
    h = _hip()
    out = _dev(h, 64)
    rc = ctx.L.lucid_render_tile_device_noalloc(ctx.h, sh, C.byref(camd), C.byref(prm), out, None)
    h.hipFree(out)
    assert rc == 1 and b"exceeds the no-allocation scratch" in ctx.L.lucid_last_error(ctx.h)
    ctx.L.lucid_scene_release(sv)
    ctx.L.lucid_scene_release(sh)
    ctx.close()
