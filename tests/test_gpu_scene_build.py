"""The device scene build (scene_build.hip) against the host packer
(scene_pack.cpp, std::nth_element): for every workload shape -- grid walls
with whole rows of equal centroids, icospheres with smooth normals, node
materials with UVs, glass and emissive meshes, tiny meshes (<= 4 triangles,
a leaf as the mesh root), empty meshes, the 4M-triangle volume scene -- the
node array, the triangle records, face and vertex normals, UVs and the
stack depth are byte-identical, and renders through both scenes are
bit-identical.
"""

import ctypes as C
import time

import numpy as np
import pytest

import lucid_capi
import oracle
import workloads

pytestmark = pytest.mark.gpu


def _geometry(ctx, sh):
    L = ctx.L
    nn, depth, dev = C.c_int64(), C.c_int32(), C.c_int32()
    ctx.check(L.lucid_scene_read_geometry(ctx.h, sh, None, None, None, None, None, C.byref(nn), C.byref(depth),
                                          C.byref(dev)))
    nt = L.lucid_scene_num_triangles(sh)
    nodes = np.zeros(nn.value * 16, np.int32)
    tri, fn, vn, uv = (np.zeros(nt * k, np.float32) for k in (12, 4, 12, 6))
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    ctx.check(L.lucid_scene_read_geometry(ctx.h, sh, nodes.ctypes.data_as(C.c_void_p), fp(tri), fp(fn), fp(vn),
                                          fp(uv), C.byref(nn), C.byref(depth), C.byref(dev)))
    return {"nodes": nodes, "tri": tri, "fn": fn, "vn": vn, "uv": uv, "depth": depth.value, "device": dev.value}


def _both(ctx, scene):
    out = {}
    for mode in (0, 1):
        ctx.check(ctx.L.lucid_set_scene_build(ctx.h, mode))
        t0 = time.perf_counter()
        sh = ctx.scene(scene)
        out[mode] = (sh, time.perf_counter() - t0)
    ctx.check(ctx.L.lucid_set_scene_build(ctx.h, 1))
    return out


def _tiny_and_empty():
    rng = np.random.default_rng(3)
    meshes = []
    for n in (1, 2, 3, 4, 5, 0, 9, 4, 17):
        v = rng.uniform(-1, 1, (3 * max(n, 1), 3))
        t = np.arange(3 * n).reshape(n, 3) if n else np.zeros((0, 3), np.int64)
        meshes.append(workloads._mesh(v[: 3 * n] if n else np.zeros((0, 3)), t,
                                      {"base_color": [0.5, 0.5, 0.5], "roughness": 0.5}))
    # duplicated triangles: every centroid tied
    v = np.tile(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float64), (40, 1))
    meshes.append(workloads._mesh(v, np.arange(120).reshape(40, 3), {"base_color": [0.8, 0.2, 0.2]}))
    return {"meshes": meshes, "native_lights": [], "environment": {"color": [0.3, 0.3, 0.3], "strength": 1.0}}


SCENES = {
    "closed_box": lambda tmp: workloads.closed_box(),
    "textured_box": lambda tmp: workloads.textured_box(panel=True),
    "glass_lights_box": lambda tmp: workloads.glass_lights_box(),
    "tiny_and_empty": lambda tmp: _tiny_and_empty(),
    "occluded_room": lambda tmp: workloads.occluded_room(),
    "icospheres_1m": lambda tmp: workloads.icospheres(seed=2, count=200, level=4),
    "volume_4m": lambda tmp: workloads.volume_scene(str(tmp / "smoke.bin")),
}


@pytest.mark.timeout(900)
@pytest.mark.parametrize("name", list(SCENES))
def test_device_build_equals_host_build(name, tmp_path):
    scene = SCENES[name](tmp_path)
    ctx = lucid_capi.Context(0)
    both = _both(ctx, scene)
    (hs, th), (ds, td) = both[0], both[1]
    h, d = _geometry(ctx, hs), _geometry(ctx, ds)
    assert h["device"] == 0 and d["device"] == 1
    assert h["depth"] == d["depth"]
    for k in ("nodes", "tri", "fn", "vn", "uv"):
        assert h[k].shape == d[k].shape, k
        assert h[k].tobytes() == d[k].tobytes(), f"{k} differs"
    ntri = ctx.L.lucid_scene_num_triangles(ds)
    print(f"{name}: {ntri} triangles, {len(h['nodes']) // 16} nodes, host build {th:.3f} s, device build {td:.3f} s")
    for sh in (hs, ds):
        ctx.L.lucid_scene_release(sh)
    ctx.close()


@pytest.mark.parametrize("name,cam,algo", [("textured_box", workloads.CLOSED_BOX_CAMERA, "mis"),
                                           ("glass_lights_box", workloads.CLOSED_BOX_CAMERA, "nee"),
                                           ("occluded_room", workloads.OCCLUDED_ROOM_CAMERA, "mis")])
def test_renders_through_both_builds_are_identical(name, cam, algo, tmp_path):
    scene = SCENES[name](tmp_path)
    ctx = lucid_capi.Context(0)
    both = _both(ctx, scene)
    camd = lucid_capi.camera(cam)
    imgs = []
    for mode in (0, 1):
        out = np.zeros(96 * 64 * 4, np.float32)
        prm = lucid_capi.params(0, 0, 96, 64, 96, 64, 8, 0, 8, algo)
        ctx.check(ctx.L.lucid_render_tile(ctx.h, both[mode][0], C.byref(camd), C.byref(prm),
                                          out.ctypes.data_as(C.POINTER(C.c_float))))
        imgs.append(out)
    assert np.isfinite(imgs[1]).all() and imgs[1].sum() > 0
    np.testing.assert_array_equal(imgs[0], imgs[1])
    for m in (0, 1):
        ctx.L.lucid_scene_release(both[m][0])
    ctx.close()
