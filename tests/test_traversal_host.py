"""The device BVH traversal code (traverse<ANY> in csrc/pathtrace.hip), run on
the CPU through lucid_debug_host_trace, against the oracle's intersectScene.

The host build of the traversal checks every node and triangle index before
using it, so a traversal that would read outside its arrays on the GPU fails
here on the CPU.  Closest hits must match the oracle bit for bit (mesh,
triangle, u, v, t), including exact ties; occlusion must match the oracle's
"any hit in [tmin, tmax]"."""

import numpy as np
import pytest

import lucid_capi
import oracle
import workloads


def _rays(rng, n, lo, hi, spread=1.0):
    o = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    tgt = rng.uniform(lo, hi, size=(n, 3)) * spread
    d = tgt - o
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    return o, d


def _camera_rays(cam, W, H):
    pos = np.float32(cam["pos"])
    f = np.float32(cam["dir"]) / np.linalg.norm(cam["dir"])
    up = np.float32(cam["up"])
    r = np.cross(f, up)
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    s = np.tan(np.radians(cam["fov"]) / 2)
    ys, xs = np.mgrid[0:H, 0:W]
    nx = (2 * (xs + 0.5) / W - 1) * (W / H)
    ny = 1 - 2 * (ys + 0.5) / H
    d = f[None] + r[None] * (nx.reshape(-1, 1) * s) + u[None] * (ny.reshape(-1, 1) * s)
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    return np.repeat(pos[None], len(d), 0).astype(np.float32), d


def _compare(scene, o, d, tmin, tmax):
    h = lucid_capi.host_trace(scene, o, d, tmin, tmax)
    c = oracle.OracleScene(scene).intersect(o, d, tmin, tmax)
    for k in ("mesh", "tri", "u", "v", "t"):
        same = h[k].view(np.int32) == c[k].view(np.int32)
        assert same.all(), f"{k}: {np.count_nonzero(~same)} of {same.size} differ"
    occ = lucid_capi.host_trace(scene, o, d, tmin, tmax, any_hit=True)
    assert np.array_equal(occ["mesh"] >= 0, c["mesh"] >= 0)
    return c


def test_closed_box_camera_and_random_rays():
    scene = workloads.closed_box()
    o, d = _camera_rays(workloads.CLOSED_BOX_CAMERA, 96, 96)
    c = _compare(scene, o, d, 0.0, 1e30)
    assert (c["mesh"] >= 0).mean() > 0.5
    rng = np.random.default_rng(1)
    o, d = _rays(rng, 20000, -1.2, 2.2)
    _compare(scene, o, d, 1e-4, 1e30)
    # shadow-like segments of random length
    _compare(scene, o, d, 1e-4, rng.uniform(0.01, 3.0, size=len(o)).astype(np.float32))


def test_triangle_soup_many_meshes():
    rng = np.random.default_rng(5)
    meshes = []
    for m in range(9):
        n = int(rng.integers(1, 300))
        c = rng.uniform(-1, 1, size=(n, 1, 3))
        v = (c + rng.normal(size=(n, 3, 3)) * 0.2).reshape(-1, 3)
        meshes.append({"vertices": v.astype(np.float32).tolist(), "triangles": np.arange(3 * n).reshape(n, 3).tolist()})
    scene = {"meshes": meshes}
    o, d = _rays(rng, 20000, -2.5, 2.5)
    _compare(scene, o, d, 1e-4, 1e30)
    _compare(scene, o, d, 1e-4, rng.uniform(0.05, 4.0, size=len(o)).astype(np.float32))


def test_exact_ties_and_axis_aligned_boxes():
    # duplicated triangles (exact t ties across meshes and within a leaf) and
    # rays parallel to the axes through flat boxes
    t = [[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]]
    tris = [[0, 1, 2], [1, 3, 2], [0, 1, 2], [1, 3, 2], [0, 1, 2], [0, 1, 2]]
    scene = {"meshes": [{"vertices": t, "triangles": tris}, {"vertices": t, "triangles": tris[:2]},
                        {"vertices": t, "triangles": tris}]}
    rng = np.random.default_rng(2)
    n = 4000
    xy = rng.uniform(-0.1, 1.1, size=(n, 2))
    o = np.c_[xy, np.full(n, 1.0)].astype(np.float32)
    d = np.tile(np.float32([0, 0, -1]), (n, 1))
    c = _compare(scene, o, d, 1e-4, 1e30)
    assert (c["mesh"] == 2).any()
    o2 = np.c_[np.full(n, -1.0), xy[:, 1], rng.uniform(-0.1, 0.1, n)].astype(np.float32)
    _compare(scene, o2, np.tile(np.float32([1, 0, 0]), (n, 1)), 1e-4, 1e30)


def test_icospheres_subset():
    scene = workloads.icospheres(seed=9, count=12, level=3)
    cam = workloads.ICOSPHERES_CAMERA
    o, d = _camera_rays(cam, 80, 80)
    _compare(scene, o, d, 0.0, 1e30)
    rng = np.random.default_rng(3)
    o, d = _rays(rng, 10000, -8, 8)
    _compare(scene, o, d, 1e-4, 1e30)


def test_one_voxel_grid_axis_is_refused(tmp_path):
    # the trilinear lookup's cell clamp [0, n-2] is empty for a 1-voxel axis
    scene = workloads.smoke_box(str(tmp_path / "density.bin"), grid_n=8)
    for m in scene["meshes"]:
        sg = (m.get("material") or {}).get("volume", {}).get("smoke_grid")
        if sg:
            sg["dims"] = [8, 8, 1]
    o = np.zeros((1, 3), np.float32)
    d = np.float32([[0, 0, 1]])
    with pytest.raises(RuntimeError, match="LUCID_ERR_INVALID"):
        lucid_capi.host_trace(scene, o, d, 1e-4, 1e30)


def test_occluded_room_rays():
    """The 1080p workload's scene: camera rays and random rays inside the
    room against the oracle."""
    scene = workloads.occluded_room()
    o, d = _camera_rays(workloads.OCCLUDED_ROOM_CAMERA, 64, 36)
    _compare(scene, o, d, 0.0, 1e30)
    rng = np.random.default_rng(4)
    o, d = _rays(rng, 4000, -1.5, 1.5)
    o[:, 1] += 1.5
    _compare(scene, o, d, 1e-4, 1e30)
