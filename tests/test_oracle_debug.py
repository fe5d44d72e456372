"""Pins the oracle's debug passes (render_debug, pybind_renderer.hpp:371-439;
traceNormal / traceAlbedo / traceEmission / traceVolume, renderer.hpp:
58-190) by known answers:

* normal: an axis-aligned floor shows (0.5, 1, 0.5); a miss shows 0.5 grey;
  a smooth icosphere's normals are unit length, face the camera and are
  (0, 0, 1) at its front pole;
* albedo / emission: the JSON base colour and emission of the hit mesh, zero
  on a miss;
* volume: a cube of homogeneous medium of density s seen through its centre
  gives colour x (1 - exp(-s L)) with L the cube's thickness less the 1e-4
  entry offset; a ray leaving no exit hit gives colour x (1 - exp(-s));
* an unknown mode paints every pixel magenta; rows are bottom first.
"""

import math

import numpy as np

import oracle

DOWN = {"pos": (0.0, 3.0, 0.0), "dir": (0.0, -1.0, 0.0), "up": (0.0, 0.0, -1.0), "fov": 20.0}
FRONT = {"pos": (0.0, 0.0, 4.0), "dir": (0.0, 0.0, -1.0), "up": (0.0, 1.0, 0.0), "fov": 30.0}


def _scene(meshes):
    return {"meshes": meshes, "environment": {"color": [0.1, 0.1, 0.1], "strength": 1.0}}


def _o(scene, cam):
    o = oracle.OracleScene(scene)
    o.set_camera(cam["pos"], cam["dir"], cam["up"], cam["fov"])
    return o


def _floor(color, emission=(0, 0, 0), half=0.5):
    h = half
    return {"vertices": [[-h, 0, -h], [h, 0, -h], [h, 0, h], [-h, 0, h]], "triangles": [[0, 2, 1], [0, 3, 2]],
            "material": {"base_color": list(color), "emission": list(emission), "roughness": 1.0}}


def test_normal_albedo_emission_known_answers():
    o = _o(_scene([_floor((0.25, 0.5, 0.75), emission=(2.0, 3.0, 4.0))]), DOWN)
    W = H = 32
    n = o.render_debug("normal", 0, 0, W, H, W, H).reshape(H, W, 4)
    a = o.render_debug("albedo", 0, 0, W, H, W, H).reshape(H, W, 4)
    e = o.render_debug("emission", 0, 0, W, H, W, H).reshape(H, W, 4)
    c = (H // 2, W // 2)  # the centre sees the floor; the 1 m floor fills ~1/3 of a 20 degree view at 3 m
    np.testing.assert_array_equal(n[c], np.float32([0.5, 1.0, 0.5, 1.0]))
    np.testing.assert_array_equal(a[c], np.float32([0.25, 0.5, 0.75, 1.0]))
    np.testing.assert_array_equal(e[c], np.float32([2.0, 3.0, 4.0, 1.0]))
    corner = (0, 0)  # misses
    np.testing.assert_array_equal(n[corner], np.float32([0.5, 0.5, 0.5, 1.0]))
    np.testing.assert_array_equal(a[corner], np.float32([0.0, 0.0, 0.0, 1.0]))
    hit = (n[..., 1] == 1.0)
    assert 0.05 < hit.mean() < 0.9
    np.testing.assert_array_equal(hit, (a[..., 0] == np.float32(0.25)))


def test_sphere_normals_equal_hit_points():
    import workloads

    v, t = workloads._icosphere(4)
    sphere = {"vertices": v.tolist(), "triangles": t.tolist(), "triangle_normals": v[t].tolist(), "smooth": True,
              "material": {"base_color": [0.5, 0.5, 0.5]}}
    o = _o(_scene([sphere]), FRONT)
    W = H = 48
    n = o.render_debug("normal", 0, 0, W, H, W, H).reshape(H, W, 4)[..., :3].astype(np.float64) * 2 - 1
    m = np.linalg.norm(n, axis=2) > 0.5
    assert m.mean() > 0.2
    # smooth normals of a level-4 icosphere: unit length, facing the camera
    assert np.allclose(np.linalg.norm(n[m], axis=1), 1.0, atol=1e-5)
    assert (n[m][:, 2] > 0).all()
    # the centre pixel looks straight at the sphere's front pole
    assert np.allclose(n[H // 2, W // 2], [0, 0, 1], atol=0.05)


def _medium_cube(density, color, half=0.5):
    h = half
    corners = [[sx * h, sy * h, sz * h] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)]
    faces = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    tris = []
    for a, b, c, d in faces:
        tris += [[a, b, c], [a, c, d]]
    return {"vertices": corners, "triangles": tris,
            "material": {"base_color": [0.8, 0.8, 0.8], "volume": {"density": density, "color": list(color)}}}


def test_volume_thickness_closed_form():
    s, col = 2.0, (0.9, 0.5, 0.3)
    o = _o(_scene([_medium_cube(s, col)]), FRONT)
    W = H = 33  # odd: the centre pixel's ray is the camera axis
    v = o.render_debug("volume", 0, 0, W, H, W, H).reshape(H, W, 4)
    L = 1.0 - 1e-4  # the cube's thickness on the axis, less the entry offset
    want = np.float64(col) * (1.0 - math.exp(-s * L))
    np.testing.assert_allclose(v[H // 2, W // 2, :3], want, rtol=2e-5)
    np.testing.assert_array_equal(v[0, 0, :3], 0.0)  # a miss
    # a medium on a single quad: no exit hit, colour x (1 - exp(-s))
    q = _floor((0.5, 0.5, 0.5))
    q["material"]["volume"] = {"density": s, "color": list(col)}
    v2 = _o(_scene([q]), DOWN).render_debug("volume", 0, 0, W, H, W, H).reshape(H, W, 4)
    np.testing.assert_allclose(v2[H // 2, W // 2, :3], np.float64(col) * (1.0 - math.exp(-s)), rtol=2e-5)


def test_unknown_mode_is_magenta_everywhere():
    o = _o(_scene([_floor((0.3, 0.3, 0.3))]), DOWN)
    x = o.render_debug("depth", 0, 0, 8, 6, 8, 6).reshape(-1, 4)
    np.testing.assert_array_equal(x, np.tile(np.float32([1, 0, 1, 1]), (48, 1)))


def test_rows_bottom_first_and_tile_offsets():
    o = _o(_scene([_floor((0.3, 0.6, 0.9))]), DOWN)
    full = o.render_debug("normal", 0, 0, 40, 30, 40, 30).reshape(30, 40, 4)
    tile = o.render_debug("normal", 8, 5, 16, 12, 40, 30).reshape(12, 16, 4)
    top_first_full = full[::-1]
    np.testing.assert_array_equal(tile[::-1], top_first_full[5:17, 8:24])
