"""Small hand-built scenes used by the known-answer and convergence tests.

They follow the reference's test fixtures (LucidRenderer/cpp_renderer/test/
test_utils.hpp:135-330 and test_gpu_path_tracer.cpp:281-349), written in the
scene-JSON schema.
"""

import math


def _quad(p0, p1, p2, p3, material):
    return {"vertices": [p0, p1, p2, p3], "triangles": [[0, 1, 2], [0, 2, 3]], "material": material}


def _mat(albedo, metallic=0.0, roughness=0.5, emission=(0.0, 0.0, 0.0), transmission=0.0, ior=1.45):
    return {"base_color": list(albedo), "metallic": metallic, "roughness": roughness, "emission": list(emission),
            "transmission": transmission, "ior": ior}


def cornell_box(light_radiance=15.0):
    """5 walls + an emissive quad under the ceiling (test_utils.hpp:159-272)."""
    white = _mat((0.73, 0.73, 0.73))
    meshes = [
        _quad([-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, 1], white),             # floor
        {"vertices": [[-1, 2, -1], [1, 2, -1], [1, 2, 1], [-1, 2, 1]], "triangles": [[0, 2, 1], [0, 3, 2]],
         "material": white},                                                      # ceiling
        _quad([-1, 0, -1], [1, 0, -1], [1, 2, -1], [-1, 2, -1], white),           # back
        _quad([-1, 0, -1], [-1, 0, 1], [-1, 2, 1], [-1, 2, -1], _mat((0.65, 0.05, 0.05))),  # left (red)
        _quad([1, 0, -1], [1, 2, -1], [1, 2, 1], [1, 0, 1], _mat((0.12, 0.45, 0.15))),      # right (green)
        _quad([-0.25, 1.99, -0.25], [0.25, 1.99, -0.25], [0.25, 1.99, 0.25], [-0.25, 1.99, 0.25],
              _mat((0.0, 0.0, 0.0), emission=(light_radiance,) * 3)),
    ]
    return {"meshes": meshes, "environment": {"color": [0.0, 0.0, 0.0], "strength": 1.0}}


def point_light_quad(energy=10.0, albedo=0.7):
    """Lambert quad at z = -1 under a delta point light at the origin
    (test_gpu_path_tracer.cpp:281-349)."""
    quad = _quad([-2, -2, -1], [2, -2, -1], [2, 2, -1], [-2, 2, -1], _mat((albedo,) * 3, roughness=1.0))
    light = {"type": "POINT", "position": [0.0, 0.0, 0.0], "color": [1.0, 1.0, 1.0], "energy": energy,
             "radius": 0.0}
    return {"meshes": [quad], "native_lights": [light], "environment": {"color": [0, 0, 0], "strength": 1.0}}


def furnace(albedo=0.5):
    """Octahedron in a uniform white environment (test_path_tracer.cpp:314-361)."""
    v = [[0, 1, 0], [0, -1, 0], [1, 0, 0], [-1, 0, 0], [0, 0, 1], [0, 0, -1]]
    t = [[0, 2, 4], [0, 4, 3], [0, 3, 5], [0, 5, 2], [1, 4, 2], [1, 3, 4], [1, 5, 3], [1, 2, 5]]
    return {"meshes": [{"vertices": v, "triangles": t, "material": _mat((albedo,) * 3, roughness=1.0)}],
            "environment": {"color": [1.0, 1.0, 1.0], "strength": 1.0}}


CAM_FRONT = {"pos": (0.0, 1.0, 3.5), "dir": (0.0, 0.0, -1.0), "up": (0.0, 1.0, 0.0), "fov": 40.0}
CAM_DOWN_Z = {"pos": (0.0, 0.0, 5.0), "dir": (0.0, 0.0, -1.0), "up": (0.0, 1.0, 0.0), "fov": 30.0}
PI = math.pi
