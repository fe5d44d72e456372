"""Procedural scenes for the path-tracing workloads, in Lucid's scene-JSON schema.

The schema is the one ``load_scene_json`` reads (reference:
LucidRenderer/cpp_renderer/include/pybind_renderer.hpp:1258-1746): ``meshes``
with ``vertices``, ``triangles``, optional ``triangle_normals`` / ``smooth``
and a constant ``material``; ``native_lights``; ``environment``.

Workloads (all seeded; the seed is part of the returned spec):

* ``closed_box``    -- 5 axis-aligned walls + 2 boxes rotated about y (34
  triangles), one 0.5 m x 0.5 m square area light of radiance 15, diffuse
  Principled materials with base colours U[0.2, 0.8].  256^2, 16 spp, depth 8.
* ``icospheres``    -- 200 seed-placed icospheres (subdivision 4, 5120
  triangles each = 1,024,000 triangles) with seeded radial displacement and
  smooth normals; per object roughness U[0.05, 1], metallic Bernoulli(0.3),
  IOR U[1.3, 1.7]; one point, one sun and 4 area lights.  1024^2, 64 spp,
  depth 12.
* ``occluded_room`` -- about 300k triangles: a glossy room (roughness
  U[0.1, 0.4]) cut by a partition whose only opening is a 5 cm slit, with an
  area light hidden behind it.  1920 x 1080, up to 1024 spp, depth 12; the
  scene of the time-to-relMSE measurement (guided vs plain).
"""

from __future__ import annotations

import json
import math

import numpy as np


def _quad(p0, p1, p2, p3):
    """Two triangles (p0,p1,p2), (p0,p2,p3)."""
    v = np.array([p0, p1, p2, p3], dtype=np.float64)
    t = np.array([[0, 1, 2], [0, 2, 3]], dtype=np.int64)
    return v, t


def _box(center, half, angle_y):
    """Axis-aligned box of half extents `half`, rotated by angle_y about +y."""
    hx, hy, hz = half
    corners = np.array([[sx * hx, sy * hy, sz * hz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)])
    c, s = math.cos(angle_y), math.sin(angle_y)
    rot = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    v = corners @ rot.T + np.asarray(center)
    # corner index = 4*ix + 2*iy + iz
    faces = [
        (0, 1, 3, 2),  # -x
        (4, 6, 7, 5),  # +x
        (0, 4, 5, 1),  # -y
        (2, 3, 7, 6),  # +y
        (0, 2, 6, 4),  # -z
        (1, 5, 7, 3),  # +z
    ]
    t = []
    for a, b, cc, d in faces:
        t.append([a, b, cc])
        t.append([a, cc, d])
    return v, np.array(t, dtype=np.int64)


def _diffuse(color):
    return {"base_color": [float(x) for x in color], "metallic": 0.0, "roughness": 1.0,
            "emission": [0.0, 0.0, 0.0], "transmission": 0.0, "ior": 1.45}


def _mesh(v, t, material, normals=None, smooth=False):
    m = {"vertices": np.asarray(v, np.float32).tolist(), "triangles": np.asarray(t, np.int64).tolist(),
         "material": material}
    if normals is not None:
        m["triangle_normals"] = np.asarray(normals, np.float32).reshape(-1, 3, 3).tolist()
        m["smooth"] = bool(smooth)
    return m


def closed_box(seed: int = 1) -> dict:
    """34-triangle closed box lit by a 0.25 m^2 area light of radiance 15."""
    rng = np.random.default_rng(seed)
    meshes = []
    walls = [
        _quad([-1, 0, -1], [-1, 0, 1], [1, 0, 1], [1, 0, -1]),      # floor  (y = 0)
        _quad([-1, 2, -1], [1, 2, -1], [1, 2, 1], [-1, 2, 1]),      # ceiling (y = 2)
        _quad([-1, 0, -1], [1, 0, -1], [1, 2, -1], [-1, 2, -1]),    # back (z = -1)
        _quad([-1, 0, -1], [-1, 2, -1], [-1, 2, 1], [-1, 0, 1]),    # left (x = -1)
        _quad([1, 0, -1], [1, 0, 1], [1, 2, 1], [1, 2, -1]),        # right (x = +1)
    ]
    for v, t in walls:
        meshes.append(_mesh(v, t, _diffuse(rng.uniform(0.2, 0.8, 3))))
    boxes = [
        _box([-0.4, 0.6, -0.3], [0.3, 0.6, 0.3], float(rng.uniform(0.1, 0.6))),
        _box([0.45, 0.3, 0.35], [0.3, 0.3, 0.3], float(-rng.uniform(0.1, 0.6))),
    ]
    for v, t in boxes:
        meshes.append(_mesh(v, t, _diffuse(rng.uniform(0.2, 0.8, 3))))
    radiance, size = 15.0, 0.5
    light = {"type": "AREA", "name": "Light", "position": [0.0, 1.98, 0.0], "direction": [0.0, -1.0, 0.0],
             "right": [1.0, 0.0, 0.0], "up": [0.0, 0.0, 1.0], "shape": "SQUARE", "size": size,
             "color": [1.0, 1.0, 1.0], "energy": radiance * math.pi * size * size}
    return {"meshes": meshes, "native_lights": [light],
            "environment": {"color": [0.05, 0.05, 0.05], "strength": 1.0}}


CLOSED_BOX_CAMERA = {"pos": (0.0, 1.0, 3.6), "dir": (0.0, 0.0, -1.0), "up": (0.0, 1.0, 0.0), "fov": 40.0}


def _sphere_mesh(center, radius, material, level=3):
    """Icosphere of `radius` at `center` with exact sphere normals per corner
    (smooth shading)."""
    if level not in _ICO_CACHE:
        _ICO_CACHE[level] = _icosphere(level)
    v, t = _ICO_CACHE[level]
    return _mesh(v * radius + np.asarray(center, np.float64), t, material, v[t], smooth=True)


def glass_lights_box(seed: int = 3) -> dict:
    """The closed box's walls with three spheres -- smooth glass
    (transmission 1, IOR 1.5, roughness 0), rough glass (roughness 0.3) and a
    partly transmissive sphere (transmission 0.4) -- lit by a sphere point
    light (radius 0.08), a DISK and an ELLIPSE area light under the ceiling
    and a spot light with a blended edge.  Covers the glass branch of
    sampleBSDF and the light shapes the closed box does not have."""
    rng = np.random.default_rng(seed)
    walls = closed_box(seed)["meshes"][:5]
    glass = {"base_color": [1.0, 1.0, 1.0], "metallic": 0.0, "roughness": 0.0, "emission": [0.0, 0.0, 0.0],
             "transmission": 1.0, "ior": 1.5}
    rough = dict(glass, roughness=0.3)
    mixed = {"base_color": rng.uniform(0.3, 0.8, 3).tolist(), "metallic": 0.0, "roughness": 0.2,
             "emission": [0.0, 0.0, 0.0], "transmission": 0.4, "ior": 1.45}
    meshes = walls + [_sphere_mesh([-0.45, 0.38, -0.2], 0.36, glass),
                      _sphere_mesh([0.45, 0.32, 0.25], 0.3, rough),
                      _sphere_mesh([0.05, 1.15, -0.45], 0.25, mixed)]
    down = {"direction": [0.0, -1.0, 0.0], "right": [1.0, 0.0, 0.0], "up": [0.0, 0.0, 1.0]}
    lights = [
        {"type": "POINT", "name": "Bulb", "position": [0.55, 1.55, 0.45], "color": [1.0, 0.9, 0.8], "energy": 25.0,
         "radius": 0.08},
        dict(down, type="AREA", name="Disk", shape="DISK", position=[-0.45, 1.98, 0.2], size=0.4, size_y=0.4,
             color=[0.9, 0.95, 1.0], energy=6.0),
        dict(down, type="AREA", name="Ellipse", shape="ELLIPSE", position=[0.35, 1.98, -0.5], size=0.5, size_y=0.25,
             color=[1.0, 1.0, 0.9], energy=5.0),
        {"type": "SPOT", "name": "Spot", "position": [-0.75, 1.9, 0.8], "direction": [0.45, -1.0, -0.55],
         "color": [1.0, 0.95, 0.85], "energy": 40.0, "spot_size": 0.9, "spot_blend": 0.3},
    ]
    return {"meshes": meshes, "native_lights": lights, "environment": {"color": [0.03, 0.03, 0.04], "strength": 1.0}}


def _icosphere(level: int):
    """Unit icosphere: (vertices (n,3), triangles (m,3)), 20*4^level triangles."""
    p = (1.0 + math.sqrt(5.0)) / 2.0
    v = [[-1, p, 0], [1, p, 0], [-1, -p, 0], [1, -p, 0], [0, -1, p], [0, 1, p], [0, -1, -p], [0, 1, -p],
         [p, 0, -1], [p, 0, 1], [-p, 0, -1], [-p, 0, 1]]
    v = [list(np.array(x, float) / np.linalg.norm(x)) for x in v]
    f = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
         [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11],
         [6, 2, 10], [8, 6, 7], [9, 8, 1]]
    for _ in range(level):
        cache = {}

        def mid(a, b):
            k = (a, b) if a < b else (b, a)
            if k not in cache:
                m = (np.array(v[a]) + np.array(v[b])) / 2.0
                v.append(list(m / np.linalg.norm(m)))
                cache[k] = len(v) - 1
            return cache[k]

        nf = []
        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        f = nf
    return np.array(v, np.float64), np.array(f, np.int64)


_ICO_CACHE = {}


def icospheres(seed: int = 2, count: int = 200, level: int = 4, smooth: bool = True) -> dict:
    """`count` displaced icospheres with per-object Principled materials."""
    rng = np.random.default_rng(seed)
    if level not in _ICO_CACHE:
        _ICO_CACHE[level] = _icosphere(level)
    base_v, base_t = _ICO_CACHE[level]
    meshes = []
    for i in range(count):
        radius = rng.uniform(0.35, 0.9)
        center = np.array([rng.uniform(-9, 9), rng.uniform(0.4, 5.0), rng.uniform(-9, 9)])
        # seeded radial displacement: a few random low-frequency lobes
        k = rng.normal(size=(4, 3))
        k /= np.linalg.norm(k, axis=1, keepdims=True)
        amp = rng.uniform(0.02, 0.12, size=4)
        disp = 1.0 + (amp[None, :] * np.sin(3.0 * base_v @ k.T + rng.uniform(0, 6.28, size=4))).sum(1)
        pos = base_v * (radius * disp)[:, None] + center
        # smooth vertex normals: area-weighted face normals
        a, b, c = pos[base_t[:, 0]], pos[base_t[:, 1]], pos[base_t[:, 2]]
        fn = np.cross(b - a, c - a)
        vn = np.zeros_like(pos)
        for j in range(3):
            np.add.at(vn, base_t[:, j], fn)
        vn /= np.linalg.norm(vn, axis=1, keepdims=True)
        normals = vn[base_t] if smooth else None
        metallic = 1.0 if rng.uniform() < 0.3 else 0.0
        mat = {"base_color": rng.uniform(0.15, 0.9, 3).tolist(), "metallic": metallic,
               "roughness": float(rng.uniform(0.05, 1.0)), "emission": [0.0, 0.0, 0.0],
               "transmission": 0.0, "ior": float(rng.uniform(1.3, 1.7))}
        meshes.append(_mesh(pos, base_t, mat, normals, smooth))
    lights = [
        {"type": "POINT", "name": "Point", "position": [0.0, 9.0, 2.0], "color": [1.0, 0.95, 0.9],
         "energy": 2500.0, "radius": 0.0},
        {"type": "SUN", "name": "Sun", "position": [0.0, 20.0, 0.0], "direction": [-0.3, -1.0, -0.4],
         "color": [1.0, 1.0, 1.0], "energy": 2.0},
    ]
    for j, (x, z) in enumerate([(-6, -6), (6, -6), (-6, 6), (6, 6)]):
        lights.append({"type": "AREA", "name": f"Area{j}", "position": [float(x), 8.0, float(z)],
                       "direction": [0.0, -1.0, 0.0], "right": [1.0, 0.0, 0.0], "up": [0.0, 0.0, 1.0],
                       "shape": "RECTANGLE", "size": 2.0, "size_y": 1.0, "color": [1.0, 1.0, 1.0],
                       "energy": 600.0})
    return {"meshes": meshes, "native_lights": lights,
            "environment": {"color": [0.05, 0.06, 0.08], "strength": 1.0}}


ICOSPHERES_CAMERA = {"pos": (0.0, 6.0, 20.0), "dir": (0.0, -0.25, -1.0), "up": (0.0, 1.0, 0.0), "fov": 55.0}


def fbm_grid(n: int, seed: int = 3, octaves: int = 4) -> np.ndarray:
    """Seeded fractal value noise on an n^3 grid (x fastest in memory), shaped
    into a soft blob, values in [0, 1]."""
    from scipy.ndimage import zoom

    rng = np.random.default_rng(seed)
    acc = np.zeros((n, n, n), np.float64)
    amp, norm = 1.0, 0.0
    for k in range(octaves):
        r = 2 ** (k + 2) + 1
        lat = rng.random((r, r, r))
        acc += amp * zoom(lat, n / r, order=1, mode="nearest")[:n, :n, :n]
        norm += amp
        amp *= 0.5
    acc /= norm
    c = (np.arange(n) + 0.5) / n - 0.5
    zz, yy, xx = np.meshgrid(c, c, c, indexing="ij")
    blob = np.clip(1.0 - np.sqrt(xx * xx + yy * yy + zz * zz) / 0.5, 0.0, 1.0)
    g = np.clip((acc - 0.35) * 2.5, 0.0, 1.0) * blob
    return np.ascontiguousarray(g, np.float32)      # index [z, y, x] -> x fastest


def boundary_node_tree():
    """Material Output with Volume linked and Surface unlinked: the mesh only
    bounds a medium (pybind_renderer.hpp:1386-1398, 1487-1503)."""
    return {"nodes": [{"name": "Material Output", "type": "ShaderNodeOutputMaterial",
                       "inputs": [{"name": "Surface"}, {"name": "Volume"}]},
                      {"name": "Principled Volume", "type": "ShaderNodePrincipledVolume", "inputs": []}],
            "links": [{"from_node": "Principled Volume", "from_socket": "Volume",
                       "to_node": "Material Output", "to_socket": "Volume"}]}


# ---- node-tree materials (the scene-JSON node_tree schema of
# scene_export.py:104-181; evaluated as src/node_evaluator.cpp does) ----

def _rgba(c):
    return [float(x) for x in c] + [1.0]


def _node(name, typ, inputs=(), outputs=()):
    return {"name": name, "type": typ, "inputs": [dict(x) for x in inputs], "outputs": [dict(x) for x in outputs]}


def _link(from_node, from_socket, to_node, to_socket):
    return {"from_node": from_node, "from_socket": from_socket, "to_node": to_node, "to_socket": to_socket}


def _output_node():
    return _node("Material Output", "ShaderNodeOutputMaterial",
                 [{"name": "Surface"}, {"name": "Volume"}, {"name": "Displacement"}])


def _principled_node(name="Principled BSDF", base=(0.8, 0.8, 0.8), metallic=0.0, roughness=0.5,
                     emission=(0.0, 0.0, 0.0), strength=0.0):
    return _node(name, "ShaderNodeBsdfPrincipled", [
        {"name": "Base Color", "default_value": _rgba(base)}, {"name": "Metallic", "default_value": metallic},
        {"name": "Roughness", "default_value": roughness}, {"name": "IOR", "default_value": 1.45},
        {"name": "Transmission Weight", "default_value": 0.0},
        {"name": "Emission Color", "default_value": _rgba(emission)},
        {"name": "Emission Strength", "default_value": strength}], [{"name": "BSDF"}])


def _checker_node(name, c1, c2, scale):
    return _node(name, "ShaderNodeTexChecker", [
        {"name": "Vector"}, {"name": "Color1", "default_value": _rgba(c1)},
        {"name": "Color2", "default_value": _rgba(c2)}, {"name": "Scale", "default_value": float(scale)}],
        [{"name": "Color"}, {"name": "Fac"}])


def _quad_uvs(q):
    """Per-triangle UVs of _quad's two triangles from its corner UVs q[0..3]."""
    return [[q[0], q[1], q[2]], [q[0], q[2], q[3]]]


def textured_box(seed: int = 6, panel: bool = True) -> dict:
    """The closed box with node-tree materials: a checker floor (negative UVs
    too), a back wall mixing a checker (one colour from an RGB node) with a
    constant, a left wall whose Principled tree folds to constants, a box
    whose tree is present but unused (use_nodes false), and (with `panel`)
    an emissive checker panel on the right wall, lit by the ceiling area
    light."""
    rng = np.random.default_rng(seed)
    scene = closed_box(seed)
    m = scene["meshes"]
    floor = {"use_nodes": True, "node_tree": {
        "nodes": [_output_node(), _principled_node(roughness=0.6),
                  _checker_node("Checker", (0.85, 0.8, 0.7), (0.1, 0.25, 0.15), 3.0)],
        "links": [_link("Principled BSDF", "BSDF", "Material Output", "Surface"),
                  _link("Checker", "Color", "Principled BSDF", "Base Color")]}}
    m[0]["material"] = dict(m[0]["material"], **floor)
    m[0]["triangle_uvs"] = _quad_uvs([[-1.0, -1.0], [-1.0, 1.0], [1.0, 1.0], [1.0, -1.0]])
    back = {"use_nodes": True, "node_tree": {
        "nodes": [_output_node(),
                  _node("Diffuse BSDF", "ShaderNodeBsdfDiffuse", [{"name": "Color", "default_value": _rgba((0.5,) * 3)},
                                                                  {"name": "Roughness", "default_value": 0.0}],
                        [{"name": "BSDF"}]),
                  _node("Mix", "ShaderNodeMixRGB", [{"name": "Fac", "default_value": 0.35},
                                                    {"name": "A", "default_value": _rgba((0, 0, 0))},
                                                    {"name": "B", "default_value": _rgba((0.2, 0.4, 0.8))}],
                        [{"name": "Color"}]),
                  _checker_node("Checker", (0.8, 0.8, 0.8), (0.05, 0.05, 0.05), 5.0),
                  _node("RGB", "ShaderNodeRGB", [], [{"name": "Color", "default_value": _rgba((0.9, 0.6, 0.2))}])],
        "links": [_link("Diffuse BSDF", "BSDF", "Material Output", "Surface"),
                  _link("Mix", "Color", "Diffuse BSDF", "Color"), _link("Checker", "Color", "Mix", "A"),
                  _link("RGB", "Color", "Checker", "Color1")]}}
    m[2]["material"] = dict(m[2]["material"], **back)
    m[2]["triangle_uvs"] = _quad_uvs([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])
    left = {"use_nodes": True, "node_tree": {
        "nodes": [_output_node(), _principled_node(base=tuple(rng.uniform(0.2, 0.8, 3)), metallic=0.3, roughness=0.25)],
        "links": [_link("Principled BSDF", "BSDF", "Material Output", "Surface")]}}
    m[3]["material"] = dict(m[3]["material"], **left)
    unused = {"use_nodes": False, "node_tree": left["node_tree"]}
    m[5]["material"] = dict(m[5]["material"], **unused)
    panel_v, panel_t = _quad([0.99, 0.5, -0.6], [0.99, 0.5, 0.2], [0.99, 1.3, 0.2], [0.99, 1.3, -0.6])
    panel_mat = {"base_color": [0.8, 0.8, 0.8], "use_nodes": True, "node_tree": {
        "nodes": [_output_node(),
                  _node("Emission", "ShaderNodeEmission", [{"name": "Color", "default_value": _rgba((1, 1, 1))},
                                                           {"name": "Strength", "default_value": 2.0}],
                        [{"name": "Emission"}]),
                  _checker_node("Checker", (6.0, 5.0, 4.0), (0.0, 0.0, 0.0), 2.0)],
        "links": [_link("Emission", "Emission", "Material Output", "Surface"),
                  _link("Checker", "Color", "Emission", "Color")]}}
    pm = _mesh(panel_v, panel_t, panel_mat)
    pm["triangle_uvs"] = _quad_uvs([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])
    if panel:
        m.append(pm)
    return scene


def smoke_box(sidecar_path: str, grid_n: int = 32, seed: int = 4, density: float = 6.0, anisotropy: float = 0.3,
              albedo: float = 0.9) -> dict:
    """The closed box with a gridded smoke medium (fBm density, albedo 0.9)
    inside a boundary-only cube, plus an emissive-mesh light and a spot light.
    Writes the float32 density grid to `sidecar_path`."""
    scene = closed_box(seed=1)
    g = fbm_grid(grid_n, seed)
    g.tofile(sidecar_path)
    lo, hi = [-0.8, 0.15, -0.8], [0.8, 1.75, 0.8]
    v, t = _box([(lo[i] + hi[i]) / 2 for i in range(3)], [(hi[i] - lo[i]) / 2 for i in range(3)], 0.0)
    vol = {"color": [albedo, albedo, albedo], "absorption_color": [0.0, 0.0, 0.0], "density": density,
           "anisotropy": anisotropy,
           "smoke_grid": {"dims": [grid_n, grid_n, grid_n], "world_min": lo, "world_max": hi,
                          "sidecar_path": sidecar_path}}
    scene["meshes"].append(_mesh(v, t, {"base_color": [0.8, 0.8, 0.8], "metallic": 0.0, "roughness": 0.5,
                                        "emission": [0.0, 0.0, 0.0], "node_tree": boundary_node_tree(),
                                        "volume": vol}))
    # emissive mesh light: a small quad on the left wall
    ev, et = _quad([-0.99, 1.2, -0.3], [-0.99, 1.2, 0.1], [-0.99, 1.6, 0.1], [-0.99, 1.6, -0.3])
    scene["meshes"].append(_mesh(ev, et, {"base_color": [0.0, 0.0, 0.0], "metallic": 0.0, "roughness": 1.0,
                                          "emission": [6.0, 4.0, 2.0], "transmission": 0.0, "ior": 1.45}))
    scene["native_lights"].append({"type": "SPOT", "name": "Spot", "position": [0.6, 1.9, 0.6],
                                   "direction": [-0.4, -1.0, -0.4], "spot_size": 0.9, "spot_blend": 0.2,
                                   "color": [1.0, 1.0, 1.0], "energy": 40.0})
    return scene


SMOKE_BOX_CAMERA = CLOSED_BOX_CAMERA


def _grid_quad(origin, du, dv, n):
    """A planar quad origin + a*du + b*dv, a, b in [0, 1], as an n x n grid
    of 2 n^2 triangles."""
    a = np.linspace(0.0, 1.0, n + 1)
    A, Bm = np.meshgrid(a, a, indexing="ij")
    v = (np.asarray(origin, np.float64)[None] + A.reshape(-1, 1) * np.asarray(du, np.float64)[None]
         + Bm.reshape(-1, 1) * np.asarray(dv, np.float64)[None])
    i = np.arange(n)[:, None] * (n + 1) + np.arange(n)[None, :]
    i = i.reshape(-1)
    t = np.concatenate([np.stack([i, i + n + 1, i + n + 2], 1), np.stack([i, i + n + 2, i + 1], 1)])
    return v, t.astype(np.int64)


def _glossy(rng, lo=0.1, hi=0.4):
    return {"base_color": [float(x) for x in rng.uniform(0.55, 0.85, 3)], "metallic": 0.0,
            "roughness": float(rng.uniform(lo, hi)), "emission": [0.0, 0.0, 0.0], "transmission": 0.0, "ior": 1.45}


def occluded_room(seed: int = 5, grid: int = 150, spheres: int = 20) -> dict:
    """An 8 m x 3 m x 8 m room cut by a partition wall at z = -3 whose only
    opening is a 5 cm horizontal slit (y in [2.00, 2.05], x in [-3, 3]).  A
    2 m x 1.2 m area light behind the partition faces it, hidden from the
    camera in the front part; the front part is lit only through the slit.
    Walls are finely tessellated glossy surfaces (roughness U[0.1, 0.4]),
    plus `spheres` glossy icospheres on the floor: about 300k triangles at
    the defaults."""
    rng = np.random.default_rng(seed)
    X0, X1, Y1, Z0, Z1 = -4.0, 4.0, 3.0, -6.0, 2.0
    meshes = []
    walls = [
        ((X0, 0, Z0), (X1 - X0, 0, 0), (0, 0, Z1 - Z0)),      # floor
        ((X0, Y1, Z0), (0, 0, Z1 - Z0), (X1 - X0, 0, 0)),     # ceiling
        ((X0, 0, Z0), (0, Y1, 0), (X1 - X0, 0, 0)),           # back wall
        ((X0, 0, Z1), (X1 - X0, 0, 0), (0, Y1, 0)),           # front wall (behind the camera)
        ((X0, 0, Z0), (0, 0, Z1 - Z0), (0, Y1, 0)),           # left
        ((X1, 0, Z0), (0, Y1, 0), (0, 0, Z1 - Z0)),           # right
    ]
    for o, du, dv in walls:
        v, t = _grid_quad(o, du, dv, grid)
        meshes.append(_mesh(v, t, _glossy(rng)))
    # partition (0.1 m thick) with the slit
    zc, hz = -3.0, 0.05
    parts = [((0.0, 1.0, zc), (4.0, 1.0, hz)),            # below the slit, y in [0, 2]
             ((0.0, 2.525, zc), (4.0, 0.475, hz)),        # above the slit, y in [2.05, 3]
             ((-3.5, 2.025, zc), (0.5, 0.025, hz)),       # slit ends
             ((3.5, 2.025, zc), (0.5, 0.025, hz))]
    pmat = _glossy(rng)
    for c, h in parts:
        v, t = _box(c, h, 0.0)
        meshes.append(_mesh(v, t, pmat))
    bv, bt = _icosphere(3)
    for _ in range(spheres):
        r = float(rng.uniform(0.25, 0.55))
        c = np.array([rng.uniform(-3.2, 3.2), r, rng.uniform(-2.5, 0.5)])
        meshes.append(_mesh(bv * r + c, bt, _glossy(rng), normals=None))
    radiance, sx, sy = 400.0, 2.0, 1.2
    light = {"type": "AREA", "name": "HiddenLight", "position": [0.0, 1.4, -5.9], "direction": [0.0, 0.0, 1.0],
             "right": [1.0, 0.0, 0.0], "up": [0.0, 1.0, 0.0], "shape": "RECTANGLE", "size": sx, "size_y": sy,
             "color": [1.0, 0.95, 0.85], "energy": radiance * math.pi * sx * sy}
    return {"meshes": meshes, "native_lights": [light],
            "environment": {"color": [0.0, 0.0, 0.0], "strength": 0.0}}


OCCLUDED_ROOM_CAMERA = {"pos": (0.0, 1.5, 1.9), "dir": (0.0, 0.08, -1.0), "up": (0.0, 1.0, 0.0), "fov": 65.0}
# a second view of the room (from the front-left corner towards the far right
# of the front part), not used to choose the guiding defaults
OCCLUDED_ROOM_HELDOUT_CAMERA = {"pos": (-3.4, 2.3, 1.6), "dir": (0.75, -0.3, -0.6), "up": (0.0, 1.0, 0.0),
                                "fov": 60.0}


def relmse_pixels(width: int = 1920, height: int = 1080, nx: int = 128, ny: int = 128):
    """Stratified pixel subset of the relMSE measurement: the centre of each
    cell of an nx x ny grid over the frame (16,384 pixels by default)."""
    xs = ((np.arange(nx) + 0.5) * width / nx).astype(np.int32)
    ys = ((np.arange(ny) + 0.5) * height / ny).astype(np.int32)
    X, Y = np.meshgrid(xs, ys)
    return X.reshape(-1), Y.reshape(-1)


def relmse(x: np.ndarray, r: np.ndarray, eps: float = 0.01) -> float:
    """mean over pixels and channels of (x - r)^2 / (r^2 + eps)."""
    x = np.asarray(x, np.float64)
    r = np.asarray(r, np.float64)
    return float(np.mean((x - r) ** 2 / (r * r + eps)))


def session_scene(seed: int, triangles: int) -> dict:
    """A scene of about `triangles` triangles for the batched-session
    workload: a ground quad and displaced icospheres (5,120 triangles each,
    1,280 below 10k), Principled or Emission materials (one object in seven
    emits), and 1-8 native lights of every type.  Vertex and triangle
    arrays stay numpy arrays (to_json converts them)."""
    rng = np.random.default_rng(seed)
    level = 4 if triangles >= 10000 else 3
    if level not in _ICO_CACHE:
        _ICO_CACHE[level] = _icosphere(level)
    base_v, base_t = _ICO_CACHE[level]
    count = max(1, int(round(triangles / len(base_t))))
    extent = 3.0 + 1.5 * math.sqrt(count)
    meshes = [_mesh(*_quad((-extent, 0, -extent), (-extent, 0, extent), (extent, 0, extent), (extent, 0, -extent)),
                    _diffuse(rng.uniform(0.3, 0.8, 3).tolist()))]
    for i in range(count):
        radius = rng.uniform(0.3, 0.9)
        center = np.array([rng.uniform(-extent, extent), rng.uniform(0.4, 3.0), rng.uniform(-extent, extent)])
        k = rng.normal(size=(3, 3))
        k /= np.linalg.norm(k, axis=1, keepdims=True)
        disp = 1.0 + (rng.uniform(0.02, 0.1, 3)[None, :] * np.sin(3.0 * base_v @ k.T)).sum(1)
        pos = base_v * (radius * disp)[:, None] + center
        if rng.uniform() < 1.0 / 7.0:
            mat = {"base_color": [0.0, 0.0, 0.0], "metallic": 0.0, "roughness": 1.0,
                   "emission": (rng.uniform(0.5, 1.0, 3) * rng.uniform(2.0, 8.0)).tolist(),
                   "transmission": 0.0, "ior": 1.5}
        else:
            mat = {"base_color": rng.uniform(0.15, 0.9, 3).tolist(),
                   "metallic": 1.0 if rng.uniform() < 0.3 else 0.0,
                   "roughness": float(rng.uniform(0.05, 1.0)), "emission": [0.0, 0.0, 0.0],
                   "transmission": 0.0, "ior": float(rng.uniform(1.3, 1.7))}
        meshes.append({"vertices": pos.astype(np.float32), "triangles": base_t.astype(np.int32), "material": mat})
    lights = []
    for j in range(int(rng.integers(1, 9))):
        kind = ("POINT", "SUN", "SPOT", "AREA")[j % 4 if j < 4 else int(rng.integers(0, 4))]
        pos = [float(rng.uniform(-extent, extent)), float(rng.uniform(4.0, 9.0)), float(rng.uniform(-extent, extent))]
        col = rng.uniform(0.7, 1.0, 3).tolist()
        if kind == "POINT":
            lights.append({"type": "POINT", "position": pos, "color": col, "energy": float(rng.uniform(300, 1500)),
                           "radius": float(rng.choice([0.0, 0.25]))})
        elif kind == "SUN":
            lights.append({"type": "SUN", "position": pos, "direction": [-0.3, -1.0, -0.2], "color": col,
                           "energy": float(rng.uniform(1.0, 3.0))})
        elif kind == "SPOT":
            lights.append({"type": "SPOT", "position": pos, "direction": [0.0, -1.0, 0.1], "color": col,
                           "energy": float(rng.uniform(500, 2000)), "spot_size": 1.0, "spot_blend": 0.2})
        else:
            lights.append({"type": "AREA", "position": pos, "direction": [0.0, -1.0, 0.0], "right": [1.0, 0.0, 0.0],
                           "up": [0.0, 0.0, 1.0], "shape": "RECTANGLE", "size": 2.0, "size_y": 1.0,
                           "color": col, "energy": float(rng.uniform(300, 900))})
    return {"meshes": meshes, "native_lights": lights,
            "environment": {"color": [0.05, 0.06, 0.08], "strength": 1.0}}


def session_camera(scene: dict) -> dict:
    """Looks at the scene from outside its ground quad, slightly above."""
    ext = float(np.max(np.abs(np.asarray(scene["meshes"][0]["vertices"], np.float64))))
    return {"pos": (0.0, 0.35 * ext + 2.0, 1.6 * ext), "dir": (0.0, -0.3, -1.0), "up": (0.0, 1.0, 0.0),
            "fov": 50.0}


def batched_sessions(seed: int = 7, count: int = 32, tri_range=(1e4, 5e5), edge_range=(128, 1024)):
    """`count` sessions: (scene, camera, width, height) with log-uniform
    edges and triangle counts."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        w, h = (int(round(math.exp(rng.uniform(math.log(edge_range[0]), math.log(edge_range[1]))))) for _ in "wh")
        tris = int(math.exp(rng.uniform(math.log(tri_range[0]), math.log(tri_range[1]))))
        sc = session_scene(int(rng.integers(1 << 30)), tris)
        out.append((sc, session_camera(sc), w, h))
    return out


def volume_scene(sidecar_path: str, seed: int = 11, spheres: int = 780, grid_n: int = 256) -> dict:
    """The 4M-triangle participating-media workload: `spheres` displaced
    icospheres (5,120 triangles each) on a ground quad around a 4 m cube of
    fBm smoke (albedo 0.9, an grid_n^3 density grid written to
    `sidecar_path`, boundary-only cube), an emissive quad and a spot light.
    Vertex and triangle arrays stay numpy arrays."""
    rng = np.random.default_rng(seed)
    if 4 not in _ICO_CACHE:
        _ICO_CACHE[4] = _icosphere(4)
    base_v, base_t = _ICO_CACHE[4]
    ext = 22.0
    meshes = [_mesh(*_quad((-ext, 0, -ext), (-ext, 0, ext), (ext, 0, ext), (ext, 0, -ext)), _diffuse([0.6, 0.6, 0.6]))]
    for i in range(spheres):
        radius = rng.uniform(0.3, 0.8)
        while True:
            c = np.array([rng.uniform(-ext, ext), rng.uniform(0.4, 4.0), rng.uniform(-ext, ext)])
            if max(abs(c[0]), abs(c[2])) > 3.0:  # keep the smoke cube free
                break
        k = rng.normal(size=(3, 3))
        k /= np.linalg.norm(k, axis=1, keepdims=True)
        disp = 1.0 + (rng.uniform(0.02, 0.1, 3)[None, :] * np.sin(3.0 * base_v @ k.T)).sum(1)
        mat = {"base_color": rng.uniform(0.15, 0.9, 3).tolist(), "metallic": 1.0 if rng.uniform() < 0.3 else 0.0,
               "roughness": float(rng.uniform(0.05, 1.0)), "emission": [0.0, 0.0, 0.0], "transmission": 0.0,
               "ior": float(rng.uniform(1.3, 1.7))}
        meshes.append({"vertices": (base_v * (radius * disp)[:, None] + c).astype(np.float32),
                       "triangles": base_t.astype(np.int32), "material": mat})
    fbm_grid(grid_n, seed).tofile(sidecar_path)
    lo, hi = [-2.0, 0.05, -2.0], [2.0, 4.05, 2.0]
    v, t = _box([(lo[i] + hi[i]) / 2 for i in range(3)], [(hi[i] - lo[i]) / 2 for i in range(3)], 0.0)
    vol = {"color": [0.9, 0.9, 0.9], "absorption_color": [0.0, 0.0, 0.0], "density": 4.0, "anisotropy": 0.3,
           "smoke_grid": {"dims": [grid_n] * 3, "world_min": lo, "world_max": hi, "sidecar_path": sidecar_path}}
    meshes.append(_mesh(v, t, {"base_color": [0.8, 0.8, 0.8], "metallic": 0.0, "roughness": 0.5,
                               "emission": [0.0, 0.0, 0.0], "node_tree": boundary_node_tree(), "volume": vol}))
    ev, et = _quad([-1.0, 7.0, -1.0], [1.0, 7.0, -1.0], [1.0, 7.0, 1.0], [-1.0, 7.0, 1.0])
    meshes.append(_mesh(ev, et, {"base_color": [0.0, 0.0, 0.0], "metallic": 0.0, "roughness": 1.0,
                                 "emission": [12.0, 11.0, 9.0], "transmission": 0.0, "ior": 1.45}))
    lights = [{"type": "SPOT", "position": [6.0, 8.0, 6.0], "direction": [-0.55, -0.6, -0.55], "spot_size": 0.8,
               "spot_blend": 0.2, "color": [1.0, 0.95, 0.9], "energy": 6000.0}]
    return {"meshes": meshes, "native_lights": lights, "environment": {"color": [0.05, 0.06, 0.08], "strength": 1.0}}


def volume_sessions(seed: int = 12, count: int = 16):
    """`count` (camera, width, height) views of volume_scene: edges
    log-uniform between 256 x 256 and 1920 x 1080, cameras on a ring around
    the smoke cube."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        w = int(round(math.exp(rng.uniform(math.log(256), math.log(1920)))))
        h = int(round(math.exp(rng.uniform(math.log(256), math.log(1080)))))
        a = 2.0 * math.pi * i / count
        pos = (9.0 * math.cos(a), 3.0 + rng.uniform(0.0, 2.0), 9.0 * math.sin(a))
        d = (-pos[0], 2.0 - pos[1], -pos[2])
        out.append(({"pos": pos, "dir": d, "up": (0.0, 1.0, 0.0), "fov": 50.0}, w, h))
    return out


def to_json(scene: dict) -> str:
    return json.dumps(scene, separators=(",", ":"), default=lambda a: np.asarray(a).tolist())


WORKLOADS = {
    "closed_box": {"make": closed_box, "camera": CLOSED_BOX_CAMERA, "width": 256, "height": 256, "spp": 16,
                   "max_depth": 8, "algorithm": "mis"},
    "icospheres": {"make": icospheres, "camera": ICOSPHERES_CAMERA, "width": 1024, "height": 1024, "spp": 64,
                   "max_depth": 12, "algorithm": "mis"},
    "occluded_room": {"make": occluded_room, "camera": OCCLUDED_ROOM_CAMERA, "width": 1920, "height": 1080,
                      "spp": 1024, "max_depth": 12, "algorithm": "mis"},
}
