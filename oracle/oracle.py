"""ctypes front end of the CPU oracle (liblucid_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the CPU-baseline leg of ``bench.py``.  The product never imports it.

The oracle restates the reference CPU tracer (LucidRenderer/cpp_renderer/
include/**) in plain float C++; see lucid_oracle.cpp for the per-function
file:line citations.  Scenes are given in the reference's scene-JSON schema
(pybind_renderer.hpp:1258-1746) as a Python dict.

How each part is pinned (checks live in tests/test_oracle_*.py):

* seed_random_xyz / pcg32 / randf -- PCG32 core checked against the published
  pcg32 demo stream (pcg32_srandom_r(42, 54)) and the textbook mixers
  (MurmurHash3 fmix32, SplitMix64, Jenkins 32-bit) recomputed independently
  in numpy; randf by its closed form.
* Moller-Trumbore, AABB slab, sphere / rectangle / ellipse tests -- the
  reference's known answers (test_geometry.cpp) plus brute-force
  plane/barycentric solutions in float64.
* Fresnel, GGX D / G1 / G2 -- reference known answers (test_bsdf.cpp:
  F(1, 1.5) = 0.04, TIR = 1) plus closed forms evaluated in float64.
* BSDF sampling and pdf -- Monte Carlo: the sampled-direction histogram
  matches pdf_bsdf, and pdf_bsdf integrates to (1 - transmission) over the
  hemisphere.
* light pdf, MIS power heuristic, bsdf_sample_weight -- reference known
  answers (test_light.cpp 2.25 sr^-1; test_integrator.cpp 0.8 weight;
  heuristic formula and complementarity).
* BVH build and closest-hit traversal -- brute force over all triangles on
  random rays (same closest t, tie broken by the reference's leaf order).
* traceSimple / traceNEE / traceMIS -- whole-path statistics from the
  reference's tests: NEE and MIS agree within 3 sigma on the Cornell box
  (test_path_tracer.cpp:193-217), env-only scene returns the env colour
  exactly (:132-152), white furnace bound (:314-361), closed-form point light
  f*I/d^2 (test_gpu_path_tracer.cpp:281-349), and a direct-lighting integral
  computed by numerical quadrature.
* node-tree materials (node_evaluator.cpp) -- checkerboard parity against
  floor(u * scale) mod 2 computed in numpy, Mix against the float32 lerp,
  colour x strength for Emission / Principled emission, the reference's
  defaults and fallbacks case by case, and the hit's texture coordinate
  against the affine UV map of a planar quad (tests/test_node_materials.py).
* traceVolumeMIS -- homogeneous-slab transmittance against exp(-sigma*d)
  closed form; scatter vs no-scatter energy bound.  Per-sample agreement with
  the reference binary is "parity unpinned" (the reference cannot be built
  here: it needs mp-units 2.4.0, nlohmann_json 3.11.3 and Dawn).
"""

from __future__ import annotations

import ctypes as C
import json
import os
import struct
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

LIGHT_TYPES = {"POINT": 1, "SUN": 2, "SPOT": 3, "AREA": 4}
AREA_SHAPES = {"SQUARE": 0, "RECTANGLE": 1, "DISK": 2, "ELLIPSE": 3}


class _NodeSocket(C.Structure):
    _fields_ = [("name", C.c_char_p), ("kind", C.c_int32), ("value", C.c_float * 4), ("linked_node", C.c_char_p),
                ("linked_socket", C.c_char_p)]


class _Node(C.Structure):
    _fields_ = [("name", C.c_char_p), ("type", C.c_char_p), ("inputs", C.POINTER(_NodeSocket)),
                ("num_inputs", C.c_int32), ("outputs", C.POINTER(_NodeSocket)), ("num_outputs", C.c_int32)]


class _MeshDesc(C.Structure):
    _fields_ = [
        ("num_vertices", C.c_int32),
        ("vertices", C.POINTER(C.c_float)),
        ("num_triangles", C.c_int32),
        ("indices", C.POINTER(C.c_int32)),
        ("tri_normals", C.POINTER(C.c_float)),
        ("num_tri_normals", C.c_int32),
        ("smooth", C.c_int32),
        ("base_color", C.c_float * 3),
        ("metallic", C.c_float),
        ("roughness", C.c_float),
        ("emission", C.c_float * 3),
        ("transmission", C.c_float),
        ("ior", C.c_float),
        ("has_volume", C.c_int32),
        ("vol_color", C.c_float * 3),
        ("vol_absorption", C.c_float * 3),
        ("vol_density", C.c_float),
        ("vol_anisotropy", C.c_float),
        ("grid_dims", C.c_int32 * 3),
        ("grid_min", C.c_float * 3),
        ("grid_max", C.c_float * 3),
        ("grid", C.POINTER(C.c_float)),
        ("surface_linked", C.c_int32),
        ("use_nodes", C.c_int32),
        ("nodes", C.POINTER(_Node)),
        ("num_nodes", C.c_int32),
        ("tri_uvs", C.POINTER(C.c_float)),
        ("num_tri_uvs", C.c_int32),
    ]


class _LightDesc(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("position", C.c_float * 3),
        ("has_direction", C.c_int32),
        ("direction", C.c_float * 3),
        ("color", C.c_float * 3),
        ("energy", C.c_float),
        ("radius", C.c_float),
        ("spot_size", C.c_float),
        ("spot_blend", C.c_float),
        ("size", C.c_float),
        ("size_y", C.c_float),
        ("shape", C.c_int32),
        ("has_right", C.c_int32),
        ("has_up", C.c_int32),
        ("right", C.c_float * 3),
        ("up", C.c_float * 3),
    ]


_libs = {}


def lib(fast: bool = False) -> C.CDLL:
    """Loads liblucid_oracle.so (parity build) or the -ffast-math baseline build."""
    name = "liblucid_oracle_fast.so" if fast else "liblucid_oracle.so"
    if name in _libs:
        return _libs[name]
    path = os.path.join(_HERE, name)
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
    L = C.CDLL(path)
    vp, fp, ip, i32p, i64p, dp = C.c_void_p, C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_longlong), C.POINTER(C.c_double)
    L.orc_scene_new.restype = vp
    L.orc_scene_free.argtypes = [vp]
    L.orc_scene_add_mesh.argtypes = [vp, C.POINTER(_MeshDesc)]
    L.orc_scene_add_light.argtypes = [vp, C.POINTER(_LightDesc)]
    L.orc_scene_set_env.argtypes = [vp, C.c_float, C.c_float, C.c_float, C.c_float]
    L.orc_scene_finalize.argtypes = [vp]
    L.orc_set_camera.argtypes = [vp] + [C.c_float] * 10
    L.orc_set_exposure.argtypes = [vp, C.c_float]
    L.orc_num_lights.argtypes = [vp]
    L.orc_native_light_start.argtypes = [vp]
    L.orc_light_cdf.argtypes = [vp, ip]
    L.orc_light_cdf.restype = C.c_float
    L.orc_light_total_area.argtypes = [vp]
    L.orc_light_total_area.restype = C.c_float
    L.orc_light_emission.argtypes = [vp, ip, fp]
    L.orc_material_eval.argtypes = [vp, C.c_int, C.c_int, fp, fp]
    L.orc_intersect_uv.argtypes = [vp, C.c_int, fp, fp, fp]
    L.orc_render_tile.argtypes = [vp, C.c_char_p] + [ip] * 11 + [fp, dp, dp, i64p]
    L.orc_trace_pixel.argtypes = [vp, C.c_char_p] + [ip] * 8 + [fp]
    L.orc_primary_hits.argtypes = [vp] + [ip] * 8 + [i32p, i32p, i32p, fp, fp, fp]
    L.orc_render_debug.argtypes = [vp] + [ip] * 7 + [fp]
    L.orc_intersect.argtypes = [vp, ip, fp, fp, fp, fp, i32p, i32p, i32p, fp, fp, fp]
    L.orc_mesh_bvh.argtypes = [vp, ip, C.POINTER(C.c_int), fp, i32p, i32p]
    L.orc_seed.argtypes = [C.c_uint32] * 3 + [C.POINTER(C.c_uint64)] * 2
    L.orc_pcg32_stream.argtypes = [C.c_uint64, C.c_uint64, ip, C.POINTER(C.c_uint32)]
    L.orc_randf_stream.argtypes = [C.c_uint32] * 3 + [ip, fp]
    for nm in ("orc_jenkins", "orc_fmix32"):
        getattr(L, nm).argtypes = [C.c_uint32]
        getattr(L, nm).restype = C.c_uint32
    L.orc_splitmix64.argtypes = [C.c_uint64]
    L.orc_splitmix64.restype = C.c_uint64
    L.orc_intersect_triangle.argtypes = [fp] * 5 + [C.c_float, C.c_float, fp]
    L.orc_intersect_aabb.argtypes = [fp] * 4
    L.orc_intersect_sphere.argtypes = [fp, fp, fp, C.c_float, fp]
    for nm, n in (("orc_fresnel_dielectric", 2), ("orc_fresnel_schlick", 2), ("orc_ggx_d", 2), ("orc_ggx_g1", 2),
                  ("orc_ggx_g2", 3), ("orc_power_heuristic", 2), ("orc_pdf_w_from_area", 3), ("orc_hg_eval", 2)):
        getattr(L, nm).argtypes = [C.c_float] * n
        getattr(L, nm).restype = C.c_float
    L.orc_bsdf_weight.argtypes = [fp, C.c_float, C.c_float, fp]
    L.orc_eval_bsdf.argtypes = [fp, fp, fp, fp, fp, fp]
    L.orc_sample_bsdf.argtypes = [fp, fp, fp, C.c_float, C.c_float, C.c_float, C.c_uint32, C.c_uint32, C.c_uint32, fp]
    L.orc_sample_light.argtypes = [vp, ip, fp, C.c_float, C.c_float, fp]
    L.orc_select_light.argtypes = [vp, C.c_float, fp]
    L.orc_pdf_light_area.argtypes = [C.c_float, fp, fp, fp]
    u8p = C.POINTER(C.c_uint8)
    L.orc_render_pixels.argtypes = [vp, C.c_char_p, ip, ip, ip, i32p, i32p, ip, ip, C.c_longlong, ip, ip, dp, dp]
    L.orc_render_tile_diag.argtypes = [vp, C.c_char_p] + [ip] * 11 + [fp, u8p, dp, i64p]
    L.orc_diag_pixel_run.argtypes = [ip, ip, ip, i32p, i32p, C.POINTER(C.c_int16), u8p, u8p, fp, u8p, u8p,
                                     C.POINTER(C.c_uint64), u8p, i64p, C.POINTER(C.c_uint64),
                                     C.POINTER(C.c_uint32), fp, fp, i32p, i32p, fp, fp, C.POINTER(C.c_uint32), fp,
                                     u8p]
    L.orc_pdf_light_area.restype = C.c_float
    _libs[name] = L
    return L


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _iptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _f3(v, default=(0.0, 0.0, 0.0)):
    v = default if v is None else v
    return (C.c_float * 3)(float(v[0]), float(v[1]), float(v[2]))


def surface_linked(material: dict) -> bool:
    """Material Output "Surface" socket linked? (pybind_renderer.hpp:1386-1398
    with parseNodeTree, node_evaluator.cpp:82-149: a socket is linked when a
    "links" entry targets (node name, socket name); a tree without nodes is
    invalid and leaves the surface linked)."""
    nt = material.get("node_tree")
    if not isinstance(nt, dict) or not isinstance(nt.get("nodes"), list) or not nt["nodes"]:
        return True
    links = nt.get("links") or []
    for n in nt["nodes"]:
        if n.get("type", "") == "ShaderNodeOutputMaterial":
            name = n.get("name", "")
            for s in n.get("inputs", []) or []:
                if s.get("name", "") == "Surface":
                    return any(l.get("to_node") == name and l.get("to_socket") == "Surface" for l in links)
            return False
    return True


def _node_tree(material: dict, keep: list):
    """parseNodeTree (node_evaluator.cpp:82-149) of a material dict: the
    nodes with their sockets' parseSocketValue kinds (:47-69; 1 float,
    2 vec3, 3 vec4, 0 anything else) and each input's link from the last
    "links" entry that targets it.  Returns (array, count, use_nodes)."""
    nt = material.get("node_tree")
    if not isinstance(nt, dict) or not isinstance(nt.get("nodes"), list) or not nt["nodes"]:
        return None, 0, 0
    links = {}
    for l in nt.get("links") or []:
        links[(l["to_node"], l["to_socket"])] = (l["from_node"], l["from_socket"])

    def enc(x):
        b = str(x).encode()
        keep.append(b)
        return b

    def kind(v):
        if isinstance(v, bool) or v is None:
            return 0, ()
        if isinstance(v, (int, float)):
            return 1, (float(v),)
        if isinstance(v, list) and len(v) in (3, 4) and all(
                isinstance(x, (int, float)) and not isinstance(x, bool) for x in v):
            return (2 if len(v) == 3 else 3), tuple(float(x) for x in v)
        return 0, ()

    arr = (_Node * len(nt["nodes"]))()
    keep.append(arr)
    for i, n in enumerate(nt["nodes"]):
        name = n.get("name", "")
        name = name if isinstance(name, str) else ""
        typ = n.get("type", "")
        arr[i].name, arr[i].type = enc(name), enc(typ if isinstance(typ, str) else "")
        for side in ("inputs", "outputs"):
            socks = n.get(side) if isinstance(n.get(side), list) else []
            sa = (_NodeSocket * max(1, len(socks)))()
            keep.append(sa)
            for k, sj in enumerate(socks):
                sn = sj.get("name", "")
                sn = sn if isinstance(sn, str) else ""
                sa[k].name = enc(sn)
                sa[k].kind, vals = kind(sj.get("default_value"))
                for j, x in enumerate(vals):
                    sa[k].value[j] = x
                if side == "inputs" and (name, sn) in links:
                    sa[k].linked_node, sa[k].linked_socket = (enc(x) for x in links[(name, sn)])
            setattr(arr[i], side, C.cast(sa, C.POINTER(_NodeSocket)))
            setattr(arr[i], "num_" + side, len(socks))
    return arr, len(nt["nodes"]), 1 if material.get("use_nodes", False) else 0


class OracleScene:
    """A scene loaded into the oracle, from a reference-schema scene dict."""

    def __init__(self, scene: dict, fast: bool = False):
        self.L = lib(fast)
        self.h = self.L.orc_scene_new()
        self._ntri = []
        for m in scene.get("meshes", []):
            self._add_mesh(m)
        lights = scene.get("native_lights", scene.get("lights"))
        for l in lights or []:
            self._add_light(l)
        env = scene.get("environment")
        if isinstance(env, dict):
            col = env.get("color", [0.05, 0.05, 0.05])
            self.L.orc_scene_set_env(self.h, float(col[0]), float(col[1]), float(col[2]), float(env.get("strength", 1.0)))
        self.L.orc_scene_finalize(self.h)

    def __del__(self):
        try:
            if self.h:
                self.L.orc_scene_free(self.h)
        except Exception:
            pass

    def _add_mesh(self, m: dict):
        verts = np.ascontiguousarray(np.asarray(m.get("vertices", []), dtype=np.float32).reshape(-1, 3))
        tris = np.ascontiguousarray(np.asarray(m.get("triangles", []), dtype=np.int32).reshape(-1, 3))
        d = _MeshDesc()
        d.num_vertices = verts.shape[0]
        d.vertices = _fptr(verts)
        d.num_triangles = tris.shape[0]
        d.indices = _iptr(tris)
        keep = [verts, tris]
        tn = m.get("triangle_normals")
        if tn is not None and len(tn) > 0:
            tn = np.ascontiguousarray(np.asarray(tn, dtype=np.float32).reshape(-1, 9)[: tris.shape[0]])
            d.tri_normals = _fptr(tn)
            d.num_tri_normals = tn.shape[0]
            keep.append(tn)
        d.smooth = 1 if m.get("smooth", False) else 0
        mat = m.get("material") or {}
        d.base_color = _f3(mat.get("base_color"), (0.8, 0.8, 0.8))
        d.metallic = float(mat.get("metallic", 0.0))
        d.roughness = float(mat.get("roughness", 0.5))
        d.emission = _f3(mat.get("emission"))
        d.transmission = float(mat.get("transmission", 0.0))
        d.ior = float(mat.get("ior", 1.45))
        d.surface_linked = 1 if surface_linked(mat) else 0
        nodes, n_nodes, use = _node_tree(mat, keep)
        if n_nodes:
            d.nodes, d.num_nodes, d.use_nodes = nodes, n_nodes, use
        tu = m.get("triangle_uvs")
        if tu is not None and len(tu) > 0:
            tu = np.ascontiguousarray(np.asarray(tu, dtype=np.float32).reshape(-1, 6))
            d.tri_uvs = _fptr(tu)
            d.num_tri_uvs = tu.shape[0]
            keep.append(tu)
        vol = mat.get("volume")
        if isinstance(vol, dict):
            d.has_volume = 1
            d.vol_color = _f3(vol.get("color"), (1.0, 1.0, 1.0))
            d.vol_absorption = _f3(vol.get("absorption_color"))
            d.vol_density = float(vol.get("density", 0.0))
            d.vol_anisotropy = float(vol.get("anisotropy", 0.0))
            sg = vol.get("smoke_grid")
            if isinstance(sg, dict) and os.path.exists(sg.get("sidecar_path", "")):
                dims = [int(x) for x in sg["dims"]]
                n = dims[0] * dims[1] * dims[2]
                g = np.fromfile(sg["sidecar_path"], dtype=np.float32, count=n)
                if g.size == n:
                    g = np.ascontiguousarray(g)
                    d.grid_dims = (C.c_int32 * 3)(*dims)
                    d.grid_min = _f3(sg["world_min"])
                    d.grid_max = _f3(sg["world_max"])
                    d.grid = _fptr(g)
                    keep.append(g)
        rc = self.L.orc_scene_add_mesh(self.h, C.byref(d))
        del keep
        if rc != 0:
            raise ValueError("oracle rejected mesh (bad triangle index)")
        self._ntri.append(int(tris.shape[0]))

    def _add_light(self, l: dict):
        d = _LightDesc()
        d.type = LIGHT_TYPES.get(l.get("type", "POINT"), 0)
        d.position = _f3(l.get("position"))
        if "direction" in l:
            d.has_direction = 1
            d.direction = _f3(l["direction"])
        d.color = _f3(l.get("color"), (1.0, 1.0, 1.0))
        d.energy = float(l.get("energy", 1.0))
        d.radius = float(l.get("radius", 0.0))
        d.spot_size = float(l.get("spot_size", np.float32(np.pi / 4.0)))
        d.spot_blend = float(l.get("spot_blend", 0.0))
        d.size = float(l.get("size", 1.0))
        d.size_y = float(l.get("size_y", d.size))
        d.shape = AREA_SHAPES.get(l.get("shape", "SQUARE"), 0)
        if "right" in l:
            d.has_right = 1
            d.right = _f3(l["right"])
        if "up" in l:
            d.has_up = 1
            d.up = _f3(l["up"])
        if self.L.orc_scene_add_light(self.h, C.byref(d)) != 0:
            raise ValueError(f"unknown light type {l.get('type')}")

    # ---- camera / render ----
    def set_camera(self, pos, direction, up, fov_deg):
        self.L.orc_set_camera(self.h, *[float(x) for x in pos], *[float(x) for x in direction],
                              *[float(x) for x in up], float(fov_deg))

    def set_exposure(self, ev: float):
        self.L.orc_set_exposure(self.h, float(ev))

    def render_tile(self, algorithm, tx, ty, tw, th, fw, fh, samples=1, sample_offset=0, max_depth=8,
                    fp64=False, threads=0, stats=False, strategies=False):
        """Returns (rgba_flat, extras).  rgba_flat mirrors PyRenderer::render_tile:
        tw*th*4 floats, bottom row first, RGB = clamp(SUM * 2^EV).
        extras: 'stats' (th*tw, 6) float64 [sum, sumsq] per channel (tile order,
        top row first), 'strat' (th*tw, 2, 6), 'counts' (closest, shadow)."""
        out = np.zeros(tw * th * 4, dtype=np.float32)
        st = np.zeros((th * tw, 6), dtype=np.float64) if stats else None
        sd = np.zeros((th * tw, 2, 6), dtype=np.float64) if strategies else None
        counts = np.zeros(2, dtype=np.int64)
        self.L.orc_render_tile(self.h, algorithm.encode(), tx, ty, tw, th, fw, fh, samples, sample_offset, max_depth,
                               1 if fp64 else 0, threads, _fptr(out),
                               st.ctypes.data_as(C.POINTER(C.c_double)) if st is not None else None,
                               sd.ctypes.data_as(C.POINTER(C.c_double)) if sd is not None else None,
                               counts.ctypes.data_as(C.POINTER(C.c_longlong)))
        return out, {"stats": st, "strat": sd, "counts": counts}

    def render_tile_diag(self, algorithm, tx, ty, tw, th, fw, fh, samples=1, sample_offset=0, max_depth=8,
                         subsample=1, threads=0):
        """render_tile with diagnostics on: (rgba bottom-first, film as a
        PIXEL_DT array (sh, sw), per-strategy FP64 Welford (sh, sw, 4, 7),
        total records)."""
        sw, sh = tw // subsample, th // subsample
        out = np.zeros(tw * th * 4, np.float32)
        film = np.zeros(max(sw * sh, 1) * PIXEL_DT.itemsize, np.uint8)
        strat = np.zeros(max(sw * sh, 1) * 28, np.float64)
        tot = np.zeros(1, np.int64)
        rc = self.L.orc_render_tile_diag(self.h, algorithm.encode(), tx, ty, tw, th, fw, fh, samples, sample_offset,
                                         max_depth, subsample, threads, _fptr(out),
                                         film.ctypes.data_as(C.POINTER(C.c_uint8)),
                                         strat.ctypes.data_as(C.POINTER(C.c_double)),
                                         tot.ctypes.data_as(C.POINTER(C.c_int64)))
        if rc != 0:
            raise RuntimeError(f"orc_render_tile_diag failed ({rc})")
        film = film[: sw * sh * PIXEL_DT.itemsize].view(PIXEL_DT).reshape(sh, sw)
        return out, film, strat[: sw * sh * 28].reshape(sh, sw, 4, 7), int(tot[0])

    def render_pixels(self, algorithm, fw, fh, xs, ys, samples, samples_in_call, sample_offset, max_depth=8,
                      threads=0, acc=None):
        """FP64 radiance sums over `samples` samples at pixels (xs, ys):
        returns (sum (n, 3), sumsq (n, 3)), added onto `acc` when given."""
        xs = np.ascontiguousarray(xs, np.int32)
        ys = np.ascontiguousarray(ys, np.int32)
        n = len(xs)
        sm, sq = acc if acc is not None else (np.zeros((n, 3)), np.zeros((n, 3)))
        rc = self.L.orc_render_pixels(self.h, algorithm.encode(), fw, fh, n, _iptr(xs), _iptr(ys), samples,
                                      samples_in_call, sample_offset, max_depth, threads,
                                      sm.ctypes.data_as(C.POINTER(C.c_double)), sq.ctypes.data_as(C.POINTER(C.c_double)))
        if rc != 0:
            raise RuntimeError("camera not set")
        return sm, sq

    def trace_pixel(self, algorithm, gx, gy, fw, fh, samples_in_call, sample_offset, nsamples, max_depth=8):
        out = np.zeros((nsamples, 3), dtype=np.float32)
        self.L.orc_trace_pixel(self.h, algorithm.encode(), gx, gy, fw, fh, samples_in_call, sample_offset,
                               nsamples, max_depth, _fptr(out))
        return out

    def primary_hits(self, tx, ty, tw, th, fw, fh, samples=1, sample_offset=0):
        n = tw * th * samples
        mesh = np.zeros(n, np.int32); tri = np.zeros(n, np.int32); pos = np.zeros(n, np.int32)
        u = np.zeros(n, np.float32); v = np.zeros(n, np.float32); t = np.zeros(n, np.float32)
        self.L.orc_primary_hits(self.h, tx, ty, tw, th, fw, fh, samples, sample_offset, _iptr(mesh), _iptr(tri),
                                _iptr(pos), _fptr(u), _fptr(v), _fptr(t))
        return {"mesh": mesh, "tri": tri, "pos": pos, "u": u, "v": v, "t": t}

    def render_debug(self, mode, tx, ty, tw, th, fw, fh):
        """render_debug (pybind_renderer.hpp:371-439): tw*th RGBA floats,
        bottom row first; mode "normal", "albedo", "emission", "volume"."""
        code = {"normal": 0, "albedo": 1, "emission": 2, "volume": 3}.get(mode, -1)
        out = np.zeros(tw * th * 4, np.float32)
        if self.L.orc_render_debug(self.h, code, tx, ty, tw, th, fw, fh, _fptr(out)) != 0:
            raise RuntimeError("camera not set")
        return out

    def intersect(self, o, d, tmin, tmax):
        o = np.ascontiguousarray(o, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(d, np.float32).reshape(-1, 3)
        n = o.shape[0]
        tmin = np.ascontiguousarray(np.broadcast_to(np.float32(tmin), (n,)), np.float32) if np.isscalar(tmin) else np.ascontiguousarray(tmin, np.float32)
        tmax = np.ascontiguousarray(np.broadcast_to(np.float32(tmax), (n,)), np.float32) if np.isscalar(tmax) else np.ascontiguousarray(tmax, np.float32)
        mesh = np.zeros(n, np.int32); tri = np.zeros(n, np.int32); pos = np.zeros(n, np.int32)
        u = np.zeros(n, np.float32); v = np.zeros(n, np.float32); t = np.zeros(n, np.float32)
        self.L.orc_intersect(self.h, n, _fptr(o), _fptr(d), _fptr(tmin), _fptr(tmax), _iptr(mesh), _iptr(tri),
                             _iptr(pos), _fptr(u), _fptr(v), _fptr(t))
        return {"mesh": mesh, "tri": tri, "pos": pos, "u": u, "v": v, "t": t}

    def intersect_uv(self, o, d):
        """(n, 2) texture coordinates of the closest hits of rays o + t d."""
        o = np.ascontiguousarray(o, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(d, np.float32).reshape(-1, 3)
        uv = np.zeros((o.shape[0], 2), np.float32)
        self.L.orc_intersect_uv(self.h, o.shape[0], _fptr(o), _fptr(d), _fptr(uv))
        return uv

    def mesh_bvh(self, mi):
        """(boxes (n,6), links (n,4) = left,right,tri_start,tri_count, tri_idx)."""
        n = C.c_int(0)
        self.L.orc_mesh_bvh(self.h, mi, C.byref(n), None, None, None)
        boxes = np.zeros((n.value, 6), np.float32)
        links = np.zeros((n.value, 4), np.int32)
        tri_idx = np.zeros(self._ntri[mi], np.int32)
        self.L.orc_mesh_bvh(self.h, mi, C.byref(n), _fptr(boxes), _iptr(links), _iptr(tri_idx))
        return boxes, links, tri_idx

    def material_eval(self, mesh: int, uv) -> np.ndarray:
        """(n, 10): albedo RGB, emission RGB, metallic, roughness,
        transmission, ior of mesh `mesh` at the (n, 2) texture coordinates."""
        uv = np.ascontiguousarray(np.asarray(uv, np.float32).reshape(-1, 2))
        out = np.zeros((uv.shape[0], 10), np.float32)
        if self.L.orc_material_eval(self.h, mesh, uv.shape[0], _fptr(uv), _fptr(out)) != 0:
            raise IndexError(mesh)
        return out

    def lights_info(self):
        n = self.L.orc_num_lights(self.h)
        em = np.zeros((n, 4), np.float32)
        for i in range(n):
            self.L.orc_light_emission(self.h, i, _fptr(em[i]))
        cdf = np.array([self.L.orc_light_cdf(self.h, i) for i in range(n)], np.float32)
        return {"n": n, "native_start": self.L.orc_native_light_start(self.h), "cdf": cdf,
                "emission_area": em, "total_area": self.L.orc_light_total_area(self.h)}

    def sample_light(self, i, p, u1, u2):
        out = np.zeros(11, np.float32)
        self.L.orc_sample_light(self.h, i, _fptr(np.asarray(p, np.float32)), u1, u2, _fptr(out))
        return out

    def select_light(self, u):
        p = np.zeros(1, np.float32)
        i = self.L.orc_select_light(self.h, u, _fptr(p))
        return i, float(p[0])


# --------------------------------------------------------------- diagnostics
# PixelPathData<16, 4> byte layout (diagnostics/pixel_path_data.hpp:145-160,
# path_stats.hpp:51-58 and 133-147, path_types.hpp:178-183).
STATS_DT = np.dtype([("count", "<u4"), ("mean", "<f4", 3), ("m2", "<f4", 3), ("variance", "<f4", 3)])
VERTEX_DT = np.dtype([("object_id", "<i4"), ("material_id", "<i2"), ("bsdf_type", "u1"), ("flags", "u1")])
GROUP_DT = np.dtype([("path_hash", "<u8"), ("stats", STATS_DT), ("depth", "u1"), ("coarse_type", "u1"),
                     ("light_type", "u1"), ("strategy", "u1"), ("pad", "u1", 4), ("vertices", VERTEX_DT, 16),
                     ("positions", "<f4", (16, 3)), ("normals", "<f4", (16, 3))])
OUTLIER_DT = np.dtype({"names": ["path_hash", "contribution"], "formats": ["<u8", ("<f4", 3)],
                       "offsets": [0, 8], "itemsize": 24})
PIXEL_DT = np.dtype([("groups", GROUP_DT, 16), ("coarse", STATS_DT, 8), ("outliers", OUTLIER_DT, 4),
                     ("num_groups", "<u8"), ("num_outliers", "<u8"), ("total_samples", "<u8"),
                     ("overflow_count", "<u8"), ("total", STATS_DT)])
assert GROUP_DT.itemsize == 568 and PIXEL_DT.itemsize == 9576


def diag_pixel_run(records, max_groups=16, max_outliers=4):
    """Feeds records [(vertices, contribution, light_type, strategy)], each
    vertex (object_id, material_id, bsdf_type, flags), through the oracle's
    PixelPathData<max_groups, max_outliers> and returns its state."""
    L = lib()
    n = len(records)
    depth = np.zeros(max(n, 1), np.int32)
    obj = np.full(max(n, 1) * 16, -1, np.int32)
    mat = np.full(max(n, 1) * 16, -1, np.int16)
    typ = np.zeros(max(n, 1) * 16, np.uint8)
    flg = np.zeros(max(n, 1) * 16, np.uint8)
    con = np.zeros(max(n, 1) * 3, np.float32)
    lty = np.zeros(max(n, 1), np.uint8)
    stg = np.zeros(max(n, 1), np.uint8)
    for i, (verts, c, lt, st) in enumerate(records):
        depth[i] = len(verts)
        for k, (o_, m_, t_, f_) in enumerate(verts[:16]):
            obj[i * 16 + k], mat[i * 16 + k], typ[i * 16 + k], flg[i * 16 + k] = o_, m_, t_, f_
        con[3 * i:3 * i + 3] = c
        lty[i], stg[i] = lt, st
    out = {"hash": np.zeros(max(n, 1), np.uint64), "coarse": np.zeros(max(n, 1), np.uint8),
           "summary": np.zeros(4, np.int64), "ghash": np.zeros(max_groups, np.uint64),
           "gcount": np.zeros(max_groups, np.uint32), "gmean": np.zeros(max_groups * 3, np.float32),
           "gvar": np.zeros(max_groups * 3, np.float32), "gdepth": np.zeros(max_groups, np.int32),
           "gtype0": np.zeros(max_groups, np.int32), "tmean": np.zeros(3, np.float32),
           "tvar": np.zeros(3, np.float32), "ccount": np.zeros(8, np.uint32), "cmean": np.zeros(24, np.float32),
           "bytes": np.zeros(PIXEL_DT.itemsize, np.uint8)}
    P = lambda a, t: a.ctypes.data_as(C.POINTER(t))
    L.orc_diag_pixel_run(max_groups, max_outliers, n, P(depth, C.c_int32), P(obj, C.c_int32), P(mat, C.c_int16),
                         P(typ, C.c_uint8), P(flg, C.c_uint8), P(con, C.c_float), P(lty, C.c_uint8),
                         P(stg, C.c_uint8), P(out["hash"], C.c_uint64), P(out["coarse"], C.c_uint8),
                         P(out["summary"], C.c_int64), P(out["ghash"], C.c_uint64), P(out["gcount"], C.c_uint32),
                         P(out["gmean"], C.c_float), P(out["gvar"], C.c_float), P(out["gdepth"], C.c_int32),
                         P(out["gtype0"], C.c_int32), P(out["tmean"], C.c_float), P(out["tvar"], C.c_float),
                         P(out["ccount"], C.c_uint32), P(out["cmean"], C.c_float),
                         P(out["bytes"], C.c_uint8) if (max_groups, max_outliers) == (16, 4) else None)
    out["hash"], out["coarse"] = out["hash"][:n], out["coarse"][:n]
    out["gmean"] = out["gmean"].reshape(-1, 3)
    out["gvar"] = out["gvar"].reshape(-1, 3)
    out["cmean"] = out["cmean"].reshape(-1, 3)
    return out


def flip_to_top_first(rgba_flat: np.ndarray, tw: int, th: int) -> np.ndarray:
    """(th, tw, 4) image with row 0 = top, from a bottom-row-first buffer."""
    return rgba_flat.reshape(th, tw, 4)[::-1]


def nth_element(keys, idx, nth, depth_limit=-1):
    """std::nth_element of the reference's BVH build on (keys, ids): both
    permuted as the library moves them.  depth_limit >= 0 runs the library's
    introselect with that partition budget."""
    k = np.ascontiguousarray(keys, np.float32).copy()
    i = np.ascontiguousarray(idx, np.int32).copy()
    lib().orc_nth_element(k.ctypes.data_as(C.POINTER(C.c_float)), i.ctypes.data_as(C.POINTER(C.c_int32)), len(k),
                          int(nth), int(depth_limit))
    return k, i
