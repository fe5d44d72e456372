"""ctypes binding of liblucid_hip.so (the C ABI declared in include/lucid_hip.h).

This is the Python-side counterpart of the pybind11 module: it talks to the
library through plain pointers and POD structs only, which is how a non-C++
host (or a benchmark that keeps its buffers in device memory) binds the path.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

ALGOS = {"simple": 0, "nee": 1, "mis": 2, "volume_simple": 3, "volume_mis": 4, "guided": 5, "guided_mis": 5}
LIGHT_TYPES = {"POINT": 1, "SUN": 2, "SPOT": 3, "AREA": 4}
SHAPES = {"SQUARE": 0, "RECTANGLE": 1, "DISK": 2, "ELLIPSE": 3}
ERRORS = {0: "LUCID_OK", 1: "LUCID_ERR_INVALID", 2: "LUCID_ERR_DEVICE", 3: "LUCID_ERR_CANCELLED",
          4: "LUCID_ERR_NOMEM"}
PROFILE_WORK, PROFILE_KERNELS = 1, 2

F3 = C.c_float * 3
I3 = C.c_int32 * 3


SOCK_NONE, SOCK_FLOAT, SOCK_VEC3, SOCK_VEC4 = 0, 1, 2, 3


class NodeSocket(C.Structure):
    _fields_ = [("name", C.c_char_p), ("kind", C.c_int32), ("value", C.c_float * 4), ("linked_node", C.c_char_p),
                ("linked_socket", C.c_char_p)]


class Node(C.Structure):
    _fields_ = [("name", C.c_char_p), ("type", C.c_char_p), ("inputs", C.POINTER(NodeSocket)),
                ("num_inputs", C.c_int32), ("outputs", C.POINTER(NodeSocket)), ("num_outputs", C.c_int32)]


class MeshDesc(C.Structure):
    _fields_ = [("num_vertices", C.c_int32), ("vertices", C.POINTER(C.c_float)), ("num_triangles", C.c_int32),
                ("indices", C.POINTER(C.c_int32)), ("tri_normals", C.POINTER(C.c_float)),
                ("num_tri_normals", C.c_int32), ("smooth", C.c_int32), ("base_color", F3), ("metallic", C.c_float),
                ("roughness", C.c_float), ("emission", F3), ("transmission", C.c_float), ("ior", C.c_float),
                ("has_volume", C.c_int32), ("vol_color", F3), ("vol_absorption", F3), ("vol_density", C.c_float),
                ("vol_anisotropy", C.c_float), ("grid_dims", I3), ("grid_min", F3), ("grid_max", F3),
                ("grid", C.POINTER(C.c_float)), ("surface_linked", C.c_int32), ("use_nodes", C.c_int32),
                ("nodes", C.POINTER(Node)), ("num_nodes", C.c_int32), ("tri_uvs", C.POINTER(C.c_float)),
                ("num_tri_uvs", C.c_int32)]


class LightDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("position", F3), ("has_direction", C.c_int32), ("direction", F3),
                ("color", F3), ("energy", C.c_float), ("radius", C.c_float), ("spot_size", C.c_float),
                ("spot_blend", C.c_float), ("size", C.c_float), ("size_y", C.c_float), ("shape", C.c_int32),
                ("has_right", C.c_int32), ("has_up", C.c_int32), ("right", F3), ("up", F3)]


class Camera(C.Structure):
    _fields_ = [("pos", F3), ("dir", F3), ("up", F3), ("fov_deg", C.c_float), ("exposure_ev", C.c_float)]


class RenderParams(C.Structure):
    _fields_ = [("tile_x", C.c_int32), ("tile_y", C.c_int32), ("tile_w", C.c_int32), ("tile_h", C.c_int32),
                ("full_w", C.c_int32), ("full_h", C.c_int32), ("samples", C.c_int32), ("sample_offset", C.c_int32),
                ("max_depth", C.c_int32), ("algorithm", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("rays_closest", C.c_uint64), ("rays_shadow", C.c_uint64), ("paths", C.c_uint64),
                ("waves", C.c_uint64), ("stack_overflows", C.c_uint64), ("ms_render", C.c_double),
                ("nodes_closest", C.c_uint64), ("tris_closest", C.c_uint64), ("nodes_shadow", C.c_uint64),
                ("tris_shadow", C.c_uint64), ("ms_intersect", C.c_double), ("ms_shade", C.c_double),
                ("ms_shadow", C.c_double), ("launches_intersect", C.c_uint64), ("launches_shade", C.c_uint64),
                ("launches_shadow", C.c_uint64), ("ms_guide_infer", C.c_double), ("ms_guide_train", C.c_double),
                ("ms_ray_binning", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# name -> (restype, argtypes); every symbol include/lucid_hip.h declares
class DiagConfig(C.Structure):
    _fields_ = [("enabled", C.c_int32), ("subsample", C.c_int32)]


class DiagInfo(C.Structure):
    _fields_ = [("valid", C.c_int32), ("width", C.c_int32), ("height", C.c_int32), ("sampled_width", C.c_int32),
                ("sampled_height", C.c_int32), ("subsample", C.c_int32), ("total_recorded", C.c_int64),
                ("pixel_bytes", C.c_int64)]


class GuidingConfig(C.Structure):
    _fields_ = [("alpha", C.c_float), ("train_steps", C.c_int32), ("batch", C.c_int32),
                ("learning_rate", C.c_float), ("record_stride", C.c_int32), ("seed", C.c_uint32),
                ("max_depth", C.c_int32), ("learn_selection", C.c_int32)]


class GuidingStats(C.Structure):
    _fields_ = [("samples_total", C.c_int64), ("train_steps", C.c_int64), ("last_loss", C.c_float),
                ("active", C.c_int32), ("num_params", C.c_int32)]


class LaunchShape(C.Structure):
    _fields_ = [("traverse_groups", C.c_int32), ("refill", C.c_int32), ("shade_groups", C.c_int32),
                ("ray_order", C.c_int32), ("shade_order", C.c_int32), ("ms", C.c_double),
                ("wave_paths", C.c_int64), ("walk_min", C.c_int32), ("leaf_tris", C.c_int32)]


SIGNATURES = {
    "lucid_create": (C.c_int, [C.c_int32, C.POINTER(C.c_void_p)]),
    "lucid_destroy": (None, [C.c_void_p]),
    "lucid_last_error": (C.c_char_p, [C.c_void_p]),
    "lucid_scene_create": (C.c_int, [C.c_void_p, C.POINTER(MeshDesc), C.c_int32, C.POINTER(LightDesc), C.c_int32,
                                     C.POINTER(C.c_float), C.c_float, C.POINTER(C.c_void_p)]),
    "lucid_scene_retain": (None, [C.c_void_p]),
    "lucid_scene_release": (None, [C.c_void_p]),
    "lucid_scene_num_triangles": (C.c_int32, [C.c_void_p]),
    "lucid_render_tile": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Camera), C.POINTER(RenderParams),
                                    C.POINTER(C.c_float)]),
    "lucid_render_tile_device": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Camera), C.POINTER(RenderParams),
                                           C.c_void_p, C.c_void_p]),
    "lucid_render_tile_device_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Camera),
                                                   C.POINTER(RenderParams), C.c_void_p, C.c_void_p]),
    "lucid_synchronize": (C.c_int, [C.c_void_p]),
    "lucid_render_samples": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Camera), C.POINTER(RenderParams),
                                       C.POINTER(C.c_float)]),
    "lucid_primary_hits": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Camera), C.POINTER(RenderParams),
                                     C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_float),
                                     C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "lucid_render_debug": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Camera), C.POINTER(RenderParams), C.c_int32,
                                     C.POINTER(C.c_float)]),
    "lucid_cancel": (None, [C.c_void_p]),
    "lucid_is_cancelled": (C.c_int32, [C.c_void_p]),
    "lucid_reset_cancel": (None, [C.c_void_p]),
    "lucid_get_stats": (C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    "lucid_set_wave_paths": (C.c_int, [C.c_void_p, C.c_int64]),
    "lucid_set_launch_shape": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "lucid_set_refill_threshold": (C.c_int, [C.c_void_p, C.c_int32]),
    "lucid_set_ray_order": (C.c_int, [C.c_void_p, C.c_int32]),
    "lucid_set_shade_order": (C.c_int, [C.c_void_p, C.c_int32]),
    "lucid_set_session_graphs": (C.c_int, [C.c_void_p, C.c_int32]),
    "lucid_set_session_fusion": (C.c_int, [C.c_void_p, C.c_int32]),
    "lucid_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Camera), C.POINTER(RenderParams),
                                 C.POINTER(LaunchShape)]),
    "lucid_set_diagnostics": (C.c_int, [C.c_void_p, C.POINTER(DiagConfig)]),
    "lucid_diag_get_info": (C.c_int, [C.c_void_p, C.POINTER(DiagInfo)]),
    "lucid_diag_read_pixels": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_void_p]),
    "lucid_diag_write_pixels": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_void_p]),
    "lucid_diag_read_strategy": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.POINTER(C.c_double)]),
    "lucid_diag_clear": (C.c_int, [C.c_void_p]),
    "lucid_set_guiding": (C.c_int, [C.c_void_p, C.POINTER(GuidingConfig)]),
    "lucid_guiding_reset": (C.c_int, [C.c_void_p]),
    "lucid_guiding_get_stats": (C.c_int, [C.c_void_p, C.POINTER(GuidingStats)]),
    "lucid_guiding_get_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_int32]),
    "lucid_guiding_set_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_int32]),
    "lucid_guiding_debug_infer": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int32,
                                            C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "lucid_guiding_debug_render_infer": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_float), C.c_int32,
                                                   C.POINTER(C.c_float), C.POINTER(C.c_int32), C.c_int32,
                                                   C.POINTER(C.c_float), C.POINTER(C.c_float),
                                                   C.POINTER(C.c_float)]),
    "lucid_guiding_debug_step": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int32,
                                           C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                           C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "lucid_set_profiling": (C.c_int, [C.c_void_p, C.c_int32]),
    "lucid_set_grid_format": (C.c_int, [C.c_void_p, C.c_int32]),
    "lucid_scene_grid_bytes": (C.c_int64, [C.c_void_p]),
    "lucid_set_scene_build": (C.c_int, [C.c_void_p, C.c_int32]),
    "lucid_set_traversal_shape": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "lucid_scene_read_geometry": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_float),
                                            C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                            C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "lucid_debug_select_nth": (C.c_int, [C.POINTER(C.c_float), C.POINTER(C.c_int32), C.c_int32, C.c_int32,
                                         C.c_int32]),
    "lucid_session_create": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Camera), C.c_int32, C.c_int32, C.c_int32,
                                       C.c_int32, C.POINTER(C.c_void_p)]),
    "lucid_session_destroy": (None, [C.c_void_p]),
    "lucid_session_set_camera": (C.c_int, [C.c_void_p, C.POINTER(Camera), C.c_int32]),
    "lucid_session_clear": (C.c_int, [C.c_void_p]),
    "lucid_session_samples": (C.c_int64, [C.c_void_p]),
    "lucid_sessions_render": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_int32, C.c_int32]),
    "lucid_session_read": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_int32, C.POINTER(C.c_int64)]),
    "lucid_session_start_async": (C.c_int, [C.c_void_p]),
    "lucid_session_stop_async": (C.c_int, [C.c_void_p]),
    "lucid_session_reset_async": (C.c_int, [C.c_void_p, C.POINTER(Camera), C.c_int32, C.c_int32, C.c_int32]),
    "lucid_session_poll_async": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_int64),
                                           C.POINTER(C.c_uint32)]),
    "lucid_session_async_running": (C.c_int32, [C.c_void_p]),
    "lucid_session_error": (C.c_char_p, [C.c_void_p]),
    "lucid_material_eval": (C.c_int, [C.POINTER(MeshDesc), C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                      C.c_char_p, C.c_int32]),
    "lucid_debug_host_trace": (C.c_int, [C.POINTER(MeshDesc), C.c_int32, C.c_int32, C.POINTER(C.c_float),
                                         C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int32,
                                         C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_float),
                                         C.POINTER(C.c_float), C.POINTER(C.c_float)]),
}

_lib = None


def load(path: str | None = None) -> C.CDLL:
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or os.path.join(_HERE, "liblucid_hip.so")
    if not os.path.exists(p):
        raise FileNotFoundError(f"{p} is not built: run `make -C blender-lucid-renderer_amd`")
    L = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = L
    return L


def _f3(v, default=(0.0, 0.0, 0.0)):
    v = default if v is None else v
    return F3(*(float(x) for x in v))


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


def _socket_value(v, sock: NodeSocket):
    """parseSocketValue (node_evaluator.cpp:47-69): a number is FLOAT, a
    numeric 3- or 4-array VEC3 / VEC4, anything else NONE."""
    sock.kind = SOCK_NONE
    if isinstance(v, bool) or v is None:
        return
    if isinstance(v, (int, float)):
        sock.kind, sock.value[0] = SOCK_FLOAT, float(v)
    elif isinstance(v, (list, tuple)) and len(v) in (3, 4) and all(
            isinstance(x, (int, float)) and not isinstance(x, bool) for x in v):
        sock.kind = SOCK_VEC3 if len(v) == 3 else SOCK_VEC4
        for k, x in enumerate(v):
            sock.value[k] = float(x)


def node_tree(material: dict, keep: list):
    """parseNodeTree (node_evaluator.cpp:82-149) of a material dict into a
    Node array: (array, count, use_nodes).  Each input socket is linked from
    the last "links" entry that targets it; useNodes = "use_nodes" and a tree
    with nodes (pybind_renderer.hpp:1381-1384).  `keep` receives every
    buffer the array points into."""
    nt = material.get("node_tree")
    if not isinstance(nt, dict) or not isinstance(nt.get("nodes"), list) or not nt["nodes"]:
        return None, 0, 0
    conn = {}
    for l in nt.get("links") or []:
        conn[(l["to_node"], l["to_socket"])] = (l["from_node"], l["from_socket"])

    def b(x):
        bs = str(x).encode()
        keep.append(bs)
        return bs

    nodes = (Node * len(nt["nodes"]))()
    keep.append(nodes)
    for i, nj in enumerate(nt["nodes"]):
        name = nj.get("name", "") if isinstance(nj.get("name", ""), str) else ""
        ntype = nj.get("type", "") if isinstance(nj.get("type", ""), str) else ""
        nodes[i].name, nodes[i].type = b(name), b(ntype)
        for field, count in (("inputs", "num_inputs"), ("outputs", "num_outputs")):
            socks = nj.get(field) if isinstance(nj.get(field), list) else []
            arr = (NodeSocket * max(1, len(socks)))()
            keep.append(arr)
            for k, sj in enumerate(socks):
                sn = sj.get("name", "") if isinstance(sj.get("name", ""), str) else ""
                arr[k].name = b(sn)
                _socket_value(sj.get("default_value"), arr[k])
                if field == "inputs" and (name, sn) in conn:
                    fn, fs = conn[(name, sn)]
                    arr[k].linked_node, arr[k].linked_socket = b(fn), b(fs)
            setattr(nodes[i], field, C.cast(arr, C.POINTER(NodeSocket)))
            setattr(nodes[i], count, len(socks))
    return nodes, len(nt["nodes"]), 1 if material.get("use_nodes", False) else 0


def material_eval(mesh_or_material: dict, uv) -> np.ndarray:
    """lucid_material_eval: (n, 10) floats -- albedo RGB, emission RGB,
    metallic, roughness, transmission, ior -- of a mesh's material (or a bare
    material dict) at the (n, 2) texture coordinates uv.  Host only."""
    mesh = mesh_or_material if "material" in mesh_or_material else {"material": mesh_or_material}
    d = SceneDescs({"meshes": [dict(mesh, vertices=[], triangles=[])]})
    uv = np.ascontiguousarray(np.asarray(uv, np.float32).reshape(-1, 2))
    out = np.zeros((uv.shape[0], 10), np.float32)
    err = C.create_string_buffer(256)
    rc = load().lucid_material_eval(C.byref(d.meshes[0]), uv.shape[0], uv.ctypes.data_as(C.POINTER(C.c_float)),
                                    out.ctypes.data_as(C.POINTER(C.c_float)), err, 256)
    if rc != 0:
        raise ValueError(err.value.decode())
    return out


class SceneDescs:
    """POD descriptors of a reference-schema scene dict, owning the arrays."""

    def __init__(self, scene: dict):
        self._keep = []
        meshes = scene.get("meshes", [])
        self.meshes = (MeshDesc * max(1, len(meshes)))()
        self.n_meshes = len(meshes)
        for i, m in enumerate(meshes):
            d = self.meshes[i]
            v = np.ascontiguousarray(np.asarray(m.get("vertices", []), np.float32).reshape(-1, 3))
            t = np.ascontiguousarray(np.asarray(m.get("triangles", []), np.int32).reshape(-1, 3))
            self._keep += [v, t]
            d.num_vertices, d.vertices = v.shape[0], v.ctypes.data_as(C.POINTER(C.c_float))
            d.num_triangles, d.indices = t.shape[0], t.ctypes.data_as(C.POINTER(C.c_int32))
            tn = m.get("triangle_normals")
            if tn is not None and len(tn):
                tn = np.ascontiguousarray(np.asarray(tn, np.float32).reshape(-1, 9))
                self._keep.append(tn)
                d.tri_normals, d.num_tri_normals = tn.ctypes.data_as(C.POINTER(C.c_float)), tn.shape[0]
            d.smooth = 1 if m.get("smooth") else 0
            mat = m.get("material") or {}
            d.base_color = _f3(mat.get("base_color"), (0.8, 0.8, 0.8))
            d.metallic = float(mat.get("metallic", 0.0))
            d.roughness = float(mat.get("roughness", 0.5))
            d.emission = _f3(mat.get("emission"))
            d.transmission = float(mat.get("transmission", 0.0))
            d.ior = float(mat.get("ior", 1.45))
            d.surface_linked = 1 if surface_linked(mat) else 0
            nodes, n_nodes, use = node_tree(mat, self._keep)
            if n_nodes:
                d.nodes, d.num_nodes, d.use_nodes = nodes, n_nodes, use
            tu = m.get("triangle_uvs")
            if tu is not None and len(tu):
                tu = np.ascontiguousarray(np.asarray(tu, np.float32).reshape(-1, 6))
                self._keep.append(tu)
                d.tri_uvs, d.num_tri_uvs = tu.ctypes.data_as(C.POINTER(C.c_float)), tu.shape[0]
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
                    g = np.fromfile(sg["sidecar_path"], np.float32, count=dims[0] * dims[1] * dims[2])
                    self._keep.append(g)
                    d.grid_dims, d.grid_min, d.grid_max = I3(*dims), _f3(sg["world_min"]), _f3(sg["world_max"])
                    d.grid = g.ctypes.data_as(C.POINTER(C.c_float))
        lights = scene.get("native_lights", scene.get("lights")) or []
        self.lights = (LightDesc * max(1, len(lights)))()
        self.n_lights = len(lights)
        for i, l in enumerate(lights):
            d = self.lights[i]
            d.type = LIGHT_TYPES[l.get("type", "POINT")]
            d.position = _f3(l.get("position"))
            if "direction" in l:
                d.has_direction, d.direction = 1, _f3(l["direction"])
            d.color = _f3(l.get("color"), (1.0, 1.0, 1.0))
            d.energy = float(l.get("energy", 1.0))
            d.radius = float(l.get("radius", 0.0))
            d.spot_size = float(l.get("spot_size", np.float32(np.pi / 4.0)))
            d.spot_blend = float(l.get("spot_blend", 0.0))
            d.size = float(l.get("size", 1.0))
            d.size_y = float(l.get("size_y", d.size))
            d.shape = SHAPES.get(l.get("shape", "SQUARE"), 0)
            if "right" in l:
                d.has_right, d.right = 1, _f3(l["right"])
            if "up" in l:
                d.has_up, d.up = 1, _f3(l["up"])
        env = scene.get("environment") or {}
        self.env = _f3(env.get("color"), (0.05, 0.05, 0.05))
        self.env_strength = float(env.get("strength", 1.0))


class Context:
    """Owns a lucid_ctx; raises RuntimeError on any non-OK status."""

    def __init__(self, device: int = 0):
        self.L = load()
        h = C.c_void_p()
        rc = self.L.lucid_create(device, C.byref(h))
        if rc != 0:
            raise RuntimeError(f"lucid_create failed: {ERRORS.get(rc, rc)}")
        self.h = h

    def check(self, rc, ok=(0,)):
        if rc not in ok:
            raise RuntimeError(f"{ERRORS.get(rc, rc)}: {self.L.lucid_last_error(self.h).decode()}")
        return rc

    def scene(self, scene: dict):
        d = SceneDescs(scene)
        out = C.c_void_p()
        self.check(self.L.lucid_scene_create(self.h, d.meshes, d.n_meshes, d.lights, d.n_lights, d.env,
                                             d.env_strength, C.byref(out)))
        return out

    def stats(self) -> dict:
        s = Stats()
        self.check(self.L.lucid_get_stats(self.h, C.byref(s)))
        return s.as_dict()

    def close(self):
        if getattr(self, "h", None):
            self.L.lucid_destroy(self.h)
            self.h = None


class Session:
    """A lucid_session: progressive accumulation of one frame of one scene."""

    def __init__(self, ctx: Context, scene_handle, cam: Camera, width: int, height: int, max_depth: int,
                 algorithm: str = "mis"):
        self.ctx, self.width, self.height = ctx, int(width), int(height)
        h = C.c_void_p()
        ctx.check(ctx.L.lucid_session_create(ctx.h, scene_handle, C.byref(cam), self.width, self.height, int(max_depth),
                                             ALGOS.get(algorithm, 1), C.byref(h)))
        self.h = h

    @property
    def samples(self) -> int:
        return int(self.ctx.L.lucid_session_samples(self.h))

    def read(self, average: bool = False) -> np.ndarray:
        out = np.zeros(self.width * self.height * 4, np.float32)
        n = C.c_int64()
        self.ctx.check(self.ctx.L.lucid_session_read(self.h, out.ctypes.data_as(C.POINTER(C.c_float)),
                                                     1 if average else 0, C.byref(n)))
        return out

    def poll(self):
        """(samples, pixels or None, revision) of the async snapshot."""
        n, rev = C.c_int64(), C.c_uint32()
        self.ctx.check(self.ctx.L.lucid_session_poll_async(self.h, None, C.byref(n), C.byref(rev)))
        if n.value <= 0:
            return 0, None, rev.value
        out = np.zeros(self.width * self.height * 4, np.float32)
        self.ctx.check(self.ctx.L.lucid_session_poll_async(self.h, out.ctypes.data_as(C.POINTER(C.c_float)),
                                                           C.byref(n), C.byref(rev)))
        return n.value, out, rev.value

    def close(self):
        if getattr(self, "h", None):
            self.ctx.L.lucid_session_destroy(self.h)
            self.h = None


def render_sessions(ctx: Context, sessions, samples: int):
    """One batched pass: `samples` more samples in every session."""
    arr = (C.c_void_p * len(sessions))(*[s.h.value for s in sessions])
    ctx.check(ctx.L.lucid_sessions_render(ctx.h, arr, len(sessions), int(samples)))


def host_trace(scene: dict, o, d, tmin, tmax, any_hit=False):
    """The device traversal code executed on the CPU (lucid_debug_host_trace)."""
    L = load()
    sd = SceneDescs(scene)
    o = np.ascontiguousarray(o, np.float32).reshape(-1, 3)
    d = np.ascontiguousarray(d, np.float32).reshape(-1, 3)
    n = o.shape[0]
    tmin = np.ascontiguousarray(np.broadcast_to(np.float32(tmin), (n,)) if np.isscalar(tmin) else tmin, np.float32)
    tmax = np.ascontiguousarray(np.broadcast_to(np.float32(tmax), (n,)) if np.isscalar(tmax) else tmax, np.float32)
    out = {k: np.zeros(n, np.int32) for k in ("mesh", "tri")}
    out.update({k: np.zeros(n, np.float32) for k in ("u", "v", "t")})
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
    rc = L.lucid_debug_host_trace(sd.meshes, sd.n_meshes, n, fp(o), fp(d), fp(tmin), fp(tmax), 1 if any_hit else 0,
                                  ip(out["mesh"]), ip(out["tri"]), fp(out["u"]), fp(out["v"]), fp(out["t"]))
    if rc != 0:
        raise RuntimeError(f"lucid_debug_host_trace failed: {ERRORS.get(rc, rc)}")
    return out


def camera(cam: dict, exposure_ev: float = 0.0) -> Camera:
    return Camera(_f3(cam["pos"]), _f3(cam["dir"]), _f3(cam["up"]), float(cam["fov"]), float(exposure_ev))


def params(tx, ty, tw, th, fw, fh, samples, offset, max_depth, algorithm) -> RenderParams:
    return RenderParams(tx, ty, tw, th, fw, fh, samples, offset, max_depth, ALGOS.get(algorithm, 1))
