"""The drop-in boundary on the CPU: the C-ABI library loads and exports every
symbol include/lucid_hip.h declares, and the lucidrenderer module mirrors the
reference's Renderer API (src/pybind_module.cpp:41-312) in everything that
does not need a GPU.  No compute call is made here."""

import os
import re

import numpy as np
import pytest

import lucid_capi
import workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lucid_hip.h")

REFERENCE_METHODS = [
    "cancel", "is_cancelled", "reset_cancel", "load_scene_json", "set_camera", "set_algorithm",
    "set_object_names", "get_object_name", "object_path_string", "render_tile", "render_tile_gpu",
    "render_tile_numpy", "is_scene_loaded", "is_camera_set", "get_algorithm", "get_mesh_count",
    "enable_diagnostics", "disable_diagnostics", "is_diagnostics_enabled", "clear_diagnostics",
    "get_diagnostic_stats", "get_diagnostic_variance_map", "export_diagnostic_json", "get_top_variance_groups",
    "get_pixel_diagnostic", "render_start_async", "render_poll_async", "render_reset_async", "render_stop_async",
    "render_snapshot_revision_async", "is_render_async_running",
]
REFERENCE_CLASSES = [
    "Renderer", "PathRecordingConfig", "ExportedGroupInfo", "TopGroupSortBy", "PixelDiagnosticInfo",
    "DiagnosticSuggestion", "GlobalDiagnosticStats", "ExportedCoarseStats", "DiagnosticFilm", "DiagnosticPathTracer",
    "DiagnosticExporter", "BsdfType", "LightSourceType", "PathTrace", "PathDiagnosticRecorder",
]


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lucid_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = lucid_capi.load()
    names = declared_functions()
    assert len(names) >= 18
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(lucid_capi.SIGNATURES) == names


def test_capi_rejects_bad_arguments_without_device():
    L = lucid_capi.load()
    import ctypes as C
    assert L.lucid_create(0, None) == 1                       # null out pointer
    assert L.lucid_set_wave_paths(None, 5) == 1
    assert L.lucid_set_profiling(None, 1) == 1
    assert L.lucid_get_stats(None, None) == 1
    assert L.lucid_last_error(None) == b"null context"
    assert L.lucid_scene_num_triangles(None) == 0
    L.lucid_cancel(None)
    L.lucid_scene_release(None)
    # sessions
    assert L.lucid_session_create(None, None, None, 4, 4, 8, 2, None) == 1
    assert L.lucid_sessions_render(None, None, 0, 1) == 1
    assert L.lucid_session_poll_async(None, None, None, None) == 1
    assert L.lucid_session_read(None, None, 1, None) == 1
    assert L.lucid_session_start_async(None) == 1 and L.lucid_session_stop_async(None) == 1
    assert L.lucid_session_reset_async(None, None, 4, 4, 8) == 1
    assert L.lucid_session_samples(None) == 0 and L.lucid_session_async_running(None) == 0
    assert L.lucid_session_error(None) == b"null session"
    L.lucid_session_destroy(None)
    assert L.lucid_set_grid_format(None, 16) == 1 and L.lucid_scene_grid_bytes(None) == 0
    assert L.lucid_set_refill_threshold(None, 48) == 1
    assert L.lucid_set_launch_shape(None, 16, 8) == 1
    assert L.lucid_set_ray_order(None, 1) == 1
    uv = np.zeros(2, np.float32)
    out = np.zeros(10, np.float32)
    assert L.lucid_material_eval(None, 1, uv.ctypes.data_as(C.POINTER(C.c_float)),
                                 out.ctypes.data_as(C.POINTER(C.c_float)), None, 0) == 1


def test_async_api_without_scene_stays_idle(lucid):
    """render_start_async without a scene or camera returns and stays idle
    (pybind_renderer.hpp:689-715); poll gives (0, empty)."""
    r = lucid.Renderer()
    r.render_start_async(64, 48, 8)
    assert r.is_render_async_running() is False
    n, px = r.render_poll_async()
    assert n == 0 and px.size == 0
    assert r.render_snapshot_revision_async() == 0
    r.render_reset_async(64, 48, 8)
    r.render_stop_async()
    assert r.is_render_async_running() is False


def test_module_surface(lucid):
    r = lucid.Renderer()
    for m in REFERENCE_METHODS:
        assert hasattr(r, m), m
    for c in REFERENCE_CLASSES:  # pybind_module.cpp:41-541
        assert hasattr(lucid, c), c
    assert lucid.__version__ == "1.0.0"
    assert lucid.openmp_enabled is False and lucid.dawn_enabled is False
    assert r.get_algorithm() == "nee"                          # C++ default (pybind_renderer.hpp:68)
    r.set_algorithm("mis")
    assert r.get_algorithm() == "mis"
    assert "scene_loaded=False" in repr(r)


def test_render_without_scene_or_camera_returns_zeros(lucid):
    r = lucid.Renderer()
    px = r.render_tile(0, 0, 4, 3, 4, 3)
    assert px == [0.0] * (4 * 3 * 4)
    assert r.load_scene_json(workloads.to_json(workloads.closed_box()))
    px = r.render_tile(0, 0, 4, 3, 4, 3, 2, 0, 8)              # camera not set
    assert px == [0.0] * 48
    a = r.render_tile_numpy(0, 0, 5, 2, 5, 2)
    assert a.shape == (2, 5, 4) and not a.any()


def test_load_scene_json_behaviour(lucid):
    r = lucid.Renderer()
    assert r.load_scene_json(workloads.to_json(workloads.closed_box()))
    assert r.is_scene_loaded() and r.get_mesh_count() == 7
    assert not r.load_scene_json("{not json")
    assert not r.is_scene_loaded() and r.get_mesh_count() == 0
    assert r.load_scene_json('{"native_lights": []}')          # no "meshes": empty scene
    assert r.get_mesh_count() == 0
    bad = {"meshes": [{"vertices": [[0, 0, 0]], "triangles": [[0, 1, 2]]}]}
    assert not r.load_scene_json(workloads.to_json(bad))      # index out of range
    ok = {"meshes": [{"vertices": [[0, 0, 0], [1, 0, 0], [0, 1, 0]], "triangles": [[0, 1, 2]],
                      "triangle_normals": [[[0, 0, 1], [0, 0, 1], [0, 0, 1]]], "smooth": True,
                      "triangle_uvs": [[[0, 0], [1, 0], [0, 1]]],
                      "material": {"base_color": [1, 0, 0], "node_tree": None, "volume": None}}],
          "lights": [{"type": "SUN", "direction": [0, -1, 0], "energy": 2.0, "name": "sé"}],
          "environment": {"color": [0.1, 0.2, 0.3], "strength": 0.5}}
    assert r.load_scene_json(workloads.to_json(ok))
    assert r.get_mesh_count() == 1


def test_object_names(lucid):
    r = lucid.Renderer()
    r.set_object_names(["Floor", "Box"])
    assert r.get_object_name(-1) == "Environment"
    assert r.get_object_name(1) == "Box"
    assert r.get_object_name(7) == "Object_7"
    assert r.object_path_string([]) == "Camera"
    assert r.object_path_string([0, 1]) == "(terminated) → Box → Floor → Camera"
    assert r.object_path_string([0, -1]) == "Environment(?) → Floor → Camera"
    assert r.object_path_string([1, -3]) == "Light(?) → Box → Camera"


def test_render_fails_loudly_without_device(lucid):
    """With a scene and camera but no HIP device there is no CPU fallback."""
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    r = lucid.Renderer()
    assert r.load_scene_json(workloads.to_json(workloads.closed_box()))
    c = workloads.CLOSED_BOX_CAMERA
    r.set_camera(*c["pos"], *c["dir"], *c["up"], c["fov"])
    with pytest.raises(RuntimeError, match="no usable HIP device"):
        r.render_tile(0, 0, 4, 4, 4, 4, 1, 0, 2)
