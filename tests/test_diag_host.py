"""Host side of the DiagnosticFilm API (lucidrenderer module, no GPU needed).

The record logic in csrc/diag_film.h is the code the device film kernel runs;
here it runs on the host behind the reference's Python classes
(DiagnosticFilm, PathDiagnosticRecorder, DiagnosticExporter; pybind_module
.cpp:316-541) and must leave every film pixel byte for byte equal to the
oracle's independent restatement of PixelPathData<16, 4>, over random record
streams that exercise merging, eviction past 16 groups and the outlier ring.
Plus the reference's own checks of the export surface
(test_pyrenderer_diagnostics.cpp:110-277, test_diagnostic_export.cpp).
"""

import numpy as np
import pytest

import oracle

lr = pytest.importorskip("lucidrenderer")

B = lr.BsdfType
LT = lr.LightSourceType


def _stream(seed, n, vocab):
    """Random completed paths from a vocabulary of path shapes."""
    rng = np.random.default_rng(seed)
    shapes = []
    for _ in range(vocab):
        d = int(rng.integers(1, 6))
        verts = [(int(rng.integers(0, 5)), int(rng.integers(-1, 3)), int(rng.integers(0, 4)), int(rng.integers(0, 4)))
                 for _ in range(d)]
        end = int(rng.integers(0, 3))
        shapes.append((verts, end, int(rng.integers(0, 3))))
    out = []
    for _ in range(n):
        verts, end, lid = shapes[int(rng.integers(0, vocab))]
        scale = float(rng.choice([0.1, 1.0, 50.0], p=[0.45, 0.5, 0.05]))
        c = tuple(float(x) for x in (rng.random(3) * scale).astype(np.float32))
        out.append((verts, end, lid, c))
    return out


def _feed_product(film, recs):
    for verts, end, lid, c in recs:
        r = lr.PathDiagnosticRecorder()
        r.begin_path()
        for o, m, t, f in verts:
            r.record_vertex(o, m, B(t), bool(f & 1), bool(f & 2))
        if end == 0:
            r.record_environment_hit()
        elif end == 1:
            r.record_light_hit(lid, LT.Area)
        else:
            r.record_emissive_hit(lid, 0)
        film.record_path(0, 0, r.end_path(*c))


def _oracle_records(recs):
    out = []
    for verts, end, lid, c in recs:
        v = list(verts)
        if end == 0:
            v.append((-1, -1, 5, 0)); lt = 5
        elif end == 1:
            v.append((-(lid + 2), -1, 4, 0)); lt = 2
        else:
            v.append((lid, 0, 4, 0)); lt = 6
        out.append((v, c, lt, 0))
    return out


@pytest.mark.parametrize("seed,n,vocab", [(1, 40, 5), (2, 400, 30), (3, 2000, 60), (4, 25, 3)])
def test_film_pixel_bytes_match_oracle(seed, n, vocab):
    cfg = lr.PathRecordingConfig()
    cfg.subsample_factor = 1
    film = lr.DiagnosticFilm(1, 1, cfg)
    recs = _stream(seed, n, vocab)
    _feed_product(film, recs)
    ref = oracle.diag_pixel_run(_oracle_records(recs))
    got = np.frombuffer(film.pixel_bytes(0, 0), np.uint8)
    assert got.tobytes() == ref["bytes"].tobytes()
    px = got.view(oracle.PIXEL_DT)[0]
    assert px["total_samples"] == n and film.total_paths_recorded() == n
    if vocab > 16:
        assert px["overflow_count"] > 0


def test_film_layout_and_exports():
    cfg = lr.PathRecordingConfig.standard()
    assert cfg.subsample_factor == 2 and cfg.max_groups_per_pixel == 16
    assert lr.PathRecordingConfig.minimal().subsample_factor == 4
    assert lr.PathRecordingConfig.detailed().max_groups_per_pixel == 32
    # path_stats_config.hpp:140-170: pixels x (8 coarse x 40 + groups x 568 + outliers x 20 + 32) + 1 MiB
    assert cfg.estimate_memory_bytes(1920, 1080) == 960 * 540 * (320 + 16 * 568 + 4 * 20 + 32) + (1 << 20)
    tracer = lr.DiagnosticPathTracer(64, 64, cfg)
    film = tracer.film()
    assert (film.width(), film.height(), film.sampled_width(), film.sampled_height()) == (64, 64, 32, 32)
    rng = np.random.default_rng(0)
    for y in range(64):
        for x in range(64):
            r = lr.PathDiagnosticRecorder()
            r.begin_path()
            r.record_vertex(1, 0, B.Diffuse, False, False)
            if (x + y) % 3 == 0:
                r.record_vertex(2, 0, B.Glossy, False, False)
            r.record_light_hit(0, LT.Area)
            film.record_path(x, y, r.end_path(*(rng.random(3) * 2)))
    assert film.total_paths_recorded() == 4096
    st = film.global_stats()
    assert st.total_samples == 4096 and st.active_pixels == 32 * 32
    ex = lr.DiagnosticExporter(film)
    top = ex.get_top_variance_groups(5)
    assert 0 < len(top) <= 5
    assert all(top[i].variance_luminance >= 0 for i in range(len(top)))
    assert top[0].signature in ("DL", "DGL")
    js = ex.export_metadata_json()
    for k in ("width", "height", "total_samples", "subsample_factor"):
        assert f'"{k}"' in js
    b = ex.export_binary()
    assert b[:4] == b"PVAR" and len(b) == 4 + 4 * 5 + 4 + 32 * 32 * 22
    assert len(ex.get_coarse_stats()) == 8
    assert isinstance(ex.generate_suggestions(), list)
    film.clear()
    assert film.total_paths_recorded() == 0 and film.global_stats().total_samples == 0


def test_renderer_diagnostics_without_render():
    r = lr.Renderer()
    assert not r.is_diagnostics_enabled()
    r.enable_diagnostics(lr.PathRecordingConfig.detailed())
    assert r.is_diagnostics_enabled()
    s = r.get_diagnostic_stats()
    assert s.total_samples == 0 and s.active_pixels == 0
    assert r.get_diagnostic_variance_map().size == 0
    assert r.export_diagnostic_json() == "{}"
    assert r.get_top_variance_groups(5) == []
    assert not r.get_pixel_diagnostic(0, 0).valid
    r.disable_diagnostics()
    assert not r.is_diagnostics_enabled()


def test_object_paths():
    r = lr.Renderer()
    r.set_object_names(["Plane", "Sphere"])
    assert r.get_object_name(0) == "Plane" and r.get_object_name(7) == "Object_7"
    assert r.get_object_name(-1) == "Environment"
    assert r.object_path_string([]) == "Camera"
    assert r.object_path_string([0, 1]) == "(terminated) → Sphere → Plane → Camera"
    assert r.object_path_string([0, -3]) == "Light(?) → Plane → Camera"
    assert r.object_path_string([1, -1]) == "Environment(?) → Sphere → Camera"
