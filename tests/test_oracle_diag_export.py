"""DiagnosticFilm exports: the oracle's restatement (oracle/diag_export.py)
pinned on its own, then the product's DiagnosticExporter compared with it.

Pins of the restatement (films built by the oracle's PixelPathData<16, 4>,
oracle.diag_pixel_run, never by the product):
* the reference's export tests (test_diagnostic_export.cpp): top variance
  group of 10 glossy samples is a 'G' path whose variance is the sample
  variance of its contributions (numpy, float64); the metadata JSON of a
  1920 x 1080 standard film names width 1920 and height 1080; PVAR starts
  with its magic; 8 coarse types; five alternating-contribution glossy path
  types give a glossy-variance suggestion;
* PVAR decoded field by field with struct against the film's own pixels,
  length 28 + 22 per sampled pixel;
* top groups against a brute-force sort of every candidate (numpy lexsort);
* the coarse merge against the count-weighted mean in float64, and its
  fixed pixel_data(2 sx, 2 sy) lookup at subsample 1, 2 and 4.

Product against restatement (host films filled through the reference's
Python classes): PVAR bytes and JSON text equal, top groups equal up to the
order of equal keys, coarse statistics and suggestions equal bit for bit.
"""

import json
import struct
from collections import Counter

import numpy as np
import pytest

import diag_export as dx
import oracle

F = np.float32


def _pixel(records):
    return oracle.diag_pixel_run(records)["bytes"].view(oracle.PIXEL_DT)[0]


def _film(pixels, sh, sw):
    film = np.zeros((sh, sw), oracle.PIXEL_DT)
    empty = _pixel([])
    for y in range(sh):
        for x in range(sw):
            film[y, x] = pixels.get((y, x), empty)
    return film


def _glossy(c):
    return ([(0, 0, 1, 0)], (float(c[0]), float(c[1]), float(c[2])), 2, 0)


def test_top_variance_group_of_glossy_samples():
    # test_diagnostic_export.cpp:36-53
    contrib = [(5.0 * i, 1.0, 1.0) for i in range(10)]
    film = _film({(0, 0): _pixel([_glossy(c) for c in contrib])}, 25, 25)
    top, last, _ = dx.top_groups(film, 5, 4)
    assert len(top) == 1
    x, y, g = top[0]
    info = dx.group_info(x, y, g)
    assert "G" in info["signature"] and info["sample_count"] == 10
    want = np.var(np.array(contrib, np.float64), axis=0, ddof=1)
    np.testing.assert_allclose(info["variance_rgb"], want, rtol=1e-5, atol=1e-6)
    assert abs(float(last) - want.sum()) <= 1e-5 * want.sum()
    assert abs(float(info["variance_luminance"]) - float((np.array([0.2126, 0.7152, 0.0722]) ** 2) @ want)) < 1e-3
    # by mean: count >= 1 suffices; a single-sample group has no variance entry
    film2 = _film({(0, 0): _pixel([_glossy((10.0, 10.0, 10.0))])}, 25, 25)
    assert dx.top_groups(film2, 5, 4)[0] == []
    (x, y, g), = dx.top_groups(film2, 5, 4, by_mean=True)[0]
    assert abs(float(dx.mean_luminance(g["stats"]["mean"])) - 10.0) < 1e-5


def test_metadata_json_and_pvar_layout():
    # test_diagnostic_export.cpp:119-150
    js = dx.metadata_json(1920, 1080, 2, 16, 4, 0)
    assert '"width": 1920' in js and '"height": 1080' in js and '"subsample_factor"' in js
    d = json.loads(js)
    assert (d["sampled_width"], d["sampled_height"], d["total_samples"]) == (960, 540, 0)
    rng = np.random.default_rng(3)
    pix = {}
    for y in range(3):
        for x in range(5):
            if (x + y) % 3:
                pix[(y, x)] = _pixel([_glossy(rng.random(3) * 4) for _ in range(int(rng.integers(1, 9)))])
    film = _film(pix, 3, 5)
    b = dx.export_binary(film, 10, 6, 2, 16, 4)
    assert b[:4] == b"PVAR" and len(b) == 28 + 22 * 15
    assert struct.unpack_from("<IIIIIBBBB", b, 4) == (1, 10, 6, 5, 3, 2, 16, 4, 0)
    for i in range(15):
        gc, ts, m0, m1, m2, v = struct.unpack_from("<HIffff", b, 28 + 22 * i)
        p = film.reshape(-1)[i]
        assert (gc, ts) == (int(p["num_groups"]), int(p["total_samples"]))
        assert (m0, m1, m2) == tuple(float(t) for t in p["total"]["mean"])
        tv = p["total"]["variance"]
        assert v == float(F(F(tv[0]) + F(tv[1])) + F(tv[2]))
        if ts:
            assert v > 0 or ts == 1


def test_top_groups_against_brute_force_sort():
    rng = np.random.default_rng(11)
    shapes = [[(int(rng.integers(0, 4)), 0, int(rng.integers(0, 4)), 0) for _ in range(int(rng.integers(1, 4)))]
              for _ in range(30)]
    pix = {}
    for y in range(6):
        for x in range(7):
            recs = []
            for _ in range(int(rng.integers(0, 40))):
                s = shapes[int(rng.integers(0, 30))]
                recs.append((s, tuple(float(v) for v in rng.random(3) * rng.choice([0.1, 1, 20])), 2, 0))
            pix[(y, x)] = _pixel(recs)
    film = _film(pix, 6, 7)
    for by_mean in (False, True):
        for n in (1, 7, 40, 10_000):
            top, last, cand = dx.top_groups(film, n, 3, by_mean)
            # brute force: every (pixel, slot) group with enough samples, numpy keys
            g = film["groups"]
            slots = np.arange(16)[None, None, :] < np.minimum(film["num_groups"], 16)[..., None]
            ok = slots & (g["stats"]["count"] >= (1 if by_mean else 2))
            if by_mean:
                m = g["stats"]["mean"].astype(np.float32)
                key = (F(0.2126) * m[..., 0] + F(0.7152) * m[..., 1]) + F(0.0722) * m[..., 2]
            else:
                v = g["stats"]["variance"].astype(np.float32)
                key = (v[..., 0] + v[..., 1]) + v[..., 2]
            ys, xs, ks = np.nonzero(ok)
            kv = key[ys, xs, ks]
            order = np.lexsort((ks, xs, ys, -kv.astype(np.float64)))[:n]
            assert len(top) == len(order)
            assert [float(dx.total_variance(t[2]["stats"]["variance"]) if not by_mean else
                          dx.mean_luminance(t[2]["stats"]["mean"])) for t in top] == [float(k) for k in kv[order]]
            assert [(t[0], t[1]) for t in top] == [(3 * int(xs[i]), 3 * int(ys[i])) for i in order]
            assert len(cand) == len(kv)


def test_coarse_merge_and_lookup():
    rng = np.random.default_rng(5)
    recs = {}
    for y in range(4):
        for x in range(4):
            recs[(y, x)] = [([(1, 0, 0, 0)], tuple(float(v) for v in rng.random(3) * (1 + x + y)), 2, 0)
                            for _ in range(int(rng.integers(1, 6)))]
    film = _film({k: _pixel(v) for k, v in recs.items()}, 4, 4)
    out = dx.coarse_stats(film, 2)
    assert [c["type_id"] for c in out] == list(range(8)) and len(out) == 8
    t = [int(ct) for ct in np.nonzero(film[0, 0]["coarse"]["count"])[0]][0]
    # subsample 2: pixel_data(2 sx, 2 sy) is pixel (sx, sy): every pixel once
    c = film["coarse"][..., t]
    n = c["count"].astype(np.float64)
    want = (c["mean"].astype(np.float64) * n[..., None]).sum(axis=(0, 1)) / n.sum()
    assert out[t]["sample_count"] == int(n.sum())
    np.testing.assert_allclose(np.array(out[t]["mean_rgb"], np.float64), want, rtol=1e-5)
    np.testing.assert_array_equal(np.array(out[t]["variance_rgb"]), c["variance"].reshape(-1, 3).max(axis=0))
    # subsample 1: only the pixels at even coordinates inside the film, each once
    o1 = dx.coarse_stats(film, 1)
    assert o1[t]["sample_count"] == int(c["count"][::2, ::2][:2, :2].sum())
    # subsample 4: pixel (sx // 2, sy // 2), each of those four times
    o4 = dx.coarse_stats(film, 4)
    assert o4[t]["sample_count"] == 4 * int(c["count"][:2, :2].sum())
    # one pixel alone: the merged mean is that pixel's mean exactly
    solo = _film({(0, 0): _pixel(recs[(1, 1)])}, 4, 4)
    s = dx.coarse_stats(solo, 2)
    assert s[t]["mean_rgb"] == [F(v) for v in solo[0, 0]["coarse"][t]["mean"]]


def test_suggestions_for_glossy_variance():
    # test_diagnostic_export.cpp:155-188: five glossy path types, contributions 0.1 / 100
    recs = []
    for p in range(5):
        for i in range(10):
            v = 0.1 if i % 2 == 0 else 100.0
            recs.append(([(p, p, 1, 0), (0, 0, 4, 0)], (v, v, v), 6, 0))
    film = _film({(0, 0): _pixel(recs)}, 25, 25)
    top, _, _ = dx.top_groups(film, 10, 4)
    assert len(top) >= 4
    sug = dx.suggestions(film, 4)
    assert sug and sug[0][:2] == ("Settings", "warning") and "glossy" in sug[0][2]
    assert dx.suggestions(_film({}, 2, 2), 1) == []


# ------------------------------------------------------------ the product --
lr = pytest.importorskip("lucidrenderer")


def _fill_host_film(seed, W, H, sub, paths):
    cfg = lr.PathRecordingConfig()
    cfg.subsample_factor = sub
    film = lr.DiagnosticFilm(W, H, cfg)
    rng = np.random.default_rng(seed)
    shapes = []
    for _ in range(40):
        verts = [(int(rng.integers(0, 5)), int(rng.integers(-1, 3)), int(rng.integers(0, 4)), int(rng.integers(0, 4)))
                 for _ in range(int(rng.integers(1, 6)))]
        shapes.append((verts, int(rng.integers(0, 3)), int(rng.integers(0, 3))))
    for _ in range(paths):
        verts, end, lid = shapes[int(rng.integers(0, 40))]
        r = lr.PathDiagnosticRecorder()
        r.begin_path()
        for o, m, t, f in verts:
            r.record_vertex(o, m, lr.BsdfType(t), bool(f & 1), bool(f & 2))
        if end == 0:
            r.record_environment_hit()
        elif end == 1:
            r.record_light_hit(lid, lr.LightSourceType.Area)
        else:
            r.record_emissive_hit(lid, 0)
        c = (rng.random(3) * rng.choice([0.05, 1.0, 30.0])).astype(np.float32)
        film.record_path(int(rng.integers(0, W)), int(rng.integers(0, H)), r.end_path(*map(float, c)))
    sw, sh = W // sub, H // sub
    arr = np.frombuffer(film.film_bytes(), np.uint8).view(oracle.PIXEL_DT).reshape(sh, sw)
    return film, arr


def _info_tuple(i):
    if isinstance(i, dict):
        return (i["pixel_x"], i["pixel_y"], i["signature"], i["sample_count"], tuple(map(float, i["mean_rgb"])),
                tuple(map(float, i["variance_rgb"])), i["depth"], i["coarse_type"], i["coarse_type_name"],
                float(i["mean_luminance"]), float(i["variance_luminance"]))
    return (i.pixel_x, i.pixel_y, i.signature, i.sample_count, tuple(map(float, i.mean_rgb)),
            tuple(map(float, i.variance_rgb)), i.depth, i.coarse_type, i.coarse_type_name, float(i.mean_luminance),
            float(i.variance_luminance))


def assert_top_groups_equal(got, film, n, sub, by_mean):
    """Product groups against the restatement, any order among equal keys
    (std::partial_sort leaves it unspecified)."""
    want, last, cand = dx.top_groups(film, n, sub, by_mean)
    key = (lambda t: t[9]) if by_mean else (lambda t: float(F(F(t[5][0]) + F(t[5][1])) + F(t[5][2])))
    g = [_info_tuple(i) for i in got]
    w = [_info_tuple(dx.group_info(*t)) for t in want]
    assert [key(t) for t in g] == [key(t) for t in w]
    if not w:
        return
    lk = key(w[-1])
    assert Counter(t for t in g if key(t) != lk) == Counter(t for t in w if key(t) != lk)
    pool = Counter(_info_tuple(dx.group_info(x, y, gg)) for _, x, y, gg in cand)
    pool = Counter({t: c for t, c in pool.items() if key(t) == lk})
    assert not (Counter(t for t in g if key(t) == lk) - pool)


@pytest.mark.parametrize("seed,W,H,sub,paths", [(1, 16, 12, 1, 3000), (2, 24, 16, 2, 5000), (3, 30, 30, 4, 4000),
                                                (4, 9, 7, 2, 800)])
def test_product_exports_equal_restatement(seed, W, H, sub, paths):
    film, arr = _fill_host_film(seed, W, H, sub, paths)
    ex = lr.DiagnosticExporter(film)
    assert ex.export_binary() == dx.export_binary(arr, W, H, sub, 16, 4)
    assert ex.export_metadata_json() == dx.metadata_json(W, H, sub, 16, 4, film.total_paths_recorded())
    st, want = ex.get_global_stats(), dx.global_stats(arr)
    assert (st.total_samples, st.active_pixels, st.total_groups, st.total_overflow) == tuple(want.values())
    for n in (1, 5, 25, 10_000):
        assert_top_groups_equal(ex.get_top_variance_groups(n), arr, n, sub, False)
        assert_top_groups_equal(ex.get_top_mean_groups(n), arr, n, sub, True)
    got_c = ex.get_coarse_stats()
    for g, w in zip(got_c, dx.coarse_stats(arr, sub)):
        assert (g.type_id, g.name, g.sample_count) == (w["type_id"], w["name"], w["sample_count"])
        assert [float(v) for v in g.mean_rgb] == [float(v) for v in w["mean_rgb"]]
        assert [float(v) for v in g.variance_rgb] == [float(v) for v in w["variance_rgb"]]
        assert float(g.mean_luminance) == float(w["mean_luminance"])
        assert float(g.variance_luminance) == float(w["variance_luminance"])
    got_s = [(s.category, s.severity, s.message, s.action) for s in ex.generate_suggestions()]
    assert got_s == dx.suggestions(arr, sub)
