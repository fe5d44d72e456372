"""The oracle's restatement of the DiagnosticFilm path (a13, a14), pinned by
the reference's own known-answer tests and by invariants of the tracers.

* PathStatistics (test_path_stats.cpp): mean/variance of 1,2,3 = 2, 1; total
  variance 6; Welford stability at a 1e6 offset; single sample variance 0.
* PathTrace::hash (test_path_types.cpp:142-180): equal paths hash equal, any
  field or the depth changes it; cross-checked against FNV-1a 64 computed
  here with Python integers (object and material ids widened with sign).
* PixelPathData (test_pixel_path_data.cpp): matching paths merge (mean 2),
  different paths make groups, eviction when full (<4,2>), high-variance
  group kept (<2,2>), coarse stats, 3-sigma outliers after 10 samples.
* Diagnostic tracers (path_tracer.hpp:880-1514): the MIS one renders the same
  image as traceMIS without volume boundaries (test_pyrenderer_diagnostics
  .cpp:353-367); every radiance term is recorded, so per pixel the FP64
  per-strategy means times n add up to the radiance sum; record counts add
  up to the film's sample counts; NEE records light hits at depth 0 only.
"""

import numpy as np
import pytest

import oracle
import workloads

DIFF, GLOSSY, MIRROR, GLASS, EMISSION, ENV = range(6)


def _rec(verts, c, lt=0, st=0):
    return (verts, c, lt, st)


def test_welford_known_answers():
    o = oracle.diag_pixel_run([_rec([(0, 0, DIFF, 0)], (v, v, v)) for v in (1.0, 2.0, 3.0)])
    np.testing.assert_allclose(o["gmean"][0], [2, 2, 2], rtol=1e-6)
    np.testing.assert_allclose(o["gvar"][0], [1, 1, 1], rtol=1e-6)
    o = oracle.diag_pixel_run([_rec([(0, 0, DIFF, 0)], (1, 2, 3)), _rec([(0, 0, DIFF, 0)], (3, 4, 5))])
    assert float(o["gvar"][0].sum()) == pytest.approx(6.0, rel=1e-6)
    off = 1e6
    o = oracle.diag_pixel_run([_rec([(0, 0, DIFF, 0)], (off + v, 0, 0)) for v in (1.0, 2.0, 3.0)])
    assert abs(o["gmean"][0][0] - (off + 2)) < 0.01 and abs(o["gvar"][0][0] - 1.0) < 0.01
    o = oracle.diag_pixel_run([_rec([(0, 0, DIFF, 0)], (1, 2, 3))])
    assert o["gcount"][0] == 1 and (o["gvar"][0] == 0).all()


def _fnv(verts):
    m = (1 << 64) - 1
    h = 14695981039346656037
    h = ((h ^ len(verts)) * 1099511628211) & m
    for o_, mt, t, f in verts:
        for x in (o_ & m, mt & m, t, f):
            h = ((h ^ x) * 1099511628211) & m
    return h


def test_path_hash():
    paths = [[(0, 1, EMISSION, 0), (1, 2, DIFF, 0)], [(0, 1, EMISSION, 0), (1, 2, GLOSSY, 0)],
             [(0, 1, EMISSION, 0)], [(0, 1, EMISSION, 0), (1, 2, DIFF, 1)], [(-1, -1, EMISSION, 0)],
             [(-5, -1, EMISSION, 0), (3, 0, GLASS, 1), (7, 0, MIRROR, 2)], [(0, 1, EMISSION, 0), (1, 2, DIFF, 0)]]
    o = oracle.diag_pixel_run([_rec(p, (1, 1, 1)) for p in paths])
    assert [int(h) for h in o["hash"]] == [_fnv(p) for p in paths]
    h = [int(x) for x in o["hash"]]
    assert h[0] == h[6] and len(set(h[:6])) == 6
    # coarse type: bit0 light-sampled flag, bit1 delta flag, bit2 depth > 2
    assert list(o["coarse"]) == [0, 0, 0, 2, 0, 7 & ~1 | 1, 0]


def test_pixel_path_data_known_answers():
    base = [(0, 0, EMISSION, 0), (1, 1, DIFF, 0)]
    o = oracle.diag_pixel_run([_rec(base, (1, 1, 1)), _rec(base, (3, 3, 3))])
    assert o["summary"][0] == 1 and o["summary"][2] == 2 and o["gmean"][0][0] == 2.0
    o = oracle.diag_pixel_run([_rec(base, (1, 1, 1)), _rec([(0, 0, EMISSION, 0), (2, 2, GLOSSY, 0)], (.5, .5, .5))])
    assert o["summary"][0] == 2
    # eviction when full, PixelPathData<4, 2>
    recs = [_rec([(i, i, DIFF, 0)], (1, 1, 1)) for i in range(4)] + [_rec([(99, 99, GLOSSY, 0)], (10, 10, 10))]
    o = oracle.diag_pixel_run(recs, 4, 2)
    assert o["summary"][0] == 4 and o["summary"][3] > 0
    # the high-variance group survives eviction, PixelPathData<2, 2>
    recs = [_rec([(0, 0, DIFF, 0)], (1, 1, 1))] * 10
    recs += [_rec([(1, 1, GLOSSY, 0)], (.1, .1, .1)), _rec([(1, 1, GLOSSY, 0)], (10, 10, 10))]
    recs += [_rec([(2, 2, MIRROR, 0)], (5, 5, 5))]
    o = oracle.diag_pixel_run(recs, 2, 2)
    assert GLOSSY in list(o["gtype0"][: o["summary"][0]])
    # coarse stats
    o = oracle.diag_pixel_run([_rec(base, (1, 2, 3))])
    assert o["ccount"][0] == 1 and o["cmean"][0][0] == 1.0
    # outliers: 20 equal samples then one far above
    o = oracle.diag_pixel_run([_rec(base, (1, 1, 1))] * 20 + [_rec(base, (100, 100, 100))])
    assert o["summary"][1] > 0
    o = oracle.diag_pixel_run([_rec(base, (1, 1, 1))] * 9 + [_rec(base, (100, 100, 100))])
    assert o["summary"][1] == 0  # fewer than 10 samples: no baseline yet


@pytest.fixture(scope="module")
def box():
    sc = workloads.closed_box()
    cam = workloads.CLOSED_BOX_CAMERA
    o = oracle.OracleScene(sc)
    o.set_camera(cam["pos"], cam["dir"], cam["up"], cam["fov"])
    return o


def test_mis_diagnostic_render_matches_plain(box):
    W = H = 16
    a, _ = box.render_tile("mis", 0, 0, W, H, W, H, samples=4, max_depth=6)
    b, film, strat, tot = box.render_tile_diag("mis", 0, 0, W, H, W, H, samples=4, max_depth=6)
    np.testing.assert_array_equal(a, b)
    assert tot == int(film["total_samples"].sum()) and tot > 0


@pytest.mark.parametrize("algo", ["simple", "nee", "mis"])
def test_records_account_for_radiance(box, algo):
    W = H = 12
    n = 6
    rgba, film, strat, tot = box.render_tile_diag(algo, 0, 0, W, H, W, H, samples=n, max_depth=6)
    img = oracle.flip_to_top_first(rgba, W, H)[..., :3].astype(np.float64)
    assert (strat[..., 0] == n).all()
    recon = (strat[..., 1:4] * n).sum(axis=2)
    np.testing.assert_allclose(recon, img, rtol=1e-4, atol=1e-5)
    used = {"simple": [0], "nee": [0, 1], "mis": [2, 3]}[algo]
    for k in range(4):
        if k not in used:
            assert (strat[:, :, k, 1:] == 0).all()
    # group sample counts add up to the records kept
    g = film["groups"]
    assert (film["num_groups"] <= 16).all()
    live = np.arange(16)[None, None, :] < film["num_groups"][..., None]
    assert (g["stats"]["count"] * live).sum() <= film["total_samples"].sum()


def test_nee_records_light_hits_only_at_depth0(box):
    W = H = 12
    _, film, _, _ = box.render_tile_diag("nee", 0, 0, W, H, W, H, samples=8, max_depth=6)
    g = film["groups"]
    live = np.arange(16)[None, None, :] < film["num_groups"][..., None]
    # BSDF-strategy groups end at the first vertex (a camera ray hitting a
    # light or emitter), everything later is NEE
    bsdf = live & (g["strategy"] == 0)
    assert (g["depth"][bsdf] == 1).all()
    assert (g["strategy"][live] <= 1).all()


def test_subsampled_film_shape(box):
    W, H = 10, 7
    rgba, film, strat, tot = box.render_tile_diag("mis", 0, 0, W, H, W, H, samples=2, max_depth=4, subsample=2)
    assert film.shape == (3, 5) and strat.shape == (3, 5, 4, 7)
    assert (strat[..., 0] == 2 * 4).all()
