"""GPU parity of the DiagnosticFilm path (traceMISWithDiagnostics and friends,
path_tracer.hpp:880-1514; DiagnosticFilm / PixelPathData<16,4>) against the
oracle, through the lucidrenderer module and the C ABI underneath.

Bars (same seed streams, closed box):
* the MIS diagnostic render equals the plain MIS render bit for bit on the
  GPU (test_pyrenderer_diagnostics.cpp:353-367), and the radiance of every
  diagnostic tracer matches the oracle to 2 % relative L1;
* per film pixel: record counts (total_samples) equal the oracle's in >= 99 %
  of pixels; where they are equal, the sets of group keys agree and the
  group / total statistics agree to 1e-3 relative (float Welford over
  contributions that differ from the oracle's in the last bits);
* group and total variances agree to 2e-3 relative (plus a small absolute
  floor where the variance is near zero);
* FP64 per-strategy Welford: n exact everywhere; means to 1e-4 relative; M2
  to 1e-3 relative in >= 99 % of pixel strategies (a path whose contribution
  differs in the last bits moves a near-zero M2 by more than 1e-3);
* independent seeds: per-strategy pixel means within 3 sigma of the oracle's
  (fraction and image-mean bounds of tests/sigma.py, with each strategy's
  own per-sample variance M2 / (n - 1)).
"""

import numpy as np
import pytest

import oracle
import workloads
import filmcheck
import sigma
from conftest import make_oracle, make_renderer

pytestmark = pytest.mark.gpu


def _render_diag(lucid, algo, W, H, spp, depth, sub=1, offset=0, scene=None):
    scene = workloads.closed_box() if scene is None else scene
    r = make_renderer(lucid, scene, workloads.CLOSED_BOX_CAMERA, algo)
    cfg = lucid.PathRecordingConfig()
    cfg.subsample_factor = sub
    r.enable_diagnostics(cfg)
    img = np.asarray(r.render_tile(0, 0, W, H, W, H, spp, offset, depth), np.float32)
    film = r.get_diagnostic_film_bytes()
    film = np.ascontiguousarray(film).view(oracle.PIXEL_DT)[..., 0]
    return r, img, film, r.get_diagnostic_strategy_stats()


# every algorithm twice, so a later render runs on storage an earlier one used;
# then the box with node-tree materials (checker floor, mixed checker wall)
@pytest.mark.parametrize("algo, textured", [("mis", False), ("nee", False), ("simple", False), ("mis", False),
                                            ("simple", False), ("mis", True), ("nee", True), ("simple", True)])
def test_film_matches_oracle(lucid, algo, textured):
    W = H = 16
    spp, depth = 8, 6
    scene = workloads.textured_box(panel=False) if textured else workloads.closed_box()
    r, img, film, strat = _render_diag(lucid, algo, W, H, spp, depth, scene=scene)
    o = make_oracle(scene, workloads.CLOSED_BOX_CAMERA)
    ref_img, ref_film, ref_strat, ref_tot = o.render_tile_diag(algo, 0, 0, W, H, W, H, samples=spp, max_depth=depth)
    rel = np.abs(img - ref_img).sum() / max(np.abs(ref_img).sum(), 1e-12)
    assert rel < 2e-2, f"radiance rel L1 {rel:.3g}"
    assert film.shape == ref_film.shape == (H, W)
    same = film["total_samples"] == ref_film["total_samples"]
    assert same.mean() >= 0.99, f"record counts agree in {same.mean():.3f} of pixels"
    st = r.get_diagnostic_stats()
    assert st.total_samples == int(film["total_samples"].sum()) and st.active_pixels > 0
    filmcheck.compare(film, ref_film, strat, ref_strat, same)
    np.testing.assert_array_equal(strat[..., 0], ref_strat[..., 0])  # strategy counts: every pixel


def test_mis_diagnostic_render_equals_plain(lucid):
    W = H = 16
    _, img, _, _ = _render_diag(lucid, "mis", W, H, 4, 6)
    r = make_renderer(lucid, workloads.closed_box(), workloads.CLOSED_BOX_CAMERA, "mis")
    plain = np.asarray(r.render_tile(0, 0, W, H, W, H, 4, 0, 6), np.float32)
    np.testing.assert_array_equal(img, plain)


def test_subsampled_film_and_api(lucid):
    W, H = 20, 14
    r, img, film, strat = _render_diag(lucid, "mis", W, H, 4, 6, sub=2)
    assert film.shape == (7, 10) and strat.shape == (7, 10, 4, 7)
    assert (strat[..., 0] == 4 * 4).all()
    vm = r.get_diagnostic_variance_map()
    assert vm.shape == (W * H,) and (vm >= 0).all()
    np.testing.assert_array_equal(vm.reshape(H, W)[::2, ::2], film["total"]["variance"].sum(-1))
    js = r.export_diagnostic_json()
    assert '"sampled_width": 10' in js and '"subsample_factor": 2' in js
    top = r.get_top_variance_groups(5)
    assert 0 < len(top) <= 5
    # ranked by total variance (the channel sum, diagnostic_film.hpp:47-49), reported as luminance variance
    tv = [sum(g.variance_rgb) for g in top]
    assert tv == sorted(tv, reverse=True)
    info = r.get_pixel_diagnostic(6, 4, "mean")
    assert info.valid and info.sample_count == int(film[2, 3]["total_samples"])
    means = [g.mean_luminance for g in info.top_groups]
    assert means == sorted(means, reverse=True)
    for g in info.top_groups:
        assert len(g.positions) == g.depth + 1 and len(g.normals) == g.depth
        assert g.strategy_name in ("MIS_BSDF", "MIS_NEE")
    r.clear_diagnostics()
    assert r.get_diagnostic_stats().total_samples == 0
    r.disable_diagnostics()
    assert r.export_diagnostic_json() == "{}"


@pytest.mark.parametrize("algo", ["simple", "nee", "mis", "volume_mis", "volume_simple"])
def test_records_with_every_algorithm(lucid, algo):
    # test_pyrenderer_diagnostics.cpp:463-477
    r, img, film, strat = _render_diag(lucid, algo, 4, 4, 2, 4)
    assert r.get_diagnostic_stats().total_samples > 0


def test_per_strategy_means_3sigma(lucid):
    W = H = 12
    n = 64
    _, _, _, g = _render_diag(lucid, "mis", W, H, n, 6, offset=7000)
    o = make_oracle(workloads.closed_box(), workloads.CLOSED_BOX_CAMERA)
    _, _, c, _ = o.render_tile_diag("mis", 0, 0, W, H, W, H, samples=n, max_depth=6)
    for k in (2, 3):  # MIS_BSDF, MIS_NEE: mean and per-sample variance (M2 / (n - 1)) of each strategy
        sigma.check(g[:, :, k, 1:4], g[:, :, k, 4:7] / (n - 1), n, c[:, :, k, 1:4], c[:, :, k, 4:7] / (n - 1), n,
                    f"strategy {k}")
