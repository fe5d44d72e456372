"""DiagnosticFilm exports of a film rendered on the GPU, against the oracle's
restatement of the reference's queries (oracle/diag_export.py):

* on the device film itself: PVAR bytes, metadata JSON, variance map and
  global statistics equal; top variance / mean groups equal up to the order
  of equal keys; coarse statistics and suggestions equal bit for bit;
* against the oracle's own film of the same render: identical PVAR headers;
  identical PVAR records for every pixel whose film bytes (padding aside)
  are identical on both sides; elsewhere, on the >= 99 % of pixels whose
  record counts agree, equal group counts and the mean and total variance
  within the film parity bars (1e-3 and 2e-3 relative: contributions differ
  from the oracle's in the last bits).
"""

import json

import numpy as np
import pytest

import diag_export as dx
import oracle
import workloads
from conftest import make_oracle, make_renderer
from test_oracle_diag_export import assert_top_groups_equal

pytestmark = pytest.mark.gpu


def _canonical(film):
    """Film pixels as bytes with the padding zeroed (each group's 4 pad
    bytes and the 4 tail bytes of each 20-byte outlier record)."""
    raw = np.ascontiguousarray(film).reshape(-1).view(np.uint8).reshape(-1, oracle.PIXEL_DT.itemsize).copy()
    g0, o0 = oracle.PIXEL_DT.fields["groups"][1], oracle.PIXEL_DT.fields["outliers"][1]
    pad = oracle.GROUP_DT.fields["pad"][1]
    for k in range(16):
        raw[:, g0 + k * oracle.GROUP_DT.itemsize + pad:g0 + k * oracle.GROUP_DT.itemsize + pad + 4] = 0
    for i in range(4):
        raw[:, o0 + 24 * i + 20:o0 + 24 * i + 24] = 0
    return raw


@pytest.mark.parametrize("algo,sub,W,H,spp", [("mis", 1, 24, 20, 8), ("nee", 2, 30, 22, 8), ("simple", 4, 40, 32, 4)])
def test_device_film_exports_equal_restatement(lucid, algo, sub, W, H, spp):
    scene = workloads.closed_box()
    cam = workloads.CLOSED_BOX_CAMERA
    r = make_renderer(lucid, scene, cam, algo)
    cfg = lucid.PathRecordingConfig()
    cfg.subsample_factor = sub
    r.enable_diagnostics(cfg)
    img = np.asarray(r.render_tile(0, 0, W, H, W, H, spp, 0, 6), np.float32)
    assert np.isfinite(img).all()
    film = np.ascontiguousarray(r.get_diagnostic_film_bytes()).view(oracle.PIXEL_DT)[..., 0]
    sh, sw = H // sub, W // sub
    assert film.shape == (sh, sw)
    st = r.get_diagnostic_stats()
    want = dx.global_stats(film)
    assert (st.total_samples, st.active_pixels, st.total_groups, st.total_overflow) == tuple(want.values())
    b = r.export_diagnostic_binary()
    assert b == dx.export_binary(film, W, H, sub, 16, 4)
    js = r.export_diagnostic_json()
    assert js == dx.metadata_json(W, H, sub, 16, 4, want["total_samples"])
    assert json.loads(js)["total_samples"] == want["total_samples"] > 0
    np.testing.assert_array_equal(np.asarray(r.get_diagnostic_variance_map()), dx.variance_map(film, W, H, sub))
    for n in (1, 10, 1000):
        assert_top_groups_equal(r.get_top_variance_groups(n), film, n, sub, False)
        assert_top_groups_equal(r.get_top_mean_groups(n), film, n, sub, True)
    for g, w in zip(r.get_diagnostic_coarse_stats(), dx.coarse_stats(film, sub)):
        assert (g.type_id, g.name, g.sample_count) == (w["type_id"], w["name"], w["sample_count"])
        assert [float(v) for v in g.mean_rgb] == [float(v) for v in w["mean_rgb"]]
        assert [float(v) for v in g.variance_rgb] == [float(v) for v in w["variance_rgb"]]
        assert float(g.variance_luminance) == float(w["variance_luminance"])
    got_s = [(s.category, s.severity, s.message, s.action) for s in r.generate_diagnostic_suggestions()]
    assert got_s == dx.suggestions(film, sub)

    # the oracle's film of the same render: identical pixels give identical PVAR records
    o = make_oracle(scene, cam)
    _, ref_film, _, _ = o.render_tile_diag(algo, 0, 0, W, H, W, H, samples=spp, max_depth=6, subsample=sub)
    assert ref_film.shape == film.shape
    same = (_canonical(film) == _canonical(ref_film)).all(axis=1)
    rec = np.frombuffer(b[28:], dx.PVAR_PIXEL_DT)
    ref_b = dx.export_binary(ref_film, W, H, sub, 16, 4)
    assert b[:28] == ref_b[:28]
    ref_rec = np.frombuffer(ref_b[28:], dx.PVAR_PIXEL_DT)
    assert rec[same].tobytes() == ref_rec[same].tobytes()
    # every pixel: counts exact where the record counts agree (>= 99 %), the
    # mean and the total variance to the film parity bars
    cnt = rec["total_samples"] == ref_rec["total_samples"]
    assert cnt.mean() >= 0.99
    np.testing.assert_array_equal(rec["group_count"][cnt], ref_rec["group_count"][cnt])
    np.testing.assert_allclose(rec["mean"][cnt], ref_rec["mean"][cnt], rtol=1e-3, atol=1e-6)
    sc = np.maximum(np.abs(ref_rec["mean"]).max(axis=1) ** 2, 1e-6)  # the film tests' variance floor, 3 channels
    dv = np.abs(rec["variance"] - ref_rec["variance"])
    assert (dv <= 2e-3 * np.abs(ref_rec["variance"]) + 3e-4 * sc)[cnt].all()
    print(f"{algo} sub {sub}: {same.mean():.3f} of film pixels byte-identical to the oracle's, "
          f"{cnt.mean():.3f} with equal record counts")
