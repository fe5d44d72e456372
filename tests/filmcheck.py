"""Comparison of a GPU DiagnosticFilm (PixelPathData<16,4> per pixel,
diagnostics/pixel_path_data.hpp) and its FP64 per-strategy Welford state
with the oracle's, on pixels whose paths are the same on both sides.

On such a pixel the records are the same up to the rounding differences of
the contributions, so: the same group keys, counts, depths, strategies and
light types; group and total means to `rtol` relative (1e-3 by default, the
per-sample bar of the parity tests) and variances to 2 rtol (float Welford,
path_stats.hpp:63-98); per-strategy n exact, means to rtol / 10, M2 to
10 rtol (a near-zero M2 moves by more than that when one contribution
differs in its last bits, so M2 is checked on 99 % of strategy slots).
"""

import numpy as np


def live_groups(px):
    n = int(px["num_groups"])
    return {int(g["path_hash"]): g for g in px["groups"][:n]}


def compare(film, ref_film, strat, ref_strat, pixels, rtol=1e-3):
    """pixels: boolean (h, w) mask of the pixels to compare.  rtol: the
    relative tolerance of group / total means (variances and per-strategy
    M2 get 2x and 10x it, per-strategy means rtol / 10): the per-sample
    tolerance of the scene's parity bar.  Returns the number of compared
    pixels whose group key sets differ (allowed: 1 %)."""
    bad = 0
    for y, x in zip(*np.nonzero(pixels)):
        a, b = live_groups(film[y, x]), live_groups(ref_film[y, x])
        if a.keys() != b.keys():
            bad += 1
            continue
        for k in a:
            assert a[k]["stats"]["count"] == b[k]["stats"]["count"]
            np.testing.assert_allclose(a[k]["stats"]["mean"], b[k]["stats"]["mean"], rtol=rtol, atol=1e-6)
            sc = max(float(np.abs(b[k]["stats"]["mean"]).max()) ** 2, 1e-6)
            np.testing.assert_allclose(a[k]["stats"]["variance"], b[k]["stats"]["variance"], rtol=2 * rtol,
                                       atol=1e-4 * sc)
            assert a[k]["depth"] == b[k]["depth"] and a[k]["strategy"] == b[k]["strategy"]
            assert a[k]["light_type"] == b[k]["light_type"]
        tot, rtot = film[y, x]["total"], ref_film[y, x]["total"]
        np.testing.assert_allclose(tot["mean"], rtot["mean"], rtol=rtol, atol=1e-6)
        sc = max(float(np.abs(rtot["mean"]).max()) ** 2, 1e-6)
        np.testing.assert_allclose(tot["variance"], rtot["variance"], rtol=2 * rtol, atol=1e-4 * sc)
    assert bad <= 0.01 * pixels.sum() + 1, f"{bad} of {int(pixels.sum())} pixels have other group keys"
    s, r = strat[pixels], ref_strat[pixels]
    np.testing.assert_array_equal(s[..., 0], r[..., 0])
    np.testing.assert_allclose(s[..., 1:4], r[..., 1:4], rtol=rtol / 10, atol=1e-6)
    m2_scale = np.maximum(r[..., 1:4] ** 2 * np.maximum(r[..., :1], 1.0), 1e-9)
    close = np.abs(s[..., 4:7] - r[..., 4:7]) <= 10 * rtol * np.abs(r[..., 4:7]) + 1e-5 * m2_scale
    assert close.mean() >= 0.99, f"M2 agrees in {close.mean():.4f} of strategy channels"
    return bad
