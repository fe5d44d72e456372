"""GPU parity on the glass and light-shapes box (workloads.glass_lights_box):
the glass branch of sampleBSDF (bsdf.hpp:205-270: Fresnel choice by the
path's 4th draw, refraction, total internal reflection, the G2/G1 weight; the
face normal and the NEE skip at transmission > 0.5, path_tracer.hpp:455-567),
a partly transmissive surface (transmission 0.4), a sphere point light
(radius > 0), DISK and ELLIPSE area lights (sampling, intersection, pdf on
the BSDF-hit MIS path) and a spot light with a blended edge.

Bars, as for the other parity scenes:
* camera-ray closest hits bit-exact (mesh, triangle, u, v, t);
* per-sample radiance on the same seed streams: >= 99 % of samples within
  1e-3 relative, tile sums within 1 % relative L1;
* independent seeds, equal sample counts: pixel means within 3 sigma
  (fraction and image-mean bounds of tests/sigma.py).
"""

import numpy as np
import pytest

import sigma
import workloads
from conftest import make_oracle, make_renderer

pytestmark = pytest.mark.gpu

SCENE = workloads.glass_lights_box()
CAM = workloads.CLOSED_BOX_CAMERA


def test_glass_box_primary_hits_bit_exact(lucid):
    r = make_renderer(lucid, SCENE, CAM)
    o = make_oracle(SCENE, CAM)
    W = H = 128
    g = r.debug_primary_hits(0, 0, W, H, W, H, 2, 0)
    c = o.primary_hits(0, 0, W, H, W, H, samples=2)
    for k in ("mesh", "tri", "u", "v", "t"):
        same = np.asarray(g[k]).view(np.int32) == np.asarray(c[k]).view(np.int32)
        assert same.all(), f"{k}: {np.count_nonzero(~same)} of {same.size} rays differ"
    assert np.isin(np.asarray(g["mesh"]), [5, 6, 7]).mean() > 0.05, "the spheres should be in view"


@pytest.mark.parametrize("algorithm", ["mis", "nee", "simple"])
def test_glass_box_per_sample_radiance(lucid, algorithm):
    r = make_renderer(lucid, SCENE, CAM, algorithm)
    o = make_oracle(SCENE, CAM)
    W = H = 48
    spp = 4
    g = r.debug_render_samples(0, 0, W, H, W, H, spp, 3, 8)
    c = np.stack([o.trace_pixel(algorithm, x, y, W, H, spp, 3, spp, 8) for y in range(H) for x in range(W)])
    assert g.shape == c.shape
    close = (np.abs(g - c) <= 1e-3 * (1.0 + np.abs(c))).all(axis=2)
    assert close.mean() >= 0.99, f"only {close.mean():.4f} of samples agree"
    rel = np.abs(g.sum(1) - c.sum(1)).sum() / np.abs(c.sum(1)).sum()
    assert rel < 1e-2, f"relative L1 {rel:.3g}"


@pytest.mark.parametrize("algorithm", ["mis", "nee"])
def test_glass_box_statistics_3sigma(lucid, algorithm):
    r = make_renderer(lucid, SCENE, CAM, algorithm)
    o = make_oracle(SCENE, CAM)
    W = H = 48
    n = 256
    g = r.debug_render_samples(0, 0, W, H, W, H, n, 100000, 8).astype(np.float64)
    _, ex = o.render_tile(algorithm, 0, 0, W, H, W, H, samples=n, sample_offset=0, max_depth=8, stats=True)
    m_c, v_c = sigma.from_sums(ex["stats"][:, :3], ex["stats"][:, 3:], n)
    out = sigma.check(g.mean(axis=1), g.var(axis=1, ddof=1), n, m_c, v_c, n, f"glass box {algorithm}")
    assert out["n_pixels"] == W * H
