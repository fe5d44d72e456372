"""Pins the oracle's participating-media path (volume/transmittance.hpp,
volume/phase.hpp, traceVolumeMIS path_tracer.hpp:586-835) by closed forms:

* HG phase function integrates to 1 over the sphere and has mean cosine g.
* A homogeneous, non-scattering (colour 0) gridded medium in front of a
  uniform environment: delta tracking survives with probability
  exp(-sigma * L_eff), L_eff = box length minus one voxel (the trilinear
  lookup is 0 in the half-voxel margins, transmittance.hpp:315-320), so the
  mean radiance equals env * exp(-sigma * L_eff) within Monte Carlo error.
* The same medium's ray-marched extinction transmittance (used for NEE
  shadow rays) equals exp(-sigma * L_eff) up to the midpoint rule.
"""

import math
import os

import numpy as np
import pytest

import oracle
import workloads


def test_hg_normalised_and_mean_cosine():
    L = oracle.lib()
    n = 20000
    mu = -1 + (np.arange(n) + 0.5) * 2 / n
    for g in (0.0, 0.3, -0.6, 0.8):
        p = np.array([L.orc_hg_eval(g, float(m)) for m in mu])
        assert abs(p.sum() * 2 / n * 2 * math.pi - 1) < 1e-3
        assert abs((p * mu).sum() * 2 / n * 2 * math.pi - g) < 2e-3


def _slab_scene(tmp_path, sigma, n=16, color=0.0):
    g = np.ones((n, n, n), np.float32)
    path = os.path.join(str(tmp_path), "grid.bin")
    g.tofile(path)
    lo, hi = [-1.0, -1.0, -1.0], [1.0, 1.0, 1.0]
    v, t = workloads._box([0, 0, 0], [1, 1, 1], 0.0)
    vol = {"color": [color] * 3, "absorption_color": [0, 0, 0], "density": sigma, "anisotropy": 0.0,
           "smoke_grid": {"dims": [n, n, n], "world_min": lo, "world_max": hi, "sidecar_path": path}}
    mesh = workloads._mesh(v, t, {"node_tree": workloads.boundary_node_tree(), "volume": vol})
    return {"meshes": [mesh], "environment": {"color": [1.0, 1.0, 1.0], "strength": 1.0}}


@pytest.mark.parametrize("sigma", [0.4, 1.5])
def test_homogeneous_absorber_survival(tmp_path, sigma):
    n = 16
    o = oracle.OracleScene(_slab_scene(tmp_path, sigma, n))
    o.set_camera((0, 0, 5), (0, 0, -1), (0, 1, 0), 10.0)
    k = 20000
    v = o.trace_pixel("volume_mis", 0, 0, 1, 1, 1, 0, k, 4)[:, 0].astype(np.float64)
    L_eff = 2.0 - 2.0 / n
    want = math.exp(-sigma * L_eff)
    se = math.sqrt(want * (1 - want) / k)
    assert abs(v.mean() - want) < 4 * se


def test_extinction_march_matches_beer_lambert(tmp_path):
    # traceVolumeSimple with colour 0 multiplies by exp(-sum sigma_a), sigma_a = sigma
    n, sigma = 16, 0.7
    o = oracle.OracleScene(_slab_scene(tmp_path, sigma, n))
    o.set_camera((0, 0, 5), (0, 0, -1), (0, 1, 0), 10.0)
    v = o.trace_pixel("volume_simple", 0, 0, 1, 1, 1, 0, 1, 2)[0, 0]
    want = math.exp(-sigma * (2.0 - 2.0 / n))
    assert abs(v - want) / want < 2e-3
