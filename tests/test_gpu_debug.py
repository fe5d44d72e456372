"""GPU debug passes (render_debug / render_debug_gpu, pybind_renderer.hpp:
371-520) against the oracle's restatement (tests/test_oracle_debug.py pins
it): normal, albedo and emission bit for bit on the closed box, the
textured box (node-tree albedo and emission at the hit's UV), the glass box
(smooth normals) and the smoke box; the volume thickness map of the smoke
box (heterogeneous grid) and of a homogeneous medium cube within 1e-5
relative (the device exp differs from the host's in the last bits); magenta
for an unknown mode; render_debug_gpu equals render_debug."""

import os

import numpy as np
import pytest

import workloads
from conftest import make_oracle, make_renderer

pytestmark = pytest.mark.gpu


def _scenes(tmp):
    cube = workloads.closed_box()
    cube["meshes"][5] = dict(cube["meshes"][5], material=dict(cube["meshes"][5]["material"],
                                                             volume={"density": 1.5, "color": [0.9, 0.6, 0.3]}))
    return {
        "closed_box": (workloads.closed_box(), workloads.CLOSED_BOX_CAMERA),
        "textured_box": (workloads.textured_box(), workloads.CLOSED_BOX_CAMERA),
        "glass_box": (workloads.glass_lights_box(), workloads.CLOSED_BOX_CAMERA),
        "smoke_box": (workloads.smoke_box(os.path.join(tmp, "d.bin"), grid_n=24), workloads.SMOKE_BOX_CAMERA),
        "medium_cube": (cube, workloads.CLOSED_BOX_CAMERA),
    }


@pytest.mark.parametrize("name", ["closed_box", "textured_box", "glass_box", "smoke_box", "medium_cube"])
def test_debug_modes_match_oracle(lucid, tmp_path, name):
    scene, cam = _scenes(str(tmp_path))[name]
    r = make_renderer(lucid, scene, cam, "mis")
    o = make_oracle(scene, cam)
    args = (16, 8, 96, 80, 128, 96)
    for mode in ("normal", "albedo", "emission", "volume", "depth"):
        g = np.asarray(r.render_debug(*args, mode), np.float32)
        c = o.render_debug(mode, *args)
        assert g.shape == c.shape == (96 * 80 * 4,)
        if mode == "volume":
            np.testing.assert_allclose(g, c, rtol=1e-5, atol=1e-7)
            if name in ("smoke_box", "medium_cube"):
                assert (c.reshape(-1, 4)[:, :3].max(axis=1) > 0).mean() > 0.02, "the medium should be in view"
        else:
            np.testing.assert_array_equal(g, c, err_msg=f"{name} {mode}")
    np.testing.assert_array_equal(np.asarray(r.render_debug_gpu(*args, "normal"), np.float32),
                                  np.asarray(r.render_debug(*args, "normal"), np.float32))
