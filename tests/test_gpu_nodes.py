"""GPU parity of node-tree materials: the device's node programs at the hit's
texture coordinate against the oracle's recursive evaluator
(src/node_evaluator.cpp) on the textured box -- a checker floor, a mixed
checker wall with an RGB node, a folded Principled tree, an unused tree and
an emissive checker panel whose light-list emission is its value at uv (0, 0).

Bars: camera-ray emission (depth 1) bit-exact; per-sample radiance as in the
untextured parity tests (99 % of samples within 1e-3 relative) and tile
sums within 1 % relative L1 on the box without the emissive panel; with the
panel, per-pixel means within 3 sigma; trees that fold to constants render
bit-identically to the equivalent constant materials.

Why the panel scene is compared statistically: a shadow ray aimed at a point
sampled on an emissive triangle ends on that triangle, and the reference's
shadow test (is_shadowed_skipping_volume_boundaries with tMax = the sampled
distance, path_tracer.hpp:103-119) counts the light's own triangle as an
occluder whenever its Moller-Trumbore t rounds to at most that distance.
Beyond the camera vertex, last-ulp differences of the device's sin / cos /
pow in the bounce directions flip that rounding for a few per cent of the
samples, so those samples differ by the whole NEE term, while their
distribution does not change.
"""

import numpy as np
import pytest

import lucid_capi
import workloads
import sigma
from conftest import make_oracle, make_renderer

pytestmark = pytest.mark.gpu

CAM = workloads.CLOSED_BOX_CAMERA


def test_emission_at_the_hit_uv_is_bit_exact(lucid):
    scene = workloads.textured_box()
    r = make_renderer(lucid, scene, CAM, "simple")
    o = make_oracle(scene, CAM)
    W, H, spp = 64, 64, 4
    g = r.debug_render_samples(0, 0, W, H, W, H, spp, 5, 1)
    c = np.stack([o.trace_pixel("simple", x, y, W, H, spp, 5, spp, 1) for y in range(H) for x in range(W)])
    assert g.shape == c.shape
    assert np.array_equal(g.view(np.int32), c.view(np.int32)), np.abs(g - c).max()
    on_panel = o.primary_hits(0, 0, W, H, W, H, samples=spp, sample_offset=5)["mesh"].reshape(W * H, spp) == 7
    lit = (c == np.float32([12.0, 10.0, 8.0])).all(axis=2)
    dark = (c == 0).all(axis=2)
    # both checker colours of the panel are seen
    assert (lit & on_panel).sum() > 20 and (dark & on_panel).sum() > 20


@pytest.mark.parametrize("algorithm", ["mis", "nee", "simple"])
def test_per_sample_radiance_textured_box(lucid, algorithm):
    scene = workloads.textured_box(panel=False)
    r = make_renderer(lucid, scene, CAM, algorithm)
    o = make_oracle(scene, CAM)
    W = H = 48
    spp = 4
    g = r.debug_render_samples(0, 0, W, H, W, H, spp, 3, 8)
    c = np.stack([o.trace_pixel(algorithm, x, y, W, H, spp, 3, spp, 8) for y in range(H) for x in range(W)])
    tol = 1e-3 * (1.0 + np.abs(c))
    close = (np.abs(g - c) <= tol).all(axis=2)
    assert close.mean() >= 0.99, f"only {close.mean():.4f} of samples agree"


@pytest.mark.parametrize("algorithm", ["mis", "nee"])
def test_textured_box_with_panel_statistics_3sigma(lucid, algorithm):
    """Independent seeds: GPU and oracle pixel means agree within 3 sigma."""
    scene = workloads.textured_box()
    r = make_renderer(lucid, scene, CAM, algorithm)
    o = make_oracle(scene, CAM)
    W = H = 40
    n = 64
    g = r.debug_render_samples(0, 0, W, H, W, H, n, 1000, 8).astype(np.float64)
    _, ex = o.render_tile(algorithm, 0, 0, W, H, W, H, samples=n, sample_offset=0, max_depth=8, stats=True)
    st = ex["stats"]
    m_c, v_c = sigma.from_sums(st[:, :3], st[:, 3:], n)
    sigma.check(g.mean(axis=1), g.var(axis=1, ddof=1), n, m_c, v_c, n, f"textured box {algorithm}")


def test_render_tile_textured_matches_oracle(lucid):
    scene = workloads.textured_box(panel=False)
    r = make_renderer(lucid, scene, CAM, "mis")
    o = make_oracle(scene, CAM)
    W = H = 96
    img = np.asarray(r.render_tile(0, 0, W, H, W, H, 16, 0, 8), np.float32)
    ref, _ = o.render_tile("mis", 0, 0, W, H, W, H, samples=16, max_depth=8)
    rel = np.abs(img - ref).sum() / np.abs(ref).sum()
    assert rel < 1e-2, f"relative L1 {rel:.3g}"


def _constant_trees(scene):
    """Every wall of the closed box as a node tree that folds to a constant
    (Principled, Diffuse, a Mix of an RGB node and a constant, an emissive
    Principled box), and the same scene with the folded values as legacy
    fields."""
    P, OUT, L = workloads._principled_node, workloads._output_node, workloads._link
    rng = np.random.default_rng(8)
    trees = []
    for i, m in enumerate(scene["meshes"]):
        c = [float(x) for x in rng.uniform(0.2, 0.8, 3)]
        k = i % 4
        if k == 0:
            nodes = [OUT(), P(base=c, metallic=0.2, roughness=0.35)]
            links = [L("Principled BSDF", "BSDF", "Material Output", "Surface")]
        elif k == 1:
            nodes = [OUT(), workloads._node("D", "ShaderNodeBsdfDiffuse", [{"name": "Color", "default_value": c}])]
            links = [L("D", "BSDF", "Material Output", "Surface")]
        elif k == 2:
            nodes = [OUT(), workloads._node("D", "ShaderNodeBsdfDiffuse", [{"name": "Color"}]),
                     workloads._node("Mix", "ShaderNodeMixRGB", [{"name": "Fac", "default_value": 0.3},
                                                                 {"name": "A"}, {"name": "B", "default_value": c}]),
                     workloads._node("RGB", "ShaderNodeRGB", [], [{"name": "Color",
                                                                   "default_value": [0.9, 0.5, 0.1, 1.0]}])]
            links = [L("D", "BSDF", "Material Output", "Surface"), L("Mix", "Color", "D", "Color"),
                     L("RGB", "Color", "Mix", "A")]
        else:
            nodes = [OUT(), P(base=c, emission=(1.0, 0.8, 0.6), strength=1.5, roughness=0.8)]
            links = [L("Principled BSDF", "BSDF", "Material Output", "Surface")]
        trees.append({"use_nodes": True, "node_tree": {"nodes": nodes, "links": links}})
    folded = {"meshes": [], "native_lights": scene["native_lights"], "environment": scene["environment"]}
    noded = dict(folded, meshes=[])
    for m, t in zip(scene["meshes"], trees):
        noded["meshes"].append(dict(m, material=t))
        v = lucid_capi.material_eval(t, [[0.0, 0.0]])[0].astype(float)
        legacy = {"base_color": list(v[:3]), "emission": list(v[3:6]), "metallic": v[6], "roughness": v[7],
                  "transmission": v[8], "ior": v[9]}
        folded["meshes"].append(dict(m, material=legacy))
    return noded, folded


def test_constant_trees_render_like_constant_materials(lucid):
    noded, folded = _constant_trees(workloads.closed_box())
    W, H = 80, 64
    out = []
    for sc in (noded, folded):
        r = make_renderer(lucid, sc, CAM, "mis")
        out.append(np.asarray(r.render_tile(0, 0, W, H, W, H, 8, 2, 8), np.float32))
    assert np.array_equal(out[0], out[1])
    assert out[0][0::4].sum() > 0


def test_node_scene_in_a_batched_session_pass():
    """Sessions of the textured box equal their own render_tile, bit for bit."""
    ctx = lucid_capi.Context(0)
    try:
        sh = ctx.scene(workloads.textured_box())
        cam = lucid_capi.camera(CAM)
        s = [lucid_capi.Session(ctx, sh, cam, 40, 32, 8, "mis"), lucid_capi.Session(ctx, sh, cam, 24, 24, 6, "nee")]
        lucid_capi.render_sessions(ctx, s, 3)
        lucid_capi.render_sessions(ctx, s, 5)
        for ss, (w, h, d, a) in zip(s, ((40, 32, 8, "mis"), (24, 24, 6, "nee"))):
            ref = np.zeros(w * h * 4, np.float32)
            pr = lucid_capi.params(0, 0, w, h, w, h, 8, 0, d, a)
            ctx.check(ctx.L.lucid_render_tile(ctx.h, sh, lucid_capi.C.byref(cam), lucid_capi.C.byref(pr),
                                              ref.ctypes.data_as(lucid_capi.C.POINTER(lucid_capi.C.c_float))))
            assert np.array_equal(ss.read(average=False), ref)
            ss.close()
        ctx.L.lucid_scene_release(sh)
    finally:
        ctx.close()
