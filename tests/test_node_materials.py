"""Node-tree materials (src/node_evaluator.cpp) on the host.

The oracle's recursive evaluator is pinned by closed forms and by the
reference's special cases (defaults, fallbacks, the Emission-albedo and
strength-read quirks): a checkerboard parity computed in numpy, the lerp of
a Mix in float32, colour x strength, and the hit's texture coordinate
against the affine UV map of a planar quad.  The product's compiled node
programs (lucid_material_eval: the host interpreter of the same programs
the device runs) must then equal the oracle bit for bit on random trees,
and the pybind module's JSON parse must hand the same trees over.
"""

import numpy as np
import pytest

import lucid_capi
import oracle
import workloads

OUT = workloads._output_node
LINK = workloads._link


def _mat(nodes, links, use_nodes=True, **legacy):
    return dict(legacy, use_nodes=use_nodes, node_tree={"nodes": nodes, "links": links})


def _oracle_eval(material, uv, **mesh_extra):
    mesh = dict({"vertices": [[0, 0, 0], [1, 0, 0], [0, 1, 0]], "triangles": [[0, 1, 2]], "material": material},
                **mesh_extra)
    return oracle.OracleScene({"meshes": [mesh]}).material_eval(0, uv)


def _same(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.int32), np.asarray(b, np.float32).view(np.int32))


def _principled_over(shader_links, extra_nodes):
    return [OUT(), workloads._principled_node()] + extra_nodes, [
        LINK("Principled BSDF", "BSDF", "Material Output", "Surface")] + shader_links


# ------------------------------------------------------------ oracle pins --

def test_oracle_checker_matches_closed_form_parity():
    c1, c2, scale = (0.9, 0.1, 0.2), (0.05, 0.6, 0.3), 3.0
    nodes, links = _principled_over([LINK("Checker", "Color", "Principled BSDF", "Base Color")],
                                    [workloads._checker_node("Checker", c1, c2, scale)])
    g = np.linspace(-2.0, 2.0, 49, dtype=np.float32)  # includes exact cell boundaries (k / 3)
    uv = np.stack(np.meshgrid(g, g), -1).reshape(-1, 2).astype(np.float32)
    uv = np.concatenate([uv, np.float32([[1.0 / 3.0, -1.0 / 3.0], [-1e-7, 1e-7], [0.0, 0.0]])])
    got = _oracle_eval(_mat(nodes, links), uv)
    # floor(u * scale) mod 2 (Python's mod is the reference's fixed-up parity)
    cu = np.floor(uv[:, 0] * np.float32(scale)).astype(np.int64) % 2
    cv = np.floor(uv[:, 1] * np.float32(scale)).astype(np.int64) % 2
    want = np.where(((cu + cv) % 2 == 0)[:, None], np.float32(c1), np.float32(c2))
    assert _same(got[:, :3], want)
    assert set(np.unique(cu)) == {0, 1} and (uv < 0).any()


def test_oracle_mix_is_float32_lerp_and_its_defaults():
    a, b = np.float32([0.7, 0.2, 0.1]), np.float32([0.1, 0.3, 0.9])
    for fac in (0.0, 0.35, 0.5, 1.0):
        mix = workloads._node("Mix", "ShaderNodeMixRGB", [{"name": "Fac", "default_value": fac},
                                                          {"name": "A", "default_value": workloads._rgba(a)},
                                                          {"name": "B", "default_value": workloads._rgba(b)}])
        nodes, links = _principled_over([LINK("Mix", "Color", "Principled BSDF", "Base Color")], [mix])
        got = _oracle_eval(_mat(nodes, links), [[0.3, 0.4]])[0, :3]
        f = np.float32(fac)
        assert _same(got, a * (np.float32(1) - f) + b * f)
    # new-style ShaderNodeMix names its factor "Factor": the default 0.5 applies,
    # and unlinked A / B without colours are black / white
    mix = workloads._node("Mix", "ShaderNodeMix", [{"name": "Factor", "default_value": 0.9},
                                                   {"name": "A", "default_value": 1.0}, {"name": "B"}])
    nodes, links = _principled_over([LINK("Mix", "Result", "Principled BSDF", "Base Color")], [mix])
    assert _same(_oracle_eval(_mat(nodes, links), [[0, 0]])[0, :3], np.float32([0.5, 0.5, 0.5]))


def test_oracle_emission_forms():
    em = workloads._node("Emission", "ShaderNodeEmission", [
        {"name": "Color", "default_value": [0.5, 0.25, 1.0, 1.0]}, {"name": "Strength", "default_value": 3.0}])
    got = _oracle_eval(_mat([OUT(), em], [LINK("Emission", "Emission", "Material Output", "Surface")]), [[0, 0]])[0]
    want = np.float32([0.5, 0.25, 1.0]) * np.float32(3.0)
    assert _same(got[3:6], want)
    assert _same(got[:3], want)  # the albedo of an Emission surface is its emission (evaluateNode on "Emission")
    # Principled emission: unlinked colour x strength; a missing strength is 0
    p = workloads._principled_node(emission=(1.0, 0.5, 0.25), strength=4.0)
    got = _oracle_eval(_mat([OUT(), p], [LINK("Principled BSDF", "BSDF", "Material Output", "Surface")]), [[0, 0]])
    assert _same(got[0, 3:6], np.float32([4.0, 2.0, 1.0]))
    p_nostr = dict(p, inputs=[x for x in p["inputs"] if x["name"] != "Emission Strength"])
    got = _oracle_eval(_mat([OUT(), p_nostr], [LINK("Principled BSDF", "BSDF", "Material Output", "Surface")]),
                       [[0, 0]])
    assert _same(got[0, 3:6], np.zeros(3))
    # linked emission colour: the strength socket's value is read even when linked
    rgb = workloads._node("RGB", "ShaderNodeRGB", [], [{"name": "Color", "default_value": [0.2, 0.4, 0.6, 1]}])
    nodes = [OUT(), p, rgb]
    links = [LINK("Principled BSDF", "BSDF", "Material Output", "Surface"),
             LINK("RGB", "Color", "Principled BSDF", "Emission Color"),
             LINK("RGB", "Color", "Principled BSDF", "Emission Strength")]
    got = _oracle_eval(_mat(nodes, links), [[0, 0]])
    assert _same(got[0, 3:6], np.float32([0.2, 0.4, 0.6]) * np.float32(4.0))
    # older trees call the colour socket "Emission"
    p_old = dict(p, inputs=[dict(x, name="Emission") if x["name"] == "Emission Color" else x for x in p["inputs"]])
    got = _oracle_eval(_mat([OUT(), p_old], [LINK("Principled BSDF", "BSDF", "Material Output", "Surface")]),
                       [[0, 0]])
    assert _same(got[0, 3:6], np.float32([4.0, 2.0, 1.0]))


def test_oracle_defaults_and_fallbacks():
    gray, black = np.float32([0.8] * 3), np.zeros(3, np.float32)
    scal_default = np.float32([0.0, 0.5, 0.0, 1.45])
    p = workloads._principled_node(base=(0.1, 0.2, 0.3), metallic=0.7, roughness=0.004)
    # no Material Output node
    got = _oracle_eval(_mat([p], []), [[0, 0]])[0]
    assert _same(got[:3], gray) and _same(got[3:6], black) and _same(got[6:], scal_default)
    # Surface not linked
    got = _oracle_eval(_mat([OUT(), p], []), [[0, 0]])[0]
    assert _same(got[:3], gray) and _same(got[6:], scal_default)
    # a linked Principled: its scalars, roughness clamped to 0.01
    got = _oracle_eval(_mat([OUT(), p], [LINK("Principled BSDF", "BSDF", "Material Output", "Surface")]),
                       [[0, 0]])[0]
    assert _same(got[:3], [0.1, 0.2, 0.3]) and _same(got[6:], [0.7, 0.01, 0.0, 1.45])
    # "Transmission" missing: "Transmission Weight"; a linked scalar input reads the default
    p2 = dict(p, inputs=[dict(x, default_value=0.6) if x["name"] == "Transmission Weight" else x
                         for x in p["inputs"]])
    rgb = workloads._node("RGB", "ShaderNodeRGB", [], [{"name": "Color"}])
    links = [LINK("Principled BSDF", "BSDF", "Material Output", "Surface"),
             LINK("RGB", "Color", "Principled BSDF", "Roughness")]
    got = _oracle_eval(_mat([OUT(), p2, rgb], links), [[0, 0]])[0]
    assert _same(got[6:], [0.7, 0.5, 0.6, 1.45])
    # an RGB node without an output value is white; with one, that colour
    links = [LINK("Principled BSDF", "BSDF", "Material Output", "Surface"),
             LINK("RGB", "Color", "Principled BSDF", "Base Color")]
    assert _same(_oracle_eval(_mat([OUT(), p, rgb], links), [[0, 0]])[0, :3], [1, 1, 1])
    rgb2 = dict(rgb, outputs=[{"name": "Color", "default_value": [0.3, 0.2, 0.1, 1.0]}])
    assert _same(_oracle_eval(_mat([OUT(), p, rgb2], links), [[0, 0]])[0, :3], [0.3, 0.2, 0.1])
    # unknown node type, link to a missing node, Principled asked for another socket
    tex = workloads._node("Image", "ShaderNodeTexImage", [])
    for nodes, links in (([OUT(), p, tex], [LINK("Principled BSDF", "BSDF", "Material Output", "Surface"),
                                            LINK("Image", "Color", "Principled BSDF", "Base Color")]),
                         ([OUT(), p], [LINK("Principled BSDF", "BSDF", "Material Output", "Surface"),
                                       LINK("Nowhere", "Color", "Principled BSDF", "Base Color")]),
                         ([OUT(), p], [LINK("Principled BSDF", "Alpha", "Material Output", "Surface")])):
        assert _same(_oracle_eval(_mat(nodes, links), [[0, 0]])[0, :3], gray)
    # Diffuse BSDF colour; a VEC3 default is read like a VEC4's RGB
    dif = workloads._node("Diffuse BSDF", "ShaderNodeBsdfDiffuse", [{"name": "Color", "default_value": [0.4, 0.5, 0.6]}])
    got = _oracle_eval(_mat([OUT(), dif], [LINK("Diffuse BSDF", "BSDF", "Material Output", "Surface")]), [[0, 0]])[0]
    assert _same(got[:3], [0.4, 0.5, 0.6]) and _same(got[6:], scal_default)
    # use_nodes false, or a tree without nodes: the legacy fields
    legacy = {"base_color": [0.2, 0.3, 0.4], "metallic": 0.1, "roughness": 0.2, "emission": [1, 2, 3]}
    for m in (_mat([OUT(), p], [LINK("Principled BSDF", "BSDF", "Material Output", "Surface")], use_nodes=False,
                   **legacy), _mat([], [], **legacy)):
        got = _oracle_eval(m, [[0.5, 0.5]])[0]
        assert _same(got[:6], [0.2, 0.3, 0.4, 1, 2, 3]) and _same(got[6:], [0.1, 0.2, 0.0, 1.45])


def test_oracle_hit_uv_follows_the_quad_parametrisation():
    """Hit::uv = uv0 w + uv1 u + uv2 v: rays straight down onto a floor quad
    whose corners carry an affine UV map read back that map at the hit point."""
    v, t = workloads._quad([-1, 0, -1], [-1, 0, 1], [1, 0, 1], [1, 0, -1])
    uvs = [[0.5, -1.0], [0.5, 3.0], [2.5, 3.0], [2.5, -1.0]]  # u = 1.5 + x, v = 1 + 2 z
    mesh = workloads._mesh(v, t, {"base_color": [0.5, 0.5, 0.5]})
    mesh["triangle_uvs"] = workloads._quad_uvs(uvs)
    o = oracle.OracleScene({"meshes": [mesh]})
    rng = np.random.default_rng(3)
    xz = rng.uniform(-0.99, 0.99, (500, 2))
    org = np.stack([xz[:, 0], np.full(500, 1.0), xz[:, 1]], -1)
    d = np.tile([0.0, -1.0, 0.0], (500, 1))
    got = o.intersect_uv(org, d)
    want = np.stack([1.5 + xz[:, 0], 1.0 + 2.0 * xz[:, 1]], -1)
    np.testing.assert_allclose(got, want, atol=2e-6)


# ------------------------------------------------- product against oracle --

_TYPES = ["ShaderNodeBsdfPrincipled", "ShaderNodeBsdfDiffuse", "ShaderNodeEmission", "ShaderNodeRGB",
          "ShaderNodeMix", "ShaderNodeMixRGB", "ShaderNodeTexChecker", "ShaderNodeTexImage"]
_SOCKETS = {"ShaderNodeBsdfPrincipled": ["Base Color", "Metallic", "Roughness", "IOR", "Transmission",
                                         "Transmission Weight", "Emission Color", "Emission", "Emission Strength"],
            "ShaderNodeBsdfDiffuse": ["Color"], "ShaderNodeEmission": ["Color", "Strength"],
            "ShaderNodeRGB": [], "ShaderNodeMix": ["Factor", "A", "B"], "ShaderNodeMixRGB": ["Fac", "A", "B"],
            "ShaderNodeTexChecker": ["Color1", "Color2", "Scale"], "ShaderNodeTexImage": ["Vector"]}
_OUT_SOCKETS = ["BSDF", "Emission", "Color", "Result", "Base Color", "Fac", "Alpha"]


def _random_value(rng):
    k = rng.integers(0, 7)
    if k == 0:
        return None
    if k == 1:
        return float(np.float32(rng.uniform(-0.5, 3.0)))
    if k == 2:
        return [float(x) for x in rng.uniform(0, 2, 3)]
    if k == 3:
        return [float(x) for x in rng.uniform(0, 2, 3)] + [1.0]
    if k == 4:
        return bool(rng.integers(0, 2))
    if k == 5:
        return "text"
    return [1.0, 2.0]


def _random_tree(rng, max_depth=3):
    """A random node graph in the exporter's schema: up to `max_depth`
    levels of links below the shader, every node type the evaluator knows
    plus an unknown one, random socket values of every JSON kind, and links
    to missing nodes."""
    nodes, links = [OUT()], []
    counter = [0]

    def make(depth):
        if depth == 0:
            typ = _TYPES[rng.integers(0, 3)]
        else:  # checkers (the uv-dependent node) a third of the time
            typ = "ShaderNodeTexChecker" if rng.random() < 0.35 else _TYPES[rng.integers(0, len(_TYPES))]
        counter[0] += 1
        name = f"N{counter[0]}"
        inputs = []
        for sn in _SOCKETS[typ]:
            if rng.random() < 0.15:
                continue  # socket absent
            inputs.append({"name": sn, "default_value": _random_value(rng)})
            if depth < max_depth and sn not in ("Scale", "Fac", "Factor", "Strength") and rng.random() < 0.6:
                child, sock = make(depth + 1)
                links.append(LINK(child, sock, name, sn))
        if rng.random() < 0.2 and depth < max_depth:  # a scalar input linked
            for x in inputs:
                if x["name"] in ("Scale", "Fac", "Roughness", "Strength", "Emission Strength"):
                    links.append(LINK("N1", "Color", name, x["name"]))
                    break
        outputs = [{"name": "Color", "default_value": _random_value(rng)}] if typ == "ShaderNodeRGB" else []
        nodes.append(workloads._node(name, typ, inputs, outputs))
        sock = _OUT_SOCKETS[rng.integers(0, len(_OUT_SOCKETS))]
        if typ == "ShaderNodeBsdfPrincipled" and rng.random() < 0.7:
            sock = "BSDF"
        if typ == "ShaderNodeEmission" and rng.random() < 0.7:
            sock = "Emission"
        return name, sock

    shader, sock = make(0)
    if rng.random() < 0.95:
        links.append(LINK(shader, sock, "Material Output", "Surface"))
    if rng.random() < 0.05:
        links.append(LINK("Missing", "Color", shader, _SOCKETS[nodes[-1]["type"]][0] if _SOCKETS[nodes[-1]["type"]]
                          else "X"))
    rng.shuffle(nodes)
    return _mat(nodes, links, use_nodes=bool(rng.random() < 0.95), base_color=[0.3, 0.2, 0.1], emission=[0, 0, 0])


def test_product_programs_equal_oracle_on_random_trees():
    rng = np.random.default_rng(17)
    uv = np.concatenate([rng.uniform(-3, 3, (96, 2)), np.float32([[0, 0], [1e9, -1e9], [0.5, 0.5]])]).astype(
        np.float32)
    textured = 0
    for _ in range(400):
        mat = _random_tree(rng)
        want = _oracle_eval(mat, uv)
        got = lucid_capi.material_eval(mat, uv)
        assert _same(got, want), mat
        textured += int(np.unique(want[:, :6], axis=0).shape[0] > 1)
    assert textured > 40  # a good share of the trees vary with uv


def test_sethi_ullman_order_and_slot_limit():
    """A texture expression deep on its B side is evaluated B first; one
    needing more than four evaluation slots is refused (the oracle, which
    recurses like the reference, still evaluates it)."""
    def checker_tree(name, depth, nodes, links):
        nodes.append(workloads._checker_node(name, (0.9, 0.1, 0.1), (0.1, 0.1, 0.9), 1.0 + depth))
        if depth > 0:
            for side, sock in (("a", "Color1"), ("b", "Color2")):
                checker_tree(name + side, depth - 1, nodes, links)
                links.append(LINK(name + side, "Color", name, sock))

    uv = np.random.default_rng(5).uniform(-2, 2, (64, 2)).astype(np.float32)
    for depth, ok in ((2, True), (3, False)):  # a complete tree of checkers needs depth + 2 slots
        nodes, links = _principled_over([LINK("c", "Color", "Principled BSDF", "Base Color")], [])
        checker_tree("c", depth, nodes, links)
        mat = _mat(nodes, links)
        want = _oracle_eval(mat, uv)
        if ok:
            assert _same(lucid_capi.material_eval(mat, uv), want)
        else:
            with pytest.raises(ValueError, match="evaluation slots"):
                lucid_capi.material_eval(mat, uv)
    # Mix(const, 4-slot checker tree) fits only with the B operand first
    nodes, links = _principled_over([LINK("Mix", "Color", "Principled BSDF", "Base Color")], [
        workloads._node("Mix", "ShaderNodeMixRGB", [{"name": "Fac", "default_value": 0.25},
                                                    {"name": "A", "default_value": [0.2, 0.3, 0.4, 1.0]},
                                                    {"name": "B"}])])
    checker_tree("c", 2, nodes, links)
    links.append(LINK("c", "Color", "Mix", "B"))
    mat = _mat(nodes, links)
    assert _same(lucid_capi.material_eval(mat, uv), _oracle_eval(mat, uv))


def test_link_cycle_is_refused():
    mix = workloads._node("Mix", "ShaderNodeMixRGB", [{"name": "Fac", "default_value": 0.5}, {"name": "A"},
                                                      {"name": "B"}])
    nodes, links = _principled_over([LINK("Mix", "Color", "Principled BSDF", "Base Color"),
                                     LINK("Mix", "Color", "Mix", "A")], [mix])
    with pytest.raises(ValueError, match="cycle"):
        lucid_capi.material_eval(_mat(nodes, links), [[0, 0]])


def test_pybind_scene_json_hands_over_the_same_trees(lucid):
    scene = workloads.textured_box()
    r = lucid.Renderer()
    assert r.load_scene_json(workloads.to_json(scene))
    o = oracle.OracleScene(scene)
    uv = np.random.default_rng(9).uniform(-2, 2, (128, 2)).astype(np.float32)
    varying = 0
    for i in range(r.get_mesh_count()):
        got = r.debug_material_eval(i, uv)
        want = o.material_eval(i, uv)
        assert _same(got, want), i
        varying += int(np.unique(want[:, :6], axis=0).shape[0] > 1)
    assert varying == 3  # floor, back wall, emissive panel
    # the panel's triangles are lights with the emission at uv (0, 0)
    info = o.lights_info()
    assert info["n"] == 3 and _same(info["emission_area"][0, :3], np.float32([6, 5, 4]) * np.float32(2))
