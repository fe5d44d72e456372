"""Pins the oracle's glass branch of sampleBSDF (bsdf.hpp:205-270, fresnel.hpp:
57-97, render_units.hpp:836-858) and its sphere-point, DISK, ELLIPSE and
blended-spot light sampling (light.hpp:187-335, scene_lights.hpp:176-246)
against mathematics that does not depend on the restatement:

* Fresnel: normal incidence ((1 - eta) / (1 + eta))^2, the float64 Fresnel
  equations at random angles, reciprocity F(i, eta) = F(t, 1/eta) through
  Snell's law, total internal reflection beyond the critical angle.
* Glass sampling at the smallest roughness (a near-delta interface): the
  refracted direction obeys Snell's law in the plane of incidence; the
  fraction of reflected samples equals F(cos theta, eta); the reflect /
  refract choice is the 4th number of the path's own stream (the draw
  inside sampleBSDF, bsdf.hpp:221); beyond the critical angle every sample
  reflects about the face normal; every weight G2 / (G1 + eps) lies in
  [0, 1].  A smooth glass sphere in a white furnace returns 1 up to path
  truncation, a rough one at most 1.
* Lights: E[1 / pdf] over the sampler's own (u1, u2) equals the solid angle
  the emitter subtends -- 2 pi (1 - sqrt(1 - r^2 / D^2)) for a sphere,
  2 pi (1 - h / sqrt(h^2 + R^2)) on the axis of a disk, a float64
  quadrature over an ellipse; sampled points lie on the emitter; the spot
  falloff is 0 outside the cone, 1 inside the blend, linear in cos between,
  and continuous.
* Whole paths: direct lighting of a floor point by a sphere, disk and
  ellipse light equals a float64 quadrature over the emitter; BSDF-only
  sampling ("simple", which reaches native lights by hitting them) and MIS
  agree within 3 sigma in a box lit by each of them (the method of
  test_path_tracer.cpp:193-217; traceNEE counts native lights twice,
  path_tracer.hpp:304-312, so "simple" is the unbiased partner); a spot
  light's NEE over a Lambert quad equals f I falloff cos / d^2.
"""

import ctypes as C
import math

import numpy as np
import pytest

import oracle
import scenes_small as S

L = oracle.lib()


def _f(a):
    return (C.c_float * len(a))(*[float(x) for x in a])


def _fresnel64(ci, eta):
    """Unpolarised dielectric Fresnel reflectance (textbook), float64."""
    s2 = eta * eta * (1.0 - ci * ci)
    if s2 >= 1.0:
        return 1.0
    ct = math.sqrt(1.0 - s2)
    rs = (eta * ci - ct) / (eta * ci + ct)
    rp = (ci - eta * ct) / (ci + eta * ct)
    return 0.5 * (rs * rs + rp * rp)


# --------------------------------------------------------------- Fresnel --
@pytest.mark.parametrize("eta", [1 / 1.3, 1 / 1.5, 1 / 1.7, 1.3, 1.5])
def test_fresnel_normal_incidence_closed_form(eta):
    want = ((1.0 - eta) / (1.0 + eta)) ** 2
    assert abs(L.orc_fresnel_dielectric(1.0, eta) - want) < 2e-6


def test_fresnel_equations_and_reciprocity():
    rng = np.random.default_rng(3)
    for _ in range(400):
        eta = float(rng.choice([1 / 1.5, 1 / 1.33, 1.5, 1.33]))
        ci = float(rng.uniform(0.0, 1.0))
        got = L.orc_fresnel_dielectric(ci, eta)
        assert abs(got - _fresnel64(ci, eta)) < 2e-4, (ci, eta)
        s2 = eta * eta * (1 - ci * ci)
        if s2 < 0.98:  # away from the critical angle: the way back through the interface reflects the same
            ct = math.sqrt(1 - s2)
            assert abs(L.orc_fresnel_dielectric(ct, 1 / eta) - got) < 2e-4


def test_fresnel_total_internal_reflection():
    for ior in (1.3, 1.5, 1.7):
        ccrit = math.sqrt(1 - 1 / ior ** 2)  # cos of the critical angle, glass -> air
        for ci in np.linspace(0.0, ccrit - 1e-3, 20):
            assert L.orc_fresnel_dielectric(float(ci), ior) == 1.0
        assert L.orc_fresnel_dielectric(ccrit + 1e-2, ior) < 1.0


# ------------------------------------------------------------ glass BSDF --
def _glass(rough=0.0, trans=1.0, ior=1.5):
    return _f([0.9, 0.9, 0.9, 0.0, rough, trans, ior])


def _sample(mat, wo, n, u1, u2, u3, seed):
    out = (C.c_float * 9)()
    L.orc_sample_bsdf(mat, _f(wo), _f(n), u1, u2, u3, seed[0], seed[1], seed[2], out)
    return np.array(out[0:3], np.float64), int(out[3]), np.array(out[4:7]), float(out[7]), int(out[8])


def _first_draw(sx, sy, ss):
    out = (C.c_float * 1)()
    L.orc_randf_stream(sx, sy, ss, 1, out)
    return out[0]


@pytest.mark.parametrize("theta_deg, ior", [(40.0, 1.5), (60.0, 1.5), (25.0, 1.33), (75.0, 1.7)])
def test_smooth_glass_snell_fresnel_and_fourth_draw(theta_deg, ior):
    th = math.radians(theta_deg)
    wo = (math.sin(th), 0.0, math.cos(th))  # outside, front face
    n = (0.0, 0.0, 1.0)
    eta = 1.0 / ior
    F0 = _fresnel64(math.cos(th), eta)
    rng = np.random.default_rng(int(theta_deg))
    N = 12000
    nrefl = off = 0
    for k in range(N):
        u1, u2 = (float(x) for x in rng.random(2))
        wi, kind, w, pdf, typ = _sample(_glass(ior=ior), wo, n, u1, u2, 0.0, (7, 9, k))
        assert kind == 1 and 0.0 <= w[0] <= 1.0 + 1e-6
        # GGX's long tail at alpha = 1e-4 tilts a few microfacet normals by
        # more than 1e-3 rad: the directions are checked on 99 % of samples
        if typ == 1:  # reflection: the mirror direction
            nrefl += 1
            off += not np.allclose(wi, (-wo[0], 0.0, wo[2]), atol=2e-3)
        else:  # refraction: Snell's law in the plane of incidence
            assert typ == 2 and wi[2] < 0
            sin_t = math.hypot(wi[0], wi[1])
            off += not (abs(sin_t - eta * math.sin(th)) < 2e-3 and abs(wi[1]) < 2e-3 and wi[0] < 0)
        u4 = _first_draw(7, 9, k)  # the branch's own draw from the path's stream
        if abs(u4 - F0) > 2e-2:
            assert (typ == 1) == (u4 < F0), (k, u4, F0, typ)
    assert off <= 0.01 * N, off
    frac = nrefl / N
    assert abs(frac - F0) < 4 * math.sqrt(F0 * (1 - F0) / N) + 1e-3, (frac, F0)


@pytest.mark.parametrize("ior", [1.33, 1.5])
def test_glass_total_internal_reflection_from_inside(ior):
    th = math.radians(60.0)  # beyond the critical angle of glass -> air
    wo = (math.sin(th), 0.0, -math.cos(th))  # inside: wo . n < 0, so the face normal is -n
    n = (0.0, 0.0, 1.0)
    rng = np.random.default_rng(5)
    off = 0
    for k in range(2000):
        u1, u2 = (float(x) for x in rng.random(2))
        wi, kind, w, pdf, typ = _sample(_glass(ior=ior), wo, n, u1, u2, 0.5, (1, 2, k))
        assert typ == 1 and kind == 1 and wi[2] < 0  # reflected, stays inside
        off += not np.allclose(wi, (-wo[0], 0.0, wo[2]), atol=2e-3)  # mirrored about the face normal
    assert off <= 20, off


def test_rough_glass_weights_bounded():
    rng = np.random.default_rng(11)
    th = math.radians(50.0)
    for wo, n in (((math.sin(th), 0.0, math.cos(th)), (0, 0, 1)), ((math.sin(th), 0.0, -math.cos(th)), (0, 0, 1))):
        ws = []
        for k in range(6000):
            u1, u2 = (float(x) for x in rng.random(2))
            wi, kind, w, pdf, typ = _sample(_glass(rough=0.3), wo, n, u1, u2, 0.0, (3, 4, k))
            ws.append(w[0] if kind == 1 else 0.0)
            assert abs(np.linalg.norm(wi) - 1) < 1e-5
        ws = np.array(ws)
        assert ws.min() >= 0.0 and ws.max() <= 1.0 + 1e-6
        assert 0.6 < ws.mean() <= 1.0  # masking only loses energy


def test_partial_transmission_picks_glass_with_probability_transmission():
    """u3 < transmission selects the glass branch (bsdf.hpp:205)."""
    wo = (0.3, 0.0, math.sqrt(1 - 0.09))
    mat = _glass(rough=0.2, trans=0.4)
    for u3, glass in ((0.0, True), (0.399, True), (0.401, False), (0.9, False)):
        wi, kind, w, pdf, typ = _sample(mat, wo, (0, 0, 1), 0.3, 0.7, u3, (1, 1, 1))
        if glass:
            assert kind == 1 and typ in (1, 2)
        else:
            assert kind == 0 and typ in (0, 1)  # the opaque lobes' pdf carries the (1 - transmission) factor
            f = (C.c_float * 3)()
            p = (C.c_float * 1)()
            L.orc_eval_bsdf(mat, _f(wo), _f(wi), _f((0, 0, 1)), f, p)
            assert abs(p[0] - pdf) <= 1e-6 * max(1.0, pdf)


def _ico_sphere(center, radius, material, level=3):
    import workloads

    v, t = workloads._icosphere(level)
    pos = v * radius + np.asarray(center)
    vn = v[t]  # exact sphere normals per corner
    return {"vertices": pos.tolist(), "triangles": t.tolist(), "triangle_normals": vn.tolist(), "smooth": True,
            "material": material}


@pytest.mark.parametrize("rough, lo, hi", [(0.0, 0.93, 1.0), (0.3, 0.6, 1.0)])
def test_glass_sphere_white_furnace(rough, lo, hi):
    glass = {"base_color": [1, 1, 1], "metallic": 0.0, "roughness": rough, "transmission": 1.0, "ior": 1.5}
    scene = {"meshes": [_ico_sphere((0, 0, 0), 1.0, glass)], "environment": {"color": [1, 1, 1], "strength": 1.0}}
    cam = {"pos": (0.0, 0.0, 4.0), "dir": (0.0, 0.0, -1.0), "up": (0.0, 1.0, 0.0), "fov": 20.0}
    o = oracle.OracleScene(scene)
    o.set_camera(cam["pos"], cam["dir"], cam["up"], cam["fov"])
    n = 6000
    for algo in ("simple", "mis"):
        v = o.trace_pixel(algo, 3, 4, 8, 8, 1, 0, n, 32).astype(np.float64)[:, 0]
        se = v.std() / math.sqrt(n)
        assert lo - 3 * se < v.mean() < hi + 3 * se, (algo, v.mean(), se)


# ---------------------------------------------------------------- lights --
def _light_scene(light):
    return oracle.OracleScene({"meshes": [], "native_lights": [light]})


def _inv_pdf_mean(o, p, k=160):
    acc = 0.0
    pts = []
    for i in range(k):
        for j in range(k):
            s = o.sample_light(0, p, (i + 0.5) / k, (j + 0.5) / k)
            pts.append(s[:3])
            if s[7] > 0:
                acc += 1.0 / s[7]
    return acc / (k * k), np.array(pts, np.float64)


def test_sphere_light_subtends_its_solid_angle():
    c, r = np.array([0.3, 2.0, -0.2]), 0.25
    o = _light_scene({"type": "POINT", "position": c.tolist(), "energy": 50.0, "radius": r})
    for p in ([0.0, 0.0, 0.0], [1.5, 1.0, 0.5]):
        D = np.linalg.norm(c - np.array(p))
        omega, pts = _inv_pdf_mean(o, p)
        want = 2 * math.pi * (1 - math.sqrt(1 - (r / D) ** 2))
        assert abs(omega - want) / want < 3e-3, (omega, want)
        assert np.allclose(np.linalg.norm(pts - c, axis=1), r, atol=1e-5)


def test_disk_light_on_axis_solid_angle():
    R, h = 0.4, 1.3
    o = _light_scene({"type": "AREA", "shape": "DISK", "size": 2 * R, "position": [0, h, 0], "direction": [0, -1, 0],
                      "right": [1, 0, 0], "up": [0, 0, 1], "energy": 10.0})
    omega, pts = _inv_pdf_mean(o, [0.0, 0.0, 0.0])
    want = 2 * math.pi * (1 - h / math.sqrt(h * h + R * R))
    assert abs(omega - want) / want < 2e-3, (omega, want)
    assert np.all(np.hypot(pts[:, 0], pts[:, 2]) <= R + 1e-6) and np.allclose(pts[:, 1], h)
    assert np.hypot(pts[:, 0], pts[:, 2]).max() > 0.99 * R


def _ellipse_solid_angle64(a, b, h, p, k=600):
    """Solid angle of the ellipse (x/a)^2 + (z/b)^2 <= 1 in the plane y = h,
    facing -y, seen from p: float64 quadrature of cos / d^2 dA."""
    rho = (np.arange(k) + 0.5) / k
    phi = 2 * math.pi * (np.arange(k) + 0.5) / k
    R, P = np.meshgrid(rho, phi)
    x, z = a * R * np.cos(P), b * R * np.sin(P)
    dx, dy, dz = x - p[0], h - p[1], z - p[2]
    d2 = dx * dx + dy * dy + dz * dz
    cos_l = dy / np.sqrt(d2)
    return float((cos_l / d2 * a * b * R).sum() * (1.0 / k) * (2 * math.pi / k))


def test_ellipse_light_solid_angle_quadrature():
    a, b, h = 0.5, 0.2, 1.1
    o = _light_scene({"type": "AREA", "shape": "ELLIPSE", "size": 2 * a, "size_y": 2 * b, "position": [0, h, 0],
                      "direction": [0, -1, 0], "right": [1, 0, 0], "up": [0, 0, 1], "energy": 10.0})
    for p in ([0.0, 0.0, 0.0], [0.7, 0.2, -0.4]):
        omega, pts = _inv_pdf_mean(o, p)
        want = _ellipse_solid_angle64(a, b, h, p)
        assert abs(omega - want) / want < 3e-3, (omega, want)
    assert np.all((pts[:, 0] / a) ** 2 + (pts[:, 2] / b) ** 2 <= 1 + 1e-5)
    assert np.abs(pts[:, 0]).max() > 0.99 * a and np.abs(pts[:, 2]).max() > 0.99 * b


def test_spot_blend_profile():
    size, blend = 1.0, 0.3
    o = _light_scene({"type": "SPOT", "position": [0, 0, 0], "direction": [0, -1, 0], "energy": 40.0,
                      "spot_size": size, "spot_blend": blend})
    I = o.lights_info()["emission_area"][0, 0]
    cc = math.cos(size / 2)
    prev = None
    for th in np.linspace(0.0, size / 2 + 0.1, 121):
        p = [2.0 * math.sin(th), -2.0 * math.cos(th), 0.0]
        e = o.sample_light(0, p, 0.5, 0.5)[8] * 4.0  # undo 1 / d^2 (d = 2)
        t = (math.cos(th) - cc) / (1 - cc)
        if th > size / 2 + 1e-4:
            assert e == 0.0
        elif t >= blend + 1e-3:
            assert abs(e - I) < 1e-4 * I
        elif t > 1e-3:
            assert 0.0 < e < I and abs(e / I - t / blend) < 2e-3
        if prev is not None:
            assert e <= prev + 1e-4 * I  # monotone away from the axis
        prev = e


# ---------------------------------------------------------- whole paths --
def _lambert_ggx_f(albedo, cos_l, cos_v, cos_h, vh, rough=1.0):
    a = rough * rough
    a2 = a * a
    den = cos_h * cos_h * (a2 - 1) + 1
    D = a2 / math.pi / (den * den + 1e-6)
    ll = math.sqrt(a2 + (1 - a2) * cos_l * cos_l)
    lv = math.sqrt(a2 + (1 - a2) * cos_v * cos_v)
    g2d = 0.5 / (cos_v * ll + cos_l * lv + 1e-6)
    F = 0.04 + 0.96 * (1 - vh) ** 5
    return albedo / math.pi * (1 - F) + D * g2d * F


_FLOOR = {"vertices": [[-3, 0, -3], [3, 0, -3], [3, 0, 3], [-3, 0, 3]], "triangles": [[0, 2, 1], [0, 3, 2]],
          "material": {"base_color": [0.6, 0.6, 0.6], "roughness": 1.0}}
_DOWN = {"pos": (0.0, 3.0, 0.0), "dir": (0.0, -1.0, 0.0), "up": (0.0, 0.0, -1.0), "fov": 30.0}


def _direct(light, n=20000):
    o = oracle.OracleScene({"meshes": [_FLOOR], "native_lights": [light],
                            "environment": {"color": [0, 0, 0], "strength": 1.0}})
    o.set_camera(_DOWN["pos"], _DOWN["dir"], _DOWN["up"], _DOWN["fov"])
    v = o.trace_pixel("nee", 0, 0, 1, 1, 1, 0, n, 1)[:, 0].astype(np.float64)
    return v.mean(), v.std() / math.sqrt(n), o.lights_info()["emission_area"][0, 0]


def _quad_points(points, normals, areas, Le):
    """float64 sum of f Le cos_s cos_l / d^2 dA at the floor point (0,0,0)."""
    q = 0.0
    for (x, y, z), nl, dA in zip(points, normals, areas):
        d = math.sqrt(x * x + y * y + z * z)
        wi = np.array([x, y, z]) / d
        cl = -float(np.dot(nl, wi))
        cs = wi[1]
        if cl <= 0 or cs <= 0:
            continue
        h = wi + np.array([0, 1, 0])
        h /= np.linalg.norm(h)
        q += _lambert_ggx_f(0.6, cs, 1.0, h[1], h[1]) * Le * cs * cl / (d * d) * dA
    return q


def test_disk_and_ellipse_direct_light_quadrature():
    for shape, a, b in (("DISK", 0.3, 0.3), ("ELLIPSE", 0.45, 0.2)):  # size_y = size for a DISK
        c = np.array([0.25, 1.0, -0.1])
        light = {"type": "AREA", "shape": shape, "size": 2 * a, "size_y": 2 * b, "position": c.tolist(),
                 "direction": [0, -1, 0], "right": [1, 0, 0], "up": [0, 0, 1], "energy": 5.0}
        m, se, Le = _direct(light)
        k = 120
        pts, nrm, dA = [], [], []
        for i in range(k):
            rho = (i + 0.5) / k
            for j in range(k):
                ph = 2 * math.pi * (j + 0.5) / k
                pts.append(c + np.array([a * rho * math.cos(ph), 0.0, b * rho * math.sin(ph)]))
                nrm.append(np.array([0.0, -1.0, 0.0]))
                dA.append(a * b * rho * (1.0 / k) * (2 * math.pi / k))
        q = _quad_points(pts, nrm, dA, Le)
        assert abs(m - q) < 4 * se + 2e-3 * q, (shape, m, q, se)


def test_sphere_light_direct_quadrature():
    c, r = np.array([-0.2, 1.2, 0.3]), 0.2
    m, se, Le = _direct({"type": "POINT", "position": c.tolist(), "energy": 20.0, "radius": r})
    k = 160
    pts, nrm, dA = [], [], []
    for i in range(k):
        th = math.pi * (i + 0.5) / k
        for j in range(2 * k):
            ph = 2 * math.pi * (j + 0.5) / (2 * k)
            u = np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
            pts.append(c + r * u)
            nrm.append(u)
            dA.append(r * r * math.sin(th) * (math.pi / k) * (math.pi / k))
    q = _quad_points(pts, nrm, dA, Le)
    assert abs(m - q) < 4 * se + 3e-3 * q, (m, q, se)


@pytest.mark.parametrize("kind", ["sphere", "disk", "ellipse"])
def test_simple_and_mis_agree_native_lights_3sigma(kind):
    scene = S.cornell_box()
    scene["meshes"] = scene["meshes"][:-1]
    if kind == "sphere":
        light = {"type": "POINT", "position": [0.2, 1.6, 0.1], "energy": 15.0, "radius": 0.12}
    else:  # the exporter sends size_y = size for a DISK (scene_export.py:733-738)
        light = {"type": "AREA", "shape": kind.upper(), "position": [0, 1.99, 0], "direction": [0, -1, 0],
                 "right": [1, 0, 0], "up": [0, 0, 1], "size": 0.6, "size_y": 0.6 if kind == "disk" else 0.3,
                 "energy": 8.0}
    scene["native_lights"] = [light]
    cam = {"pos": (0.0, 1.0, 0.8), "dir": (0.0, -0.5, -0.5), "up": (0.0, 1.0, 0.0), "fov": 10.0}
    o = oracle.OracleScene(scene)
    o.set_camera(cam["pos"], cam["dir"], cam["up"], cam["fov"])
    n = 4096
    lum = np.float64([0.2126, 0.7152, 0.0722])
    a = o.trace_pixel("simple", 0, 0, 1, 1, 1, 0, 8 * n, 8).astype(np.float64) @ lum
    b = o.trace_pixel("mis", 0, 0, 1, 1, 1, 50000, n, 8).astype(np.float64) @ lum
    se = math.sqrt(a.var() / (8 * n) + b.var() / n)
    assert b.mean() > 0.02 and abs(a.mean() - b.mean()) < 3 * se, (a.mean(), b.mean(), se)


def test_spot_light_closed_form_on_lambert_quad():
    """NEE at the quad point straight below a tilted spot: f * I * falloff *
    cos / d^2 with the falloff of the angle between the spot axis and the
    shading point (the spot is a delta light: NEE is exact per sample)."""
    energy, size, blend = 10.0, 1.2, 0.4
    quad = {"vertices": [[-2, -2, -1], [2, -2, -1], [2, 2, -1], [-2, 2, -1]], "triangles": [[0, 1, 2], [0, 2, 3]],
            "material": {"base_color": [0.7, 0.7, 0.7], "roughness": 1.0}}
    for tilt in (0.0, 0.3, 0.45, 0.55, 0.7):
        axis = [math.sin(tilt), 0.0, -math.cos(tilt)]
        scene = {"meshes": [quad], "environment": {"color": [0, 0, 0], "strength": 1.0},
                 "native_lights": [{"type": "SPOT", "position": [0, 0, 0], "direction": axis, "energy": energy,
                                    "spot_size": size, "spot_blend": blend}]}
        o = oracle.OracleScene(scene)
        o.set_camera(S.CAM_DOWN_Z["pos"], S.CAM_DOWN_Z["dir"], S.CAM_DOWN_Z["up"], S.CAM_DOWN_Z["fov"])
        I = o.lights_info()["emission_area"][0, 0]
        v = o.trace_pixel("nee", 0, 0, 1, 1, 1, 0, 4, 1)
        cc = math.cos(size / 2)
        t = (math.cos(tilt) - cc) / (1 - cc)
        fall = 0.0 if math.cos(tilt) < cc else min(1.0, t / blend)
        want = _lambert_ggx_f(0.7, 1.0, 1.0, 1.0, 1.0) * I * fall
        np.testing.assert_allclose(v[:, 0], want, rtol=2e-5, atol=1e-7)
