"""Pins the oracle's geometry (geometry/intersection.hpp, geometry/bvh.hpp,
core/scene.hpp BVH build).

* Moller-Trumbore: the reference's known answers (test_geometry.cpp:58-214)
  and, on random rays, the float64 solution of o + t d = a + u e1 + v e2.
* AABB slab and sphere: reference known answers (test_geometry.cpp:221-415).
* BVH build + closest-hit traversal: brute force over all triangles of a
  random soup, and the reference's tie rule (the triangle tested last, i.e.
  the largest (mesh, BVH position), wins exact ties).
"""

import ctypes as C

import numpy as np
import pytest

import oracle

L = oracle.lib()
FP = C.POINTER(C.c_float)


def _f(a):
    return (C.c_float * len(a))(*a)


def tri(o, d, a=(0, 0, 0), b=(1, 0, 0), c=(0, 1, 0), tmin=1e-4, tmax=1e30):
    out = (C.c_float * 3)()
    hit = L.orc_intersect_triangle(_f(o), _f(d), _f(a), _f(b), _f(c), tmin, tmax, out)
    return (tuple(out) if hit else None)


def test_triangle_known_answers():
    t, u, v = tri((0.25, 0.25, 1.0), (0, 0, -1))            # HitCenter
    assert abs(t - 1.0) < 1e-6 and abs(u - 0.25) < 1e-6 and abs(v - 0.25) < 1e-6
    t, u, v = tri((0.0, 0.0, 1.0), (0, 0, -1))              # HitVertex0
    assert u == 0.0 and v == 0.0
    t, u, v = tri((1.0, 0.0, 1.0), (0, 0, -1))              # HitVertex1
    assert abs(u - 1.0) < 1e-6 and v == 0.0
    t, u, v = tri((0.5, 0.0, 1.0), (0, 0, -1))              # HitEdge01
    assert abs(u - 0.5) < 1e-6 and v == 0.0
    assert tri((2.0, 2.0, 1.0), (0, 0, -1)) is None          # MissOutside
    assert tri((0.25, 0.25, 1.0), (1, 0, 0)) is None         # MissParallel
    assert tri((0.25, 0.25, 1.0), (0, 0, 1)) is None         # MissBehind
    assert tri((0.25, 0.25, -1.0), (0, 0, 1)) is not None    # HitFromBehind
    assert tri((0.5, 0.0, 1.0), (0, 0, -1), (0, 0, 0), (1, 0, 0), (2, 0, 0)) is None  # Degenerate
    assert tri((0.25, 0.25, 1.0), (0, 0, -1), tmin=2.0) is None   # TMinConstraint
    assert tri((0.25, 0.25, 1.0), (0, 0, -1), tmax=0.5) is None   # TMaxConstraint
    assert tri((0.25, 0.25, 1.0), (0, 0, -1), tmax=1.0) is not None  # tMax is inclusive


def test_triangle_against_float64_solution():
    rng = np.random.default_rng(3)
    n_checked = 0
    for _ in range(400):
        a, b, c = rng.normal(size=(3, 3))
        o = rng.normal(size=3) * 3
        target = a + 0.2 * (b - a) + 0.3 * (c - a) + rng.normal(size=3) * 0.05
        d = target - o
        d /= np.linalg.norm(d)
        a32, b32, c32, o32, d32 = (np.float32(x) for x in (a, b, c, o, d))
        A = np.stack([b32 - a32, c32 - a32, -d32], axis=1).astype(np.float64)
        if abs(np.linalg.det(A)) < 1e-3:
            continue
        u, v, t = np.linalg.solve(A, (o32 - a32).astype(np.float64))
        inside = u > 1e-4 and v > 1e-4 and u + v < 1 - 1e-4 and t > 1e-3
        got = tri(o32.tolist(), d32.tolist(), a32.tolist(), b32.tolist(), c32.tolist())
        if inside:
            assert got is not None
            assert abs(got[0] - t) < 1e-4 * max(1, t) and abs(got[1] - u) < 1e-4 and abs(got[2] - v) < 1e-4
            n_checked += 1
        elif u < -1e-3 or v < -1e-3 or u + v > 1 + 1e-3:
            assert got is None
    assert n_checked > 100


def aabb(o, d, bmin=(-1, -1, -1), bmax=(1, 1, 1)):
    return bool(L.orc_intersect_aabb(_f(o), _f(d), _f(bmin), _f(bmax)))


def test_aabb_known_answers():
    s = 1 / np.sqrt(3)
    assert aabb((0, 0, 5), (0, 0, -1))                  # HitFromFront
    assert aabb((5, 0, 0), (-1, 0, 0))                  # HitFromSide
    assert aabb((5, 5, 5), (-s, -s, -s))                # HitDiagonal
    assert not aabb((0, 5, 5), (0, 0, -1))              # MissAbove
    assert not aabb((5, 5, 0), (-1, 0, 0))              # MissSide
    assert aabb((0, 0, 0), (1, 0, 0))                   # RayInsideBox
    assert aabb((-5, 0, 0), (1, 0, 0))                  # ParallelToXAxis
    assert not aabb((-5, 2, 0), (1, 0, 0))              # ParallelToXAxisMiss
    assert not aabb((0, 0, 5), (0, 0, 1))               # box behind the origin


def test_sphere_known_answers():
    t = (C.c_float * 1)()
    assert L.orc_intersect_sphere(_f((0, 0, 5)), _f((0, 0, -1)), _f((0, 0, 0)), 1.0, t) and abs(t[0] - 4) < 1e-6
    assert L.orc_intersect_sphere(_f((0, 0, 0)), _f((0, 0, -1)), _f((0, 0, 0)), 1.0, t) and abs(t[0] - 1) < 1e-6
    assert not L.orc_intersect_sphere(_f((0, 5, 5)), _f((0, 0, -1)), _f((0, 0, 0)), 1.0, t)
    assert not L.orc_intersect_sphere(_f((0, 0, 5)), _f((0, 0, 1)), _f((0, 0, 0)), 1.0, t)


def _soup(rng, n_meshes=3, n=150):
    meshes = []
    for m in range(n_meshes):
        c = rng.uniform(-1, 1, size=(n, 1, 3))
        v = (c + rng.normal(size=(n, 3, 3)) * 0.15).reshape(-1, 3)
        meshes.append({"vertices": v.astype(np.float32).tolist(),
                       "triangles": np.arange(3 * n).reshape(n, 3).tolist(),
                       "material": {"base_color": [0.5, 0.5, 0.5]}})
    return {"meshes": meshes}


def test_closest_hit_matches_brute_force():
    rng = np.random.default_rng(11)
    scene = _soup(rng)
    o = oracle.OracleScene(scene)
    n = 3000
    org = rng.uniform(-3, 3, size=(n, 3)).astype(np.float32)
    dst = rng.uniform(-1, 1, size=(n, 3)).astype(np.float32)
    d = dst - org
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    got = o.intersect(org, d, 1e-4, 1e30)
    # brute force with the same float32 Moller-Trumbore over every triangle
    tris = [(mi, np.asarray(m["vertices"], np.float32).reshape(-1, 3, 3)) for mi, m in enumerate(scene["meshes"])]
    mism = 0
    for i in range(0, n, 7):
        best = (np.inf, -1, -1)
        for mi, T in tris:
            for ti, (a, b, c) in enumerate(T):
                h = tri(org[i].tolist(), d[i].tolist(), a.tolist(), b.tolist(), c.tolist())
                if h is not None and h[0] < best[0]:
                    best = (h[0], mi, ti)
        if best[1] < 0:
            mism += got["mesh"][i] != -1
        else:
            mism += not (got["mesh"][i] == best[1] and got["tri"][i] == best[2] and got["t"][i] == np.float32(best[0]))
    assert mism == 0


def test_exact_tie_goes_to_last_tested_triangle():
    # two identical triangles in one mesh and a copy in a later mesh: the hit
    # must come from the later mesh (largest mesh index, then BVH position).
    t = [[0, 0, 0], [1, 0, 0], [0, 1, 0]]
    scene = {"meshes": [{"vertices": t + t, "triangles": [[0, 1, 2], [3, 4, 5]]},
                        {"vertices": t, "triangles": [[0, 1, 2]]}]}
    o = oracle.OracleScene(scene)
    h = o.intersect(np.float32([[0.2, 0.2, 1.0]]), np.float32([[0, 0, -1]]), 1e-4, 1e30)
    assert h["mesh"][0] == 1 and h["tri"][0] == 0
    one = oracle.OracleScene({"meshes": [scene["meshes"][0]]})
    h = one.intersect(np.float32([[0.2, 0.2, 1.0]]), np.float32([[0, 0, -1]]), 1e-4, 1e30)
    boxes, links, order = one.mesh_bvh(0)
    assert h["pos"][0] == 1 and h["tri"][0] == order[1]


def test_bvh_structure_median_split():
    rng = np.random.default_rng(2)
    scene = _soup(rng, n_meshes=1, n=100)
    o = oracle.OracleScene(scene)
    boxes, links, order = o.mesh_bvh(0)
    assert sorted(order.tolist()) == list(range(100))
    leaves = links[links[:, 3] > 0]
    assert leaves[:, 3].max() <= 4 and leaves[:, 3].sum() == 100
    # left-first DFS: leaf ranges appear in increasing order and tile [0, n)
    def walk(i, acc):
        l, r, s, c = links[i]
        if c > 0:
            acc.append((s, c))
        else:
            walk(l, acc)
            walk(r, acc)
    acc = []
    walk(0, acc)
    pos = 0
    for s, c in acc:
        assert s == pos
        pos += c
    V = np.asarray(scene["meshes"][0]["vertices"], np.float32).reshape(-1, 3, 3)
    for i, (l, r, s, c) in enumerate(links):  # node boxes bound their triangles
        if c > 0:
            pts = V[order[s:s + c]].reshape(-1, 3)
            assert np.all(pts.min(0) == boxes[i, :3]) and np.all(pts.max(0) == boxes[i, 3:])
