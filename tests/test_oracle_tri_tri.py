"""The oracle's triangle-triangle test and BVH mesh self-intersection,
checked against the reference's own known answers and against an exact
rational separating-axis test.

Known answers: primal/tests/primal_intersect.cpp:826-1106 (3D cases, each over
the 36 corner permutations of its permuteCornersTest, :52-133),
primal_intersect.cpp:568-752 (2D cases, run here through the 3D test with the
triangles placed in a coordinate plane, which reaches the same 2D decision
tree through intersectCoplanar3DTriangles), quest/tests/quest_meshtester.cpp:
228-306 (small meshes). Fixture: tests/golden/tri_tri_known_answers.json.
"""
import json
import os
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))
KA = json.load(open(os.path.join(HERE, "golden", "tri_tri_known_answers.json")))
EPS = KA["eps"]


def permutations(a, b):
    """The 36 (t1, t2) orders of permuteCornersTest: rolls of a and b, rolls of
    their reflections, and the same with a and b exchanged."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    ap, bp = a[[0, 2, 1]], b[[0, 2, 1]]
    roll = lambda t, i: t[[i % 3, (i + 1) % 3, (i + 2) % 3]]
    out = []
    for (x, y) in ((a, b), (ap, bp), (b, a), (bp, ap)):
        for i in range(3):
            for j in range(3):
                out.append((roll(x, i), roll(y, j)))
    return out


@pytest.mark.parametrize("case", KA["tri_tri_3d"], ids=lambda c: c["name"])
def test_tri_tri_3d_known_answers(case):
    for inc, want in ((True, case["expect_include_boundary"]),
                      (False, case["expect_exclude_boundary"])):
        got = [orc.tri_tri(t1, t2, inc, EPS) for t1, t2 in permutations(case["a"], case["b"])]
        assert all(g == want for g in got), (case["name"], inc, got.count(want))


def _embed(t2d, plane):
    t = np.zeros((3, 3))
    u, v = {"xy": (0, 1), "xz": (0, 2), "zy": (2, 1)}[plane]
    t[:, u] = np.asarray(t2d, float)[:, 0]
    t[:, v] = np.asarray(t2d, float)[:, 1]
    return t


@pytest.mark.parametrize("plane", ["xy", "xz", "zy"])
@pytest.mark.parametrize("case", KA["tri_tri_2d_in_plane"], ids=lambda c: c["name"])
def test_tri_tri_coplanar_known_answers(case, plane):
    a, b = _embed(case["a"], plane), _embed(case["b"], plane)
    for inc, want in ((True, case["expect_include_boundary"]),
                      (False, case["expect_exclude_boundary"])):
        got = [orc.tri_tri(t1, t2, inc, EPS) for t1, t2 in permutations(a, b)]
        assert all(g == want for g in got), (case["name"], plane, inc, got.count(want))


# ---- exact separating-axis test (rational arithmetic) ------------------------
def _sub(p, q):
    return [p[0] - q[0], p[1] - q[1], p[2] - q[2]]


def _cross(u, v):
    return [u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]]


def _dot(u, v):
    return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]


def sat_margin(t1, t2):
    """Exact test for two closed, non-coplanar triangles: returns (intersect,
    margin) where margin is the smallest |gap or overlap| over the 11 axes,
    scaled by the axis length (how far the answer is from a touching case)."""
    P = [[Fraction(float(c)) for c in p] for p in t1]
    Q = [[Fraction(float(c)) for c in p] for p in t2]
    e1 = [_sub(P[1], P[0]), _sub(P[2], P[1]), _sub(P[0], P[2])]
    e2 = [_sub(Q[1], Q[0]), _sub(Q[2], Q[1]), _sub(Q[0], Q[2])]
    axes = [_cross(e1[0], e1[1]), _cross(e2[0], e2[1])]
    axes += [_cross(u, v) for u in e1 for v in e2]
    sep = False
    margin = None
    for ax in axes:
        n2 = _dot(ax, ax)
        if n2 == 0:
            continue
        a = [_dot(ax, p) for p in P]
        b = [_dot(ax, p) for p in Q]
        gap = max(min(b) - max(a), min(a) - max(b))  # > 0: separated on this axis
        if gap > 0:
            sep = True
        m = abs(float(gap)) / float(n2) ** 0.5
        margin = m if margin is None else min(margin, m)
    return (not sep), margin


def test_tri_tri_against_exact_separating_axes():
    rng = np.random.default_rng(7)
    checked = hits = 0
    while checked < 1500:
        t1 = rng.uniform(-1, 1, (3, 3))
        t2 = rng.uniform(-1, 1, (3, 3)) * rng.uniform(0.5, 1.2) + rng.uniform(-0.25, 0.25, 3)
        n1 = np.cross(t1[1] - t1[0], t1[2] - t1[0])
        n2 = np.cross(t2[1] - t2[0], t2[2] - t2[0])
        if np.linalg.norm(n1) < 1e-2 or np.linalg.norm(n2) < 1e-2:
            continue
        if np.linalg.norm(np.cross(n1, n2)) < 1e-3 * np.linalg.norm(n1) * np.linalg.norm(n2):
            continue
        # stay clear of the EPS-fuzzy cases: every corner well off the other plane
        d1 = (t1 - t2[2]) @ (n2 / np.linalg.norm(n2))
        d2 = (t2 - t1[2]) @ (n1 / np.linalg.norm(n1))
        if np.min(np.abs(np.concatenate([d1, d2]))) < 1e-4:
            continue
        want, margin = sat_margin(t1, t2)
        if margin is None or margin < 1e-6:
            continue
        for a, b in permutations(t1, t2)[::5]:
            assert orc.tri_tri(a, b, False, EPS) == want
            assert orc.tri_tri(a, b, True, EPS) == want
        checked += 1
        hits += want
    # both answers well represented
    assert 300 < hits < 1200


def test_tri_tri_symmetric_and_permutation_invariant_random():
    """Generic random pairs: the answer does not depend on corner order or on
    which triangle comes first (the reference's permuteCornersTest property)."""
    rng = np.random.default_rng(11)
    for _ in range(400):
        t1 = rng.uniform(-1, 1, (3, 3))
        t2 = rng.uniform(-1, 1, (3, 3)) * 0.7 + rng.uniform(-0.5, 0.5, 3)
        got = {orc.tri_tri(a, b, False, EPS) for a, b in permutations(t1, t2)}
        assert len(got) == 1


# ---- mesh self-intersection --------------------------------------------------
@pytest.mark.parametrize("case", KA["mesh_tester"], ids=lambda c: c["name"])
def test_mesh_self_intersections_known_answers(case):
    pairs, degen = orc.mesh_self_intersections(np.array(case["verts"]), np.array(case["tris"]))
    assert sorted(map(tuple, degen.reshape(-1, 1).tolist())) == [(d,) for d in case["degenerate"]]
    got = sorted(map(tuple, pairs.tolist()))
    if "note" in case:
        # The reference's BVH tester does not skip degenerate triangles (its
        # uniform-grid tester does, and the expectation is that tester's):
        # the pairs among non-degenerate triangles must agree.
        deg = set(case["degenerate"])
        got = [p for p in got if not (deg & set(p))]
    assert got == sorted(map(tuple, case["pairs"]))


def test_mesh_self_intersections_brute_force():
    """The BVH broad phase loses no pair: the same pairs as testing every
    i < j pair directly, in the same (i, then BVH order) sequence per i."""
    rng = np.random.default_rng(3)
    n = 300
    c = rng.uniform(-1, 1, (n, 3))
    v = (c[:, None, :] + rng.normal(0, 0.12, (n, 3, 3))).reshape(-1, 3)
    t = np.arange(3 * n, dtype=np.int32).reshape(n, 3)
    pairs, degen = orc.mesh_self_intersections(v, t)
    tris = v[t]
    brute = [(i, j) for i in range(n) for j in range(i + 1, n)
             if orc.tri_tri(tris[i], tris[j], False, 1e-8)]
    assert sorted(map(tuple, pairs.tolist())) == brute
    assert len(brute) > 50
    assert np.all(np.diff(pairs[:, 0]) >= 0)  # grouped by the first index
    assert degen.size == 0
