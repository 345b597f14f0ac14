"""Pins the CPU oracle against the known answers held by the reference's tests.

Each check names the reference test it transcribes (see
tests/golden/reference_known_answers.json for file:line).
"""
import json
import math
import os

import numpy as np
import pytest

from oracle import oracle as orc
from tests import refmeshes

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_known_answers.json")))


def interleave_by_bits(x, y, z):
    """Bit-by-bit interleave (x at bit 0, y at 1, z at 2) -- independent of the
    magic-number spread used by the oracle."""
    code = 0
    for b in range(10):
        code |= ((x >> b) & 1) << (3 * b)
        code |= ((y >> b) & 1) << (3 * b + 1)
        code |= ((z >> b) & 1) << (3 * b + 2)
    return code


def test_morton_known_answer():
    for case in GOLD["morton"]:
        assert orc.morton3(*case["point"]) == case["code"]


def test_morton_matches_bitwise_interleave():
    rng = np.random.default_rng(7)
    pts = rng.integers(0, 1024, size=(2000, 3))
    pts = np.concatenate([pts, [[0, 0, 0], [1023, 1023, 1023], [1023, 0, 0], [0, 1023, 0]]])
    for x, y, z in pts:
        assert orc.morton3(x, y, z) == interleave_by_bits(int(x), int(y), int(z))


def test_closest_point_known_answers():
    for group in GOLD["closest_point_triangle"]:
        a, b, c = (np.array(p, float) for p in group["tri"])
        for case in group["cases"]:
            cp, loc = orc.closest_point(case["q"], a, b, c, group["eps"])
            assert loc == case["loc"], (group["source"], case)
            np.testing.assert_allclose(cp, case["cp"], rtol=0, atol=case["tol"])


def _grid_bvh():
    g = GOLD["bvh_uniform_grid"]
    boxes = refmeshes.uniform_grid_cell_boxes(g["mesh"]["lo"], g["mesh"]["hi"],
                                              g["mesh"]["nodes"][0])
    return g, boxes, orc.OracleBVH(boxes, scale=g["scale"])


def test_bvh_grid_bounds():
    g, boxes, bvh = _grid_bvh()
    b = bvh.export()["bounds"]
    np.testing.assert_allclose(b[:3], g["bounds"]["lo"], atol=g["bounds"]["tol"])
    np.testing.assert_allclose(b[3:], g["bounds"]["hi"], atol=g["bounds"]["tol"])


def test_bvh_grid_find_boxes():
    g, boxes, bvh = _grid_bvh()
    q = np.array([qb["min"] + qb["max"] for qb in g["query_boxes"]])
    off, cnt, cand = bvh.find("boxes", q)
    for i, qb in enumerate(g["query_boxes"]):
        got = sorted(cand[off[i]:off[i] + cnt[i]].tolist())
        assert got == qb["expected"]


def test_bvh_grid_find_rays():
    g, boxes, bvh = _grid_bvh()
    rays = []
    for r in g["query_rays"]:
        rays.append(list(r["origin"]) + list(orc.unit_vector(r["direction"])))
    off, cnt, cand = bvh.find("rays", np.array(rays))
    for i, r in enumerate(g["query_rays"]):
        got = sorted(cand[off[i]:off[i] + cnt[i]].tolist())
        assert got == r["expected"]


def test_bvh_grid_find_points():
    g, boxes, bvh = _grid_bvh()
    cent = 0.5 * (boxes[:, :3] + boxes[:, 3:])
    off, cnt, cand = bvh.find("points", cent)
    assert np.all(cnt == 1)
    assert np.array_equal(cand[off], np.arange(boxes.shape[0]))
    off, cnt, cand = bvh.find("points", cent + 10.0)
    assert np.all(cnt == 0) and cand.size == 0


def test_bvh_single_and_empty():
    # spin_bvh.cpp:1009-1160 / :1401-1556: 0 or 1 box pads to 2 leaves.
    one = np.array([[0.0, 0.0, 0.0, 1.0, 1.0, 1.0]])
    bvh = orc.OracleBVH(one, scale=1.0)
    assert bvh.n == 2
    off, cnt, cand = bvh.find("points", np.array([[0.5, 0.5, 0.5], [2.0, 2.0, 2.0]]))
    assert cnt.tolist() == [1, 0] and cand.tolist() == [0]
    empty = orc.OracleBVH(np.zeros((0, 6)), scale=1.0)
    off, cnt, cand = empty.find("points", np.array([[0.5, 0.5, 0.5]]))
    assert cnt.tolist() == [0]


def test_sdf_sphere_known_norms():
    s = GOLD["sdf_sphere"]
    v, t = refmeshes.sphere_uv_mesh(s["center"], s["radius"], s["theta_res"], s["phi_res"])
    sdf = orc.OracleSDF(v, t, s["watertight"], s["compute_signs"])
    lo = v.min(0) - s["grid_expand"]
    hi = v.max(0) + s["grid_expand"]
    q = refmeshes.uniform_grid_nodes(lo, hi, s["grid_nodes"])
    phi = sdf.eval(q)["dist"]
    exact = np.linalg.norm(q - np.array(s["center"]), axis=1) - s["radius"]
    diff = phi - exact
    assert np.all(np.abs(diff) <= s["per_point_tol"])
    l1 = np.sum(np.abs(diff))
    l2 = math.sqrt(np.sum(diff))  # the reference sums signed differences
    linf = np.max(np.abs(diff))
    assert abs(l1 - s["l1"]) <= s["tol"]
    assert abs(l2 - s["l2_of_sum"]) <= s["tol"]
    assert abs(linf - s["linf"]) <= s["tol"]


def test_sdf_sphere_normals_and_cp():
    # quest_signed_distance.cpp:172-292 checks on the same sphere.
    s = GOLD["sdf_sphere"]
    v, t = refmeshes.sphere_uv_mesh(s["center"], s["radius"], s["theta_res"], s["phi_res"])
    sdf = orc.OracleSDF(v, t, True, True)
    lo = v.min(0) - s["grid_expand"]
    hi = v.max(0) + s["grid_expand"]
    q = refmeshes.uniform_grid_nodes(lo, hi, s["grid_nodes"])
    out = sdf.eval(q)
    cp_exact = s["radius"] * q / np.linalg.norm(q, axis=1)[:, None]
    assert np.all(np.sum((out["cp"] - cp_exact) ** 2, axis=1) <= 1e-2)
    n_exact = q / np.linalg.norm(q, axis=1)[:, None]
    assert np.all(np.sum(out["normal"] * n_exact, axis=1) >= 0.99)
    np.testing.assert_allclose(out["dist"] ** 2, np.sum((q - out["cp"]) ** 2, axis=1),
                               rtol=0, atol=1e-8)


def _ulps(a, b):
    ia = np.asarray(a, np.float64).view(np.int64)
    ib = np.asarray(b, np.float64).view(np.int64)
    ia = np.where(ia < 0, np.int64(-(2 ** 63)) - ia, ia)
    ib = np.where(ib < 0, np.int64(-(2 ** 63)) - ib, ib)
    return np.abs(ia - ib)


def test_sdf_plane_exact():
    p = GOLD["sdf_plane"]
    v, t = refmeshes.plane_soup(p["triangles"])
    sdf = orc.OracleSDF(v, t, watertight=p["closed_surface"], compute_sign=True)
    q = refmeshes.uniform_grid_nodes(p["grid"]["lo"], p["grid"]["hi"], p["grid"]["nodes"])
    phi = sdf.eval(q)["dist"]
    assert np.all(_ulps(phi, q[:, 2]) <= 4)
