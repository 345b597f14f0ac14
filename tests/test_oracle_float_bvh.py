"""Pins the oracle's float tree (spin::BVH<NDIMS, Exec, float>).

The reference instantiates every BVH test with FloatType = float as well
(spin/tests/spin_bvh.cpp:1560-1669) and expects the double answers: the
known answers below run through the float tree. What the known answers
leave open -- where float rounding of the scaling, the Morton normalisation
and the ray test differs from double -- is pinned by an independent NumPy
float32 statement of those steps and by brute force over the float leaf
boxes, on inputs chosen so that float and double give different trees."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as orc
from tests import refmeshes

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_known_answers.json")))
F = np.float32


def unit_vector_f32(d):
    """Vector<float>::unitVector (Vector.hpp:477-490, NumericArray.hpp:498-514):
    the squared norm summed in float, then v * (1 / double(float(sqrt)))."""
    d = np.asarray(d, F)
    len_sq = np.zeros(d.shape[:-1], F)
    for k in range(d.shape[-1]):
        len_sq = (len_sq + d[..., k] * d[..., k]).astype(F)
    s = np.sqrt(len_sq.astype(np.float64)).astype(F)
    return (d.astype(np.float64) * (1.0 / s.astype(np.float64))[..., None]).astype(F)


def numpy_build_f32(boxes, scale, nd):
    """transform_boxes + reduce + get_mcodes + stable sort in float32
    (build_radix_tree.hpp:85-260 with FloatType = float)."""
    mn, mx = boxes[:, :nd].astype(F), boxes[:, nd:].astype(F)
    s = F(np.float64(F(scale)) * 0.5)
    mid = (0.5 * (mn + mx).astype(np.float64)).astype(F)
    r = ((mx - mn).astype(np.float64) * np.float64(s)).astype(F)
    smin, smax = mid - r, mid + r
    lo, hi = smin.min(0), smax.max(0)
    ext = hi - lo
    inv = np.where(np.abs(ext) <= F(1e-8), F(0), F(1) / np.where(ext == 0, F(1), ext)).astype(F)
    c = (0.5 * (smin + smax).astype(np.float64)).astype(F)
    u = ((c - lo) * inv).astype(F)
    bits = 10 if nd == 3 else 16
    q = np.minimum(np.maximum(u * F(1 << bits), F(0)), F((1 << bits) - 1)).astype(np.int64)
    code = np.zeros(len(boxes), np.int64)
    for b in range(bits):
        for d in range(nd):
            code |= ((q[:, d] >> b) & 1) << (nd * b + d)
    order = np.lexsort((np.arange(len(boxes)), code))
    return np.concatenate([lo, hi]), code, order, np.concatenate([smin, smax], 1)


def ray_hits_f32(o, d, bmin, bmax, eps):
    """intersect_ray with T = float (intersect_ray_impl.hpp:150-187,321-351),
    one ray against many boxes, in float32."""
    tmin = np.full(bmin.shape[0], np.finfo(F).tiny, F)
    tmax = np.full(bmin.shape[0], np.finfo(F).max, F)
    hit = np.ones(bmin.shape[0], bool)
    for k in range(bmin.shape[1]):
        n, x0 = F(d[k]), F(o[k])
        if np.abs(n) <= F(eps):
            hit &= ~((x0 < bmin[:, k]) | (x0 > bmax[:, k]))
            continue
        invn = F(1) / n
        t1 = (bmin[:, k] - x0) * invn
        t2 = (bmax[:, k] - x0) * invn
        lo_t, hi_t = np.minimum(t1, t2), np.maximum(t1, t2)
        new_min = np.where(lo_t < tmin, tmin, lo_t)
        new_max = np.where(hi_t < tmax, hi_t, tmax)
        ok = ~(new_min > new_max)
        tmin = np.where(hit, new_min, tmin)
        tmax = np.where(hit, new_max, tmax)
        hit &= ok
    return hit


def _grid(nd):
    g = GOLD["bvh_uniform_grid"]
    boxes = refmeshes.uniform_grid_cell_boxes(g["mesh"]["lo"], g["mesh"]["hi"],
                                              g["mesh"]["nodes"][0])
    return g, boxes.astype(F), orc.OracleBVH(boxes.astype(F), scale=g["scale"], dtype=F)


def test_grid_known_answers_float():
    """spin_bvh.cpp:279-404 (boxes), :519-648 (rays), :792-888 (points) with
    FloatType = float."""
    g, boxes, bvh = _grid(3)
    b = bvh.export()["bounds"]
    np.testing.assert_allclose(b[:3], g["bounds"]["lo"], atol=g["bounds"]["tol"])
    np.testing.assert_allclose(b[3:], g["bounds"]["hi"], atol=g["bounds"]["tol"])
    q = np.array([qb["min"] + qb["max"] for qb in g["query_boxes"]], F)
    off, cnt, cand = bvh.find("boxes", q)
    for i, qb in enumerate(g["query_boxes"]):
        assert sorted(cand[off[i]:off[i] + cnt[i]].tolist()) == qb["expected"]
    rays = np.array([list(r["origin"]) + list(unit_vector_f32(r["direction"]))
                     for r in g["query_rays"]], F)
    off, cnt, cand = bvh.find("rays", rays)
    for i, r in enumerate(g["query_rays"]):
        assert sorted(cand[off[i]:off[i] + cnt[i]].tolist()) == r["expected"]
    cent = (0.5 * (boxes[:, :3] + boxes[:, 3:]).astype(np.float64)).astype(F)
    off, cnt, cand = bvh.find("points", cent)
    assert np.all(cnt == 1) and np.array_equal(cand[off], np.arange(boxes.shape[0]))


def test_single_and_empty_float():
    """spin_bvh.cpp:1009-1160, :1401-1556 with float: 0 or 1 box pads to 2
    leaves with BoundingBox<float>'s invalid box."""
    one = np.array([[0, 0, 0, 1, 1, 1]], F)
    bvh = orc.OracleBVH(one, scale=1.0, dtype=F)
    assert bvh.n == 2
    ex = bvh.export()
    pos = int(np.nonzero(ex["leafs"] == 1)[0][0])  # the padding box is box 1
    pad = ex["nodes"][ex["children"] == -(pos + 1)]
    assert pad.shape[0] == 1 and np.all(pad[0, :3] == np.finfo(F).max)
    assert np.all(pad[0, 3:] == -np.finfo(F).max)
    off, cnt, cand = bvh.find("points", np.array([[0.5, 0.5, 0.5], [2, 2, 2]], F))
    assert cnt.tolist() == [1, 0] and cand.tolist() == [0]
    empty = orc.OracleBVH(np.zeros((0, 6), F), scale=1.0, dtype=F)
    assert empty.find("points", np.array([[0.5, 0.5, 0.5]], F))[1].tolist() == [0]


def _skewed_boxes(n, seed, nd, hmax=2e-2):
    """Boxes whose float and double scaling and normalisation round apart:
    coordinates with many significant bits, far from the origin."""
    rng = np.random.default_rng(seed)
    c = 1000.0 + rng.random((n, nd)) * 3.0
    h = rng.uniform(1e-4, hmax, (n, nd))
    return np.concatenate([c - h, c + h], 1).astype(F)


@pytest.mark.parametrize("nd", [3, 2])
def test_build_matches_numpy_float32(nd):
    boxes = _skewed_boxes(20_000, 5 + nd, nd)
    bvh = orc.OracleBVH(boxes, scale=1.000123, ndims=nd, dtype=F)
    ex = bvh.export()
    bounds, code, order, scaled = numpy_build_f32(boxes, 1.000123, nd)
    assert np.array_equal(ex["bounds"], bounds.astype(np.float64))
    assert np.array_equal(ex["mcodes"].astype(np.int64), code[order])
    assert np.array_equal(ex["leafs"], order)
    # the same boxes as a double tree are scaled and ordered differently
    dbl = orc.OracleBVH(boxes.astype(np.float64), scale=1.000123, ndims=nd).export()
    assert not np.array_equal(dbl["bounds"], ex["bounds"])
    assert not np.array_equal(dbl["mcodes"], ex["mcodes"])


@pytest.mark.parametrize("kind", ["points", "boxes", "rays"])
def test_candidates_match_float32_brute_force(kind):
    nd = 3
    boxes = _skewed_boxes(3000, 11, nd, hmax=0.2)
    bvh = orc.OracleBVH(boxes, scale=1.000123, dtype=F)
    _, _, _, scaled = numpy_build_f32(boxes, 1.000123, nd)
    smin, smax = scaled[:, :nd], scaled[:, nd:]
    rng = np.random.default_rng(12)
    eps = np.finfo(F).eps
    if kind == "points":
        q = (1000.0 + rng.random((300, 3)) * 3.0).astype(F)
    elif kind == "boxes":
        c = 1000.0 + rng.random((300, 3)) * 3.0
        h = rng.uniform(1e-3, 5e-2, (300, 3))
        q = np.concatenate([c - h, c + h], 1).astype(F)
    else:
        o = (1000.0 + rng.random((100, 3)) * 3.0).astype(F)
        d = unit_vector_f32(rng.normal(size=(100, 3)))
        d[:5, 2] = 0.0  # rays in a plane: the skipped axis
        q = np.concatenate([o, d], 1).astype(F)
    off, cnt, cand = bvh.find(kind, q)
    for i in range(q.shape[0]):
        if kind == "points":
            hit = np.all((q[i] >= smin) & (q[i] <= smax), 1)
        elif kind == "boxes":
            hit = np.all((q[i, 3:] >= smin) & (q[i, :3] <= smax), 1)
        else:
            hit = ray_hits_f32(q[i, :3], q[i, 3:], smin, smax, eps)
        assert sorted(cand[off[i]:off[i] + cnt[i]].tolist()) == np.nonzero(hit)[0].tolist()
    assert cnt.sum() > 0


def ray_hits_f64(o, d, bmin, bmax, eps):
    """The same slab test in double, to show where float rounding decides."""
    tmin = np.full(bmin.shape[0], np.finfo(np.float64).tiny)
    tmax = np.full(bmin.shape[0], np.finfo(np.float64).max)
    hit = np.ones(bmin.shape[0], bool)
    for k in range(bmin.shape[1]):
        n, x0 = float(d[k]), float(o[k])
        if abs(n) <= eps:
            hit &= ~((x0 < bmin[:, k]) | (x0 > bmax[:, k]))
            continue
        t1 = (bmin[:, k] - x0) * (1.0 / n)
        t2 = (bmax[:, k] - x0) * (1.0 / n)
        lo_t, hi_t = np.minimum(t1, t2), np.maximum(t1, t2)
        new_min = np.where(lo_t < tmin, tmin, lo_t)
        new_max = np.where(hi_t < tmax, hi_t, tmax)
        ok = ~(new_min > new_max)
        tmin = np.where(hit, new_min, tmin)
        tmax = np.where(hit, new_max, tmax)
        hit &= ok
    return hit


def corner_rays(smin, smax, n, seed):
    """Rays that touch a box at one corner only, in exact arithmetic: through
    (max x, min y, mid z) with direction (a, b, 0), a and b multiples of 2^-11,
    from an origin 1/8 back along it (exact in float for coordinates in
    [512, 1024)). Whether such a ray hits is decided by the rounding of
    1 / a and 1 / b, so float and double answers differ."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        b = rng.integers(0, smin.shape[0])
        a_, b_ = (rng.integers(1 << 10, 1 << 11, 2) * 2.0 ** -11)
        p = np.array([smax[b, 0], smin[b, 1], 0.5 * (smin[b, 2] + smax[b, 2])], F)
        o = (p - F(0.125) * np.array([a_, b_, 0.0], F)).astype(F)
        assert np.all(o.astype(np.float64) == p.astype(np.float64) -
                      0.125 * np.array([a_, b_, 0.0]))  # exact
        out.append(np.concatenate([o, np.array([a_, b_, 0.0], F)]))
    return np.array(out, F)


def test_corner_rays_follow_float_rounding():
    rng = np.random.default_rng(31)
    c = 600.0 + rng.random((400, 3)) * 3.0
    h = rng.uniform(1e-2, 5e-2, (400, 3))
    boxes = np.concatenate([c - h, c + h], 1).astype(F)
    bvh = orc.OracleBVH(boxes, scale=1.000123, dtype=F)
    _, _, _, scaled = numpy_build_f32(boxes, 1.000123, 3)
    smin, smax = scaled[:, :3], scaled[:, 3:]
    rays = corner_rays(smin, smax, 2000, 32)
    eps = np.finfo(F).eps
    off, cnt, cand = bvh.find("rays", rays)
    differs = 0
    for i in range(rays.shape[0]):
        want = ray_hits_f32(rays[i, :3], rays[i, 3:], smin, smax, eps)
        assert sorted(cand[off[i]:off[i] + cnt[i]].tolist()) == np.nonzero(want)[0].tolist()
        differs += int((want != ray_hits_f64(rays[i, :3], rays[i, 3:], smin.astype(np.float64),
                                             smax.astype(np.float64), eps)).sum())
    assert differs > 50  # the inputs do separate float from double
