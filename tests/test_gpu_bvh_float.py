"""GPU parity of the float tree (spin::BVH<NDIMS, MI355X_EXEC, float>) against
the oracle's float tree: tree arrays bit-exact (node boxes, child links, leaf
ids, Morton codes, bounds) and candidate lists bit-exact, order included, for
points, boxes and rays in 3-D and 2-D; the reference's known answers run with
FloatType = float (spin/tests/spin_bvh.cpp:1560-1669)."""
import ctypes
import json
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from axom_amd import _lib as L  # noqa: E402
from axom_amd import spin  # noqa: E402
from oracle import oracle as orc  # noqa: E402
from tests import refmeshes  # noqa: E402
from tests.test_oracle_float_bvh import corner_rays, numpy_build_f32  # noqa: E402

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_known_answers.json")))
DEV = "cuda"
F = np.float32


def _t(a):
    return torch.as_tensor(np.ascontiguousarray(a), device=DEV)


def _boxes(n, seed, nd, origin=1000.0, span=3.0, hmax=0.02):
    rng = np.random.default_rng(seed)
    c = origin + rng.random((n, nd)) * span
    h = rng.uniform(1e-4, hmax, (n, nd))
    return np.concatenate([c - h, c + h], 1).astype(F)


def _pair(boxes, nd, scale=1.000123):
    g = spin.BVH(NDIMS=nd, FloatType=torch.float32, scaleFactor=scale)
    g.initialize(_t(boxes) if boxes.shape[0] else torch.zeros((0, 2 * nd), dtype=torch.float32,
                                                              device=DEV))
    return g, orc.OracleBVH(boxes, scale, ndims=nd, dtype=F)


def _check_tree(g, o, nd):
    ex, oe = g.export(), o.export()
    lo, hi = g.getBounds()
    assert lo.dtype == F
    assert np.array_equal(np.concatenate([lo, hi]).astype(np.float64), oe["bounds"])
    assert np.array_equal(ex["mcodes"].cpu().numpy().view(np.uint32), oe["mcodes"])
    assert np.array_equal(ex["leafs"].cpu().numpy(), oe["leafs"])
    assert np.array_equal(ex["children"].cpu().numpy(), oe["children"])
    assert np.array_equal(ex["nodes"].cpu().numpy(), oe["nodes"])


@pytest.mark.parametrize("nd", [3, 2])
@pytest.mark.parametrize("n", [0, 1, 2, 3, 1000, 100_000, 262_145])
def test_float_build_bit_exact(nd, n):
    boxes = _boxes(n, n + nd, nd)
    g, o = _pair(boxes, nd)
    _check_tree(g, o, nd)
    assert L.lib().axsq_bvh_precision(g._h) == 32
    if n >= 1000:  # the float tree is not the double tree of the same boxes
        d = spin.BVH(_t(boxes.astype(np.float64)), NDIMS=nd)
        assert not np.array_equal(d.export()["mcodes"].cpu().numpy(),
                                  g.export()["mcodes"].cpu().numpy())


def _queries(kind, nq, seed, nd, boxes):
    rng = np.random.default_rng(seed)
    lo = boxes[:, :nd].min(0).astype(np.float64)
    span = boxes[:, nd:].max(0).astype(np.float64) - lo
    if kind == "points":
        return (lo + rng.random((nq, nd)) * span).astype(F)
    if kind == "boxes":
        c = lo + rng.random((nq, nd)) * span
        h = rng.uniform(1e-3, 3e-2, (nq, nd))
        return np.concatenate([c - h, c + h], 1).astype(F)
    o = (lo + rng.random((nq, nd)) * span).astype(F)
    d = rng.normal(size=(nq, nd)).astype(F)
    if nd == 3:
        d[: nq // 10, 2] = 0.0  # some rays in a plane: the skipped axis
    return o, d


def _find(g, kind, q, nd):
    n = q.shape[0] if kind != "rays" else q[0].shape[0]
    off = torch.empty(n, dtype=torch.int32, device=DEV)
    cnt = torch.empty_like(off)
    if kind == "points":
        cand = g.findPoints(off, cnt, n, _t(q))
        prims = q
    elif kind == "boxes":
        cand = g.findBoundingBoxes(off, cnt, n, _t(q))
        prims = q
    else:
        mk = spin.make_rays2d if nd == 2 else spin.make_rays
        r = mk(_t(q[0]), _t(q[1]), dtype=torch.float32)
        cand = g.findRays(off, cnt, n, r)
        prims = r.cpu().numpy()
    return off.cpu().numpy(), cnt.cpu().numpy(), cand.cpu().numpy(), prims


@pytest.mark.parametrize("nd", [3, 2])
@pytest.mark.parametrize("kind,nq", [("points", 200_000), ("boxes", 50_000), ("rays", 3_000)])
def test_float_candidates_bit_exact(nd, kind, nq):
    boxes = _boxes(30_000, 40 + nd, nd, hmax=0.05)
    g, o = _pair(boxes, nd)
    q = _queries(kind, nq, 50 + nd, nd, boxes)
    off, cnt, cand, prims = _find(g, kind, q, nd)
    ooff, ocnt, ocand = o.find(kind, prims)
    assert np.array_equal(off, ooff) and np.array_equal(cnt, ocnt)
    assert np.array_equal(cand, ocand)
    assert cand.size > 0


def test_float_corner_rays_bit_exact():
    """Rays that touch boxes at one corner, where float rounding decides the
    answer (and double would decide otherwise for many of them)."""
    rng = np.random.default_rng(31)
    c = 600.0 + rng.random((400, 3)) * 3.0
    h = rng.uniform(1e-2, 5e-2, (400, 3))
    boxes = np.concatenate([c - h, c + h], 1).astype(F)
    g, o = _pair(boxes, 3)
    _, _, _, scaled = numpy_build_f32(boxes, 1.000123, 3)
    rays = corner_rays(scaled[:, :3], scaled[:, 3:], 2000, 32)
    n = rays.shape[0]
    off = torch.empty(n, dtype=torch.int32, device=DEV)
    cnt = torch.empty_like(off)
    cand = g.findRays(off, cnt, n, _t(rays)).cpu().numpy()
    ooff, ocnt, ocand = o.find("rays", rays)
    assert np.array_equal(off.cpu().numpy(), ooff) and np.array_equal(cand, ocand)


def test_float_known_answers():
    """spin_bvh.cpp:279-404, :519-648, :792-888 with FloatType = float."""
    gd = GOLD["bvh_uniform_grid"]
    boxes = refmeshes.uniform_grid_cell_boxes(gd["mesh"]["lo"], gd["mesh"]["hi"],
                                              gd["mesh"]["nodes"][0]).astype(F)
    g = spin.BVH(_t(boxes), scaleFactor=gd["scale"], FloatType=torch.float32)
    lo, hi = g.getBounds()
    np.testing.assert_allclose(lo, gd["bounds"]["lo"], atol=gd["bounds"]["tol"])
    np.testing.assert_allclose(hi, gd["bounds"]["hi"], atol=gd["bounds"]["tol"])
    q = np.array([qb["min"] + qb["max"] for qb in gd["query_boxes"]], F)
    off, cnt, cand, _ = _find(g, "boxes", q, 3)
    for i, qb in enumerate(gd["query_boxes"]):
        assert sorted(cand[off[i]:off[i] + cnt[i]].tolist()) == qb["expected"]
    o = np.array([r["origin"] for r in gd["query_rays"]], F)
    d = np.array([r["direction"] for r in gd["query_rays"]], F)
    off, cnt, cand, _ = _find(g, "rays", (o, d), 3)
    for i, r in enumerate(gd["query_rays"]):
        assert sorted(cand[off[i]:off[i] + cnt[i]].tolist()) == r["expected"]
    cent = (0.5 * (boxes[:, :3] + boxes[:, 3:]).astype(np.float64)).astype(F)
    off, cnt, cand, _ = _find(g, "points", cent, 3)
    assert np.all(cnt == 1) and np.array_equal(cand[off], np.arange(boxes.shape[0]))


def test_float_rebuild_and_precision_checks():
    boxes = _boxes(5000, 61, 3)
    g, _ = _pair(boxes, 3)
    boxes2 = _boxes(5000, 62, 3)
    g.initialize(_t(boxes2))  # in place, same count
    _check_tree(g, orc.OracleBVH(boxes2, 1.000123, dtype=F), 3)
    boxes3 = _boxes(7000, 63, 3)
    g.initialize(_t(boxes3))  # another count
    _check_tree(g, orc.OracleBVH(boxes3, 1.000123, dtype=F), 3)
    # double queries on a float tree, and float ones on a double tree, are refused
    pts = torch.zeros((4, 3), dtype=torch.float64, device=DEV)
    off = torch.empty(4, dtype=torch.int32, device=DEV)
    cnt = torch.empty_like(off)
    tot = ctypes.c_int64()
    base = pts.data_ptr()
    rc = L.lib().axsq_bvh_find_count(g._h, L.QUERY_POINTS, ctypes.c_void_p(base),
                                     ctypes.c_void_p(base + 8), ctypes.c_void_p(base + 16), 3, 4,
                                     1e-7, spin._ptr(off), spin._ptr(cnt), ctypes.byref(tot), None)
    assert rc == L.ERRORS_BY_NAME["AXSQ_ERR_STATE"] and b"float" in L.lib().axsq_last_error()
    d = spin.BVH(_t(boxes.astype(np.float64)))
    assert L.lib().axsq_bvh_precision(d._h) == 64
    p32 = torch.zeros((4, 3), dtype=torch.float32, device=DEV)
    base = p32.data_ptr()
    rc = L.lib().axsq_bvh_find_count_f32(d._h, L.QUERY_POINTS, ctypes.c_void_p(base),
                                         ctypes.c_void_p(base + 4), ctypes.c_void_p(base + 8), 3,
                                         4, 1e-7, spin._ptr(off), spin._ptr(cnt),
                                         ctypes.byref(tot), None)
    assert rc == L.ERRORS_BY_NAME["AXSQ_ERR_STATE"]
    with pytest.raises(L.AxsqError):
        L.check(L.lib().axsq_bvh_rebuild(g._h, spin._ptr(_t(boxes.astype(np.float64))), 5000,
                                         1.000123, 0, None))
