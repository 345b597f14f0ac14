"""GPU parity of the 2-D BVH (spin::BVH<2>) against the oracle.

Bar: bit-exact tree arrays (2-D inner boxes, child links, leaf ids, Morton
codes, bounds) and bit-exact candidate lists (offsets, counts, ids in
traversal order) for findPoints / findBoundingBoxes / findRays, FP64 and
FP32-traversal forms; the reference's 2-D known answers through the GPU
(spin/tests/spin_bvh.cpp:409-517, :651-790, :906-1006, :1010-1088).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from axom_amd import spin  # noqa: E402
from oracle import oracle as orc  # noqa: E402
from tests.test_oracle_bvh2d import _boxes2d, grid_cells_2d, ray2d  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _t(a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=DEV)


def _pair(boxes, scale, fp_mode="fp64"):
    g = spin.BVH(scaleFactor=scale, fp_mode=fp_mode, NDIMS=2)
    g.initialize(_t(boxes.reshape(-1, 4)))
    return g, orc.OracleBVH(boxes, scale, ndims=2)


@pytest.mark.parametrize("n,scale", [(0, 1.0), (1, 1.0), (2, 1.000123), (3, 1.0), (257, 1.000123),
                                     (5000, 1.000123), (200_000, 1.000123)])
def test_build2d_bit_exact(n, scale):
    boxes = _boxes2d(n, n + 1)
    g, o = _pair(boxes, scale)
    ex, oe = g.export(), o.export()
    lo, hi = g.getBounds()
    assert np.array_equal(np.concatenate([lo, hi]), oe["bounds"])
    assert np.array_equal(ex["mcodes"].cpu().numpy().view(np.uint32), oe["mcodes"])
    assert np.array_equal(ex["leafs"].cpu().numpy(), oe["leafs"])
    assert np.array_equal(ex["children"].cpu().numpy(), oe["children"])
    assert np.array_equal(ex["nodes"].cpu().numpy(), oe["nodes"])


def test_build2d_duplicate_codes_and_invalid_boxes():
    boxes = np.concatenate([np.tile([[0.2, 0.2, 0.3, 0.3]], (300, 1)), _boxes2d(700, 9)])
    boxes[::97] = [np.finfo(np.float64).max] * 2 + [-np.finfo(np.float64).max] * 2
    g, o = _pair(boxes, 1.0)
    ex, oe = g.export(), o.export()
    for k in ("children", "nodes", "leafs"):
        assert np.array_equal(ex[k].cpu().numpy(), oe[k]), k


def _find(g, kind, prims, n):
    off = torch.empty(n, dtype=torch.int32, device=DEV)
    cnt = torch.empty(n, dtype=torch.int32, device=DEV)
    f = {"points": g.findPoints, "boxes": g.findBoundingBoxes, "rays": g.findRays}[kind]
    cand = f(off, cnt, n, _t(prims))
    return off.cpu().numpy(), cnt.cpu().numpy(), cand.cpu().numpy()


def _queries(kind, nq, seed):
    rng = np.random.default_rng(seed)
    if kind == "points":
        return rng.random((nq, 2))
    if kind == "boxes":
        c = rng.random((nq, 2))
        h = rng.uniform(0.0, 0.02, (nq, 2))
        return np.concatenate([c - h, c + h], 1)
    o = rng.uniform(-0.2, 1.2, (nq, 2))
    d = rng.normal(size=(nq, 2))
    d[::11, 1] = 0.0
    return np.stack([ray2d(a, b) for a, b in zip(o, d)])


@pytest.mark.parametrize("fp_mode", ["fp64", "fp32"])
@pytest.mark.parametrize("kind,nq", [("points", 200_000), ("boxes", 50_000), ("rays", 3_000)])
def test_find2d_bit_exact(kind, nq, fp_mode):
    boxes = _boxes2d(20_000, 17)
    g, o = _pair(boxes, 1.000123, fp_mode)
    prims = _queries(kind, nq, 11)
    goff, gcnt, gcand = _find(g, kind, prims, nq)
    ooff, ocnt, ocand = o.find(kind, prims, nthreads=8)
    assert np.array_equal(gcnt, ocnt)
    assert np.array_equal(goff, ooff)
    assert np.array_equal(gcand, ocand)
    assert gcand.size > 100


def test_find2d_soa_points_match_aos():
    boxes = _boxes2d(5000, 3)
    g, _ = _pair(boxes, 1.000123)
    p = np.random.default_rng(2).random((10_000, 2))
    off = torch.empty(p.shape[0], dtype=torch.int32, device=DEV)
    cnt = torch.empty_like(off)
    a = g.findPoints(off, cnt, p.shape[0], _t(p)).cpu().numpy()
    b = g.findPoints(off, cnt, p.shape[0], (_t(p[:, 0]), _t(p[:, 1]))).cpu().numpy()
    assert np.array_equal(a, b)


def test_reference_2d_known_answers_on_gpu():
    cells = grid_cells_2d()
    g = spin.BVH(_t(cells), scaleFactor=1.0, NDIMS=2)
    lo, hi = g.getBounds()
    assert lo.tolist() == [0.0, 0.0] and hi.tolist() == [3.0, 3.0]
    off, cnt, cand = _find(g, "boxes", np.array([[-1.0, -1.0, 2.5, 1.5],
                                                 [-1.0, -1.0, -0.1, -0.1]]), 2)
    assert cnt.tolist() == [6, 0] and sorted(cand.tolist()) == [0, 1, 2, 3, 4, 5]
    rays = np.stack([ray2d((-1.0, -1.0), (1.0, 1.0)), ray2d((-1.0, -1.0), (-1.0, -1.0))])
    off, cnt, cand = _find(g, "rays", rays, 2)
    assert cnt.tolist() == [7, 0] and sorted(cand.tolist()) == [0, 1, 3, 4, 5, 7, 8]
    c = 0.5 * (cells[:, :2] + cells[:, 2:])
    off, cnt, cand = _find(g, "points", c, 9)
    assert cnt.tolist() == [1] * 9 and cand.tolist() == list(range(9))
    off, cnt, cand = _find(g, "points", c + 10.0, 9)
    assert cnt.tolist() == [0] * 9
    one = spin.BVH(_t(np.array([[0.0, 0.0, 1.0, 1.0]])), scaleFactor=1.0, NDIMS=2)
    off, cnt, cand = _find(one, "points", np.array([[0.5, 0.5], [10.5, 10.5]]), 2)
    assert cnt.tolist() == [1, 0] and cand.tolist() == [0]
