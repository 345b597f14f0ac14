"""The oracle's 2-D BVH (spin::BVH<2>): the reference's 2-D known answers
(spin/tests/spin_bvh.cpp:200-233 build, :409-517 findBoundingBoxes, :651-790
findRays, :906-1006 findPoints, :1010-1088 single box; spin/tests/
spin_morton.cpp:83-107 2-D Morton codes), an independent NumPy statement of
the 2-D build (16 bits per axis), tree invariants and brute force."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.test_oracle_properties import _check_subtree_ranges

EPS = float(np.finfo(np.float64).eps)


def grid_cells_2d():
    """mint::UniformMesh((0,0), (3,3), 4, 4): 3 x 3 unit cells, i fastest."""
    return np.array([[i, j, i + 1.0, j + 1.0] for j in range(3) for i in range(3)], float)


def ray2d(o, d):
    d = np.asarray(d, float)
    inv = 1.0 / np.sqrt((0.0 + d[0] * d[0]) + d[1] * d[1])
    return np.array([o[0], o[1], d[0] * inv, d[1] * inv])


def test_morton2_known_answers():
    assert orc.morton2(2, 2) == 0xC      # spin_morton.cpp:92-95
    assert orc.morton2(6, 3) == 30       # spin_morton.cpp:100-107
    rng = np.random.default_rng(1)
    for x, y in rng.integers(0, 65536, (500, 2)):
        want = 0
        for b in range(16):
            want |= ((int(x) >> b) & 1) << (2 * b) | ((int(y) >> b) & 1) << (2 * b + 1)
        assert orc.morton2(x, y) == want


def test_build2d_bounds():
    b = orc.OracleBVH(np.array([[0.0, 0.0, 1.0, 1.0], [1.0, 1.0, 2.0, 2.0]]), 1.0, ndims=2)
    assert np.array_equal(b.export()["bounds"], [0.0, 0.0, 2.0, 2.0])


def test_find_bounding_boxes_2d_known_answer():
    bvh = orc.OracleBVH(grid_cells_2d(), 1.0, ndims=2)
    assert np.array_equal(bvh.export()["bounds"], [0.0, 0.0, 3.0, 3.0])
    q = np.array([[-1.0, -1.0, 2.5, 1.5], [-1.0, -1.0, -0.1, -0.1]])
    off, cnt, cand = bvh.find("boxes", q)
    assert cnt.tolist() == [6, 0]
    assert sorted(cand.tolist()) == [0, 1, 2, 3, 4, 5]


def test_find_rays_2d_known_answer():
    bvh = orc.OracleBVH(grid_cells_2d(), 1.0, ndims=2)
    rays = np.stack([ray2d((-1.0, -1.0), (1.0, 1.0)), ray2d((-1.0, -1.0), (-1.0, -1.0))])
    off, cnt, cand = bvh.find("rays", rays, tol=EPS)
    assert cnt.tolist() == [7, 0]
    assert sorted(cand.tolist()) == [0, 1, 3, 4, 5, 7, 8]


def test_find_points_2d_known_answer():
    cells = grid_cells_2d()
    bvh = orc.OracleBVH(cells, 1.0, ndims=2)
    c = 0.5 * (cells[:, :2] + cells[:, 2:])
    off, cnt, cand = bvh.find("points", c)
    assert cnt.tolist() == [1] * 9 and cand.tolist() == list(range(9))
    off, cnt, cand = bvh.find("points", c + 10.0)
    assert cnt.tolist() == [0] * 9


def test_single_box_2d():
    bvh = orc.OracleBVH(np.array([[0.0, 0.0, 1.0, 1.0]]), 1.0, ndims=2)
    assert bvh.n == 2  # padded with an invalid box
    assert np.array_equal(bvh.export()["bounds"], [0.0, 0.0, 1.0, 1.0])
    off, cnt, cand = bvh.find("points", np.array([[0.5, 0.5]]))
    assert cnt.tolist() == [1] and cand.tolist() == [0]
    off, cnt, cand = bvh.find("points", np.array([[10.5, 10.5]]))
    assert cnt.tolist() == [0]


def _boxes2d(n, seed):
    rng = np.random.default_rng(seed)
    c = rng.random((n, 2))
    h = rng.uniform(0.001, 0.05, (n, 2))
    return np.concatenate([c - h, c + h], 1)


def numpy_build2d(boxes, scale):
    """transform + reduce + get_mcodes + stable sort for NDIMS = 2, written
    independently: scale about the midpoint, bounds, codes with 16 bits per
    axis (x even bits, y odd bits), lexicographic (code, id) order."""
    mn, mx = boxes[:, :2], boxes[:, 2:]
    mid = 0.5 * (mn + mx)
    r = (mx - mn) * (scale * 0.5)
    smin, smax = mid - r, mid + r
    lo, hi = smin.min(0), smax.max(0)
    ext = hi - lo
    inv = np.where(np.abs(ext) <= 1e-8, 0.0, 1.0 / np.where(ext == 0, 1, ext))
    u = (0.5 * (smin + smax) - lo) * inv
    q = np.minimum(np.maximum(u * 65536.0, 0.0), 65535.0).astype(np.int64)
    code = np.zeros(len(boxes), np.int64)
    for b in range(16):
        for d in range(2):
            code |= ((q[:, d] >> b) & 1) << (2 * b + d)
    order = np.lexsort((np.arange(len(boxes)), code))
    return np.concatenate([lo, hi]), code, order, np.concatenate([smin, smax], 1)


@pytest.mark.parametrize("n,scale", [(2, 1.0), (3, 1.000123), (999, 1.000123), (20000, 1.0)])
def test_build2d_matches_independent_statement_and_invariants(n, scale):
    boxes = _boxes2d(n, n)
    ex = orc.OracleBVH(boxes, scale, ndims=2).export()
    bounds, code, order, scaled = numpy_build2d(boxes, scale)
    assert np.array_equal(ex["bounds"], bounds)
    assert np.array_equal(ex["leafs"], order)
    assert np.array_equal(ex["mcodes"].astype(np.int64), code[order])
    # every inner box is the exact union of its children; leaves are the scaled boxes
    nodes, ch = ex["nodes"], ex["children"]
    assert sorted(-c - 1 for c in ch if c < 0) == list(range(n))
    for k in range(2 * (n - 1)):
        c = ch[k]
        if c < 0:
            assert np.array_equal(nodes[k], scaled[ex["leafs"][-c - 1]])
        else:
            l, r = nodes[c], nodes[c + 1]
            assert np.array_equal(nodes[k], np.concatenate([np.minimum(l[:2], r[:2]),
                                                            np.maximum(l[2:], r[2:])]))
    _check_subtree_ranges(ch, n - 1)


@pytest.mark.parametrize("kind", ["points", "boxes", "rays"])
def test_find2d_equals_brute_force(kind):
    boxes = _boxes2d(3000, 7)
    bvh = orc.OracleBVH(boxes, ndims=2)
    rng = np.random.default_rng(8)
    if kind == "points":
        q = rng.random((2000, 2))
    elif kind == "boxes":
        c = rng.random((2000, 2))
        h = rng.uniform(0.0, 0.03, (2000, 2))
        q = np.concatenate([c - h, c + h], 1)
    else:
        q = np.stack([ray2d(o, d) for o, d in zip(rng.uniform(-0.2, 1.2, (300, 2)),
                                                  rng.normal(size=(300, 2)))])
    off, cnt, cand = bvh.find(kind, q, tol=EPS)
    bcnt, bcand = bvh.brute(kind, q, tol=EPS)
    assert np.array_equal(cnt, bcnt)
    o2 = np.concatenate([[0], np.cumsum(bcnt)])
    for i in range(len(cnt)):
        assert sorted(cand[off[i]:off[i] + cnt[i]]) == list(bcand[o2[i]:o2[i + 1]])
    assert cnt.sum() > 100
