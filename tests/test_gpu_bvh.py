"""GPU parity of the BVH build and candidate search against the oracle.

Bar: bit-exact tree arrays (inner boxes, child links, leaf ids, Morton codes,
bounds) and bit-exact candidate lists (offsets, counts, ids in traversal
order), for the FP64 form and the FP32-traversal form.
"""
import ctypes
import json
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from axom_amd import spin, synth  # noqa: E402
from oracle import oracle as orc  # noqa: E402
from tests import refmeshes  # noqa: E402

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_known_answers.json")))
DEV = "cuda"


def _boxes(n, seed, lo=0.001, hi=0.05):
    rng = np.random.default_rng(seed)
    c = rng.random((n, 3))
    h = rng.uniform(lo, hi, (n, 3))
    return np.concatenate([c - h, c + h], 1)


def _t(a):
    return torch.as_tensor(np.ascontiguousarray(a), device=DEV)


def _build_pair(boxes, scale, fp_mode="fp64"):
    g = spin.BVH(scaleFactor=scale, fp_mode=fp_mode)
    g.initialize(_t(boxes) if boxes.shape[0] else torch.zeros((0, 6), dtype=torch.float64,
                                                                 device=DEV))
    return g, orc.OracleBVH(boxes, scale)


# sizes either side of the build's internal limits: one refit span (16,384
# leaves), 1,024-key sort tiles (262,144 keys), a sort that scans its own
# histogram (at most 256 tiles of 4,096 keys) and one that does not
@pytest.mark.parametrize("n,scale", [(0, 1.0), (1, 1.0), (2, 1.000123), (3, 1.0), (257, 1.000123),
                                     (5000, 1.000123), (16_384, 1.0), (16_385, 1.000123),
                                     (100_000, 1.000123), (262_144, 1.000123), (262_145, 1.0),
                                     (1_048_577, 1.000123)])
def test_build_bit_exact(n, scale):
    boxes = _boxes(n, n + 1) if n < 100_000 else synth.workload1()[0]
    g, o = _build_pair(boxes, scale)
    ex = g.export()
    oe = o.export()
    lo, hi = g.getBounds()
    assert np.array_equal(np.concatenate([lo, hi]), oe["bounds"])
    assert np.array_equal(ex["mcodes"].cpu().numpy().view(np.uint32), oe["mcodes"])
    assert np.array_equal(ex["leafs"].cpu().numpy(), oe["leafs"])
    assert np.array_equal(ex["children"].cpu().numpy(), oe["children"])
    assert np.array_equal(ex["nodes"].cpu().numpy().view(np.int64), oe["nodes"].view(np.int64))


@pytest.mark.parametrize("n", [3000, 300_000])
def test_build_signed_zero_boxes_bit_exact(n):
    """Boxes flat at z = 0, some written with -0.0 (after scaling their z
    range is [-0.0, +0.0], where box unions depend on how leaves are grouped),
    others with +0.0, and invalid boxes: node boxes equal the oracle's SEQ
    refit bit for bit, signs of zeros included."""
    rng = np.random.default_rng(n)
    boxes = _boxes(n, n + 7)
    flat = rng.random(n) < 0.3
    boxes[flat, 2] = boxes[flat, 5] = np.where(rng.random(flat.sum()) < 0.5, -0.0, 0.0)
    boxes[rng.random(n) < 0.01] = [0.5, 0.5, 0.5, 0.4, 0.4, 0.4]
    for scale in (1.000123, 1.0):
        g, o = _build_pair(boxes, scale)
        ex, oe = g.export(), o.export()
        assert np.array_equal(ex["children"].cpu().numpy(), oe["children"])
        assert np.array_equal(ex["nodes"].cpu().numpy().view(np.int64),
                              oe["nodes"].view(np.int64))
    plus = boxes.copy()  # only +0.0: no signed-zero pair
    plus[flat, 2] = plus[flat, 5] = 0.0
    g, o = _build_pair(plus, 1.0)
    assert np.array_equal(g.export()["nodes"].cpu().numpy().view(np.int64),
                          o.export()["nodes"].view(np.int64))


@pytest.mark.parametrize("n", [700, 100_000])
def test_rebuild_replays_match_oracle(n):
    """Rebuilds of the same input buffer (the small-build path replays one HIP
    graph of its launches while buffers and sizes hold) with the contents
    changed in place between builds, searches and host syncs in between:
    every rebuild equals the oracle's tree of the current contents."""
    g = spin.BVH(scaleFactor=1.000123)
    bx = _t(_boxes(n, 100))
    p = np.random.default_rng(7).random((3000, 3))
    off = torch.empty(3000, dtype=torch.int32, device=DEV)
    cnt = torch.empty_like(off)
    for rep in range(5):
        boxes = _boxes(n, 100 + rep)
        bx.copy_(_t(boxes))
        g.initialize(bx)
        cand = g.findPoints(off, cnt, 3000, _t(p)).cpu().numpy()
        ex, oe = g.export(), orc.OracleBVH(boxes, 1.000123).export()
        assert np.array_equal(ex["nodes"].cpu().numpy().view(np.int64), oe["nodes"].view(np.int64))
        assert np.array_equal(ex["children"].cpu().numpy(), oe["children"])
        lo, hi = g.getBounds()
        assert np.array_equal(np.concatenate([lo, hi]), oe["bounds"])
        _, ocnt, ocand = orc.OracleBVH(boxes, 1.000123).find("points", p)
        assert np.array_equal(cnt.cpu().numpy(), ocnt) and np.array_equal(cand, ocand)


def test_build_duplicate_codes_and_invalid_boxes():
    boxes = np.concatenate([np.tile([[0.2, 0.2, 0.2, 0.3, 0.3, 0.3]], (300, 1)), _boxes(700, 9)])
    boxes[::97] = [np.finfo(np.float64).max] * 3 + [-np.finfo(np.float64).max] * 3
    g, o = _build_pair(boxes, 1.0)
    ex, oe = g.export(), o.export()
    assert np.array_equal(ex["children"].cpu().numpy(), oe["children"])
    assert np.array_equal(ex["nodes"].cpu().numpy().view(np.int64), oe["nodes"].view(np.int64))
    assert np.array_equal(ex["leafs"].cpu().numpy(), oe["leafs"])


def test_initialize_again_rebuilds_in_place():
    """BVH::initialize on an initialized tree replaces it (BVH.hpp:426-477): the
    same count (device arrays kept), another count, the FP32 nodes added, and
    an empty input, each bit-exact against a fresh oracle build."""
    g = spin.BVH(scaleFactor=1.000123)
    for n, seed, mode in ((20_000, 1, "fp64"), (20_000, 2, "fp64"), (70_001, 3, "fp64"),
                          (5000, 4, "fp32"), (0, 5, "fp64"), (300, 6, "fp64")):
        g._fp = spin.FP_MODES[mode]  # the traversal precision the next build emits
        boxes = _boxes(n, seed)
        g.initialize(_t(boxes) if n else torch.zeros((0, 6), dtype=torch.float64, device=DEV))
        ex, oe = g.export(), orc.OracleBVH(boxes.reshape(-1, 6), 1.000123).export()
        for k in ("children", "nodes", "leafs"):
            assert np.array_equal(ex[k].cpu().numpy(), oe[k]), (n, k)
        lo, hi = g.getBounds()
        assert np.array_equal(np.concatenate([lo, hi]), oe["bounds"]), n
        if n:
            p = np.random.default_rng(seed).random((2000, 3))
            off = torch.empty(2000, dtype=torch.int32, device=DEV)
            cnt = torch.empty_like(off)
            cand = g.findPoints(off, cnt, 2000, _t(p)).cpu().numpy()
            ooff, ocnt, ocand = orc.OracleBVH(boxes, 1.000123).find("points", p)
            assert np.array_equal(cnt.cpu().numpy(), ocnt) and np.array_equal(cand, ocand)


def _abi_count(g, t):
    from axom_amd import _lib as L
    n = t.shape[0]
    off = torch.empty(n, dtype=torch.int32, device=DEV)
    cnt = torch.empty_like(off)
    tot = ctypes.c_int64()
    base = t.data_ptr()
    L.check(L.lib().axsq_bvh_find_count(g._h, L.QUERY_POINTS, ctypes.c_void_p(base),
                                        ctypes.c_void_p(base + 8), ctypes.c_void_p(base + 16),
                                        3, n, g._tol, spin._ptr(off), spin._ptr(cnt),
                                        ctypes.byref(tot), None))
    return off, cnt, tot.value


def _abi_fill(g, t, off, cnt, total, guard=4096):
    """Fills into a buffer with `guard` words after the total, all -7 first;
    returns (rc, ids, the guard words)."""
    from axom_amd import _lib as L
    buf = torch.full((total + guard,), -7, dtype=torch.int32, device=DEV)
    base = t.data_ptr()
    rc = L.lib().axsq_bvh_find_fill(g._h, L.QUERY_POINTS, ctypes.c_void_p(base),
                                    ctypes.c_void_p(base + 8), ctypes.c_void_p(base + 16), 3,
                                    t.shape[0], g._tol, spin._ptr(off), spin._ptr(cnt),
                                    spin._ptr(buf), None)
    b = buf.cpu().numpy()
    return rc, b[:total], b[total:]


def test_two_phase_fill_is_bounded_by_the_counts():
    """Phase 2 searches again and writes at most counts[i] ids per query. With
    the queries unchanged it gives the oracle's lists; with the queries moved
    in place after phase 1 (so some now have more candidates) it returns
    AXSQ_ERR_INVALID_ARG, and no id lands past the buffer of phase 1's total:
    the guard words after it keep their value."""
    from axom_amd import _lib as L
    boxes = _boxes(30_000, 8, 0.02, 0.08)  # many candidates per point
    g = spin.BVH(_t(boxes))
    o = orc.OracleBVH(boxes, 1.000123)
    pa = np.random.default_rng(1).random((4000, 3))
    ta = _t(pa)
    ooff, ocnt, ocand = o.find("points", pa)
    off, cnt, tot = _abi_count(g, ta)
    assert tot == ocand.size and np.array_equal(cnt.cpu().numpy(), ocnt)
    rc, ids, guard = _abi_fill(g, ta, off, cnt, tot)
    assert rc == 0 and np.array_equal(ids, ocand) and np.all(guard == -7)
    # move the queries in place: the same buffer now holds other points
    ta.copy_(_t(np.random.default_rng(2).random((4000, 3)) * 0.2 + 0.4))  # a denser region
    rc, ids, guard = _abi_fill(g, ta, off, cnt, tot)
    assert rc == L.ERRORS_BY_NAME["AXSQ_ERR_INVALID_ARG"]
    assert "count" in L.last_error()
    assert np.all(guard == -7)
    # after the error the tree's status is clear: a fresh count and fill agree again
    off2, cnt2, tot2 = _abi_count(g, ta)
    rc, ids, guard = _abi_fill(g, ta, off2, cnt2, tot2)
    _, _, want = o.find("points", ta.cpu().numpy())
    assert rc == 0 and np.array_equal(ids, want) and np.all(guard == -7)


def test_find_into_capacity():
    """The one-call search into caller storage writes nothing when the room is
    too small (only the counts, offsets and total), and the oracle's lists when
    it is enough; rebuilding the tree in between changes nothing."""
    from axom_amd import _lib as L
    boxes = _boxes(30_000, 8, 0.02, 0.08)
    g = spin.BVH(_t(boxes))
    pa = np.random.default_rng(3).random((5000, 3))
    ta = _t(pa)
    ooff, ocnt, ocand = orc.OracleBVH(boxes, 1.000123).find("points", pa)
    n = pa.shape[0]
    base = ta.data_ptr()
    for cap in (0, ocand.size - 1, ocand.size, ocand.size + 100):
        off = torch.empty(n, dtype=torch.int32, device=DEV)
        cnt = torch.empty_like(off)
        buf = torch.full((max(cap, 1) + 64,), -7, dtype=torch.int32, device=DEV)
        tot = ctypes.c_int64()
        g.initialize(_t(boxes))
        L.check(L.lib().axsq_bvh_find_into(g._h, L.QUERY_POINTS, ctypes.c_void_p(base),
                                           ctypes.c_void_p(base + 8),
                                           ctypes.c_void_p(base + 16), 3, n, g._tol,
                                           spin._ptr(off), spin._ptr(cnt), spin._ptr(buf), cap,
                                           ctypes.byref(tot), None))
        b = buf.cpu().numpy()
        assert tot.value == ocand.size
        assert np.array_equal(off.cpu().numpy(), ooff) and np.array_equal(cnt.cpu().numpy(), ocnt)
        if cap < ocand.size:
            assert np.all(b == -7), cap
        else:
            assert np.array_equal(b[:ocand.size], ocand) and np.all(b[ocand.size:] == -7), cap


@pytest.mark.parametrize("allocator", [0, 1, 2])
def test_one_call_find_allocator(allocator):
    """axsq_bvh_find allocates the candidate array in the memory the tree's
    allocator id names (BVH::setAllocatorID): device, managed or pinned host;
    axsq_free releases each."""
    from axom_amd import _lib as L
    boxes = _boxes(5000, 4)
    g = spin.BVH(_t(boxes))
    L.check(L.lib().axsq_bvh_set_allocator(g._h, allocator))
    assert L.lib().axsq_bvh_get_allocator(g._h) == allocator
    pa = np.random.default_rng(5).random((3000, 3))
    ta = _t(pa)
    _, ocnt, ocand = orc.OracleBVH(boxes, 1.000123).find("points", pa)
    off = torch.empty(3000, dtype=torch.int32, device=DEV)
    cnt = torch.empty_like(off)
    ptr, tot = ctypes.c_void_p(), ctypes.c_int64()
    base = ta.data_ptr()
    L.check(L.lib().axsq_bvh_find(g._h, L.QUERY_POINTS, ctypes.c_void_p(base),
                                  ctypes.c_void_p(base + 8), ctypes.c_void_p(base + 16), 3, 3000,
                                  g._tol, spin._ptr(off), spin._ptr(cnt), ctypes.byref(ptr),
                                  ctypes.byref(tot), None))
    assert tot.value == ocand.size
    host = np.empty(tot.value, np.int32)
    torch.cuda.synchronize()
    if allocator == 0:
        view = torch.as_tensor(_DevI32(ptr.value, tot.value), device=DEV)
        host[:] = view.cpu().numpy()
    else:  # managed and pinned host memory are readable from the host
        ctypes.memmove(host.ctypes.data, ptr.value, 4 * tot.value)
    assert np.array_equal(host, ocand)
    L.check(L.lib().axsq_free(ptr))
    with pytest.raises(L.AxsqError):
        L.check(L.lib().axsq_bvh_set_allocator(g._h, 9))


class _DevI32:
    def __init__(self, ptr, n):
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": "<i4", "data": (ptr, False),
                                         "version": 2}


def test_build_is_deterministic():
    boxes = _boxes(50_000, 3)
    a = spin.BVH(_t(boxes)).export()
    b = spin.BVH(_t(boxes)).export()
    for k in a:
        assert torch.equal(a[k], b[k])


def _find(g, kind, prims, n):
    off = torch.empty(n, dtype=torch.int32, device=DEV)
    cnt = torch.empty(n, dtype=torch.int32, device=DEV)
    if kind == "points":
        cand = g.findPoints(off, cnt, n, _t(prims))
    elif kind == "boxes":
        cand = g.findBoundingBoxes(off, cnt, n, _t(prims))
    else:
        cand = g.findRays(off, cnt, n, _t(prims))
    return off.cpu().numpy(), cnt.cpu().numpy(), cand.cpu().numpy()


def _queries(kind, nq, seed):
    rng = np.random.default_rng(seed)
    if kind == "points":
        return rng.random((nq, 3))
    if kind == "boxes":
        c = rng.random((nq, 3))
        h = rng.uniform(0.0, 0.02, (nq, 3))
        return np.concatenate([c - h, c + h], 1)
    o = rng.uniform(-0.2, 1.2, (nq, 3))
    d = rng.normal(size=(nq, 3))
    d[::11, 1] = 0.0
    d = np.array([orc.unit_vector(x) for x in d])
    return np.concatenate([o, d], 1)


@pytest.mark.parametrize("fp_mode", ["fp64", "fp32"])
@pytest.mark.parametrize("kind,nq", [("points", 200_000), ("boxes", 50_000), ("rays", 3_000)])
def test_find_bit_exact(kind, nq, fp_mode):
    boxes = synth.workload1(scale=0.2)[0]
    g, o = _build_pair(boxes, 1.000123, fp_mode)
    prims = _queries(kind, nq, 11)
    goff, gcnt, gcand = _find(g, kind, prims, nq)
    ooff, ocnt, ocand = o.find(kind, prims, nthreads=8)
    assert np.array_equal(gcnt, ocnt)
    assert np.array_equal(goff, ooff)
    assert np.array_equal(gcand, ocand)


def test_find_points_full_workload():
    """findPoints at the full 100,000 boxes x 1,000,000 points workload."""
    boxes, pts = synth.workload1()
    g, o = _build_pair(boxes, 1.000123)
    goff, gcnt, gcand = _find(g, "points", pts, pts.shape[0])
    ooff, ocnt, ocand = o.find("points", pts, nthreads=8)
    assert np.array_equal(gcnt, ocnt) and np.array_equal(goff, ooff)
    assert np.array_equal(gcand, ocand)


def test_find_empty_and_single():
    g = spin.BVH(_t(np.array([[0.0, 0.0, 0.0, 1.0, 1.0, 1.0]])), scaleFactor=1.0)
    off, cnt, cand = _find(g, "points", np.array([[0.5, 0.5, 0.5], [3.0, 3.0, 3.0]]), 2)
    assert cnt.tolist() == [1, 0] and cand.tolist() == [0]
    off, cnt, cand = _find(g, "points", np.zeros((0, 3)), 0)
    assert cand.size == 0


def _grid():
    gd = GOLD["bvh_uniform_grid"]
    boxes = refmeshes.uniform_grid_cell_boxes(gd["mesh"]["lo"], gd["mesh"]["hi"],
                                              gd["mesh"]["nodes"][0])
    return gd, boxes, spin.BVH(_t(boxes), scaleFactor=gd["scale"])


def test_known_answer_grid_boxes_and_bounds():
    gd, boxes, g = _grid()
    lo, hi = g.getBounds()
    np.testing.assert_allclose(lo, gd["bounds"]["lo"], atol=gd["bounds"]["tol"])
    np.testing.assert_allclose(hi, gd["bounds"]["hi"], atol=gd["bounds"]["tol"])
    q = np.array([qb["min"] + qb["max"] for qb in gd["query_boxes"]])
    off, cnt, cand = _find(g, "boxes", q, q.shape[0])
    for i, qb in enumerate(gd["query_boxes"]):
        assert sorted(cand[off[i]:off[i] + cnt[i]].tolist()) == qb["expected"]


def test_known_answer_grid_rays():
    gd, boxes, g = _grid()
    o = _t(np.array([r["origin"] for r in gd["query_rays"]]))
    d = _t(np.array([r["direction"] for r in gd["query_rays"]]))
    rays = spin.make_rays(o, d)
    n = rays.shape[0]
    off = torch.empty(n, dtype=torch.int32, device=DEV)
    cnt = torch.empty(n, dtype=torch.int32, device=DEV)
    cand = g.findRays(off, cnt, n, rays).cpu().numpy()
    off, cnt = off.cpu().numpy(), cnt.cpu().numpy()
    for i, r in enumerate(gd["query_rays"]):
        assert sorted(cand[off[i]:off[i] + cnt[i]].tolist()) == r["expected"]


def test_known_answer_grid_points():
    gd, boxes, g = _grid()
    cent = 0.5 * (boxes[:, :3] + boxes[:, 3:])
    off, cnt, cand = _find(g, "points", cent, cent.shape[0])
    assert np.all(cnt == 1) and np.array_equal(cand[off], np.arange(boxes.shape[0]))
    off, cnt, cand = _find(g, "points", cent + 10.0, cent.shape[0])
    assert np.all(cnt == 0)


def test_soa_points_match_aos():
    boxes, pts = synth.workload1(scale=0.05)
    g = spin.BVH(_t(boxes))
    n = pts.shape[0]
    a = _find(g, "points", pts, n)
    off = torch.empty(n, dtype=torch.int32, device=DEV)
    cnt = torch.empty(n, dtype=torch.int32, device=DEV)
    cols = tuple(_t(pts[:, d]) for d in range(3))
    cand = g.findPoints(off, cnt, n, cols)
    assert np.array_equal(cand.cpu().numpy(), a[2])
    assert np.array_equal(off.cpu().numpy(), a[0])


def test_bad_arguments_raise():
    g = spin.BVH(_t(_boxes(10, 1)))
    off = torch.empty(3, dtype=torch.int32, device=DEV)
    cnt = torch.empty(4, dtype=torch.int32, device=DEV)
    with pytest.raises(ValueError):
        g.findPoints(off, cnt, 4, _t(np.zeros((4, 3))))


def test_inverted_leaf_box_is_never_reported():
    """An inverted leaf box (min > max on x) that the query box would overlap
    if its bounds were taken as they stand: the traversal tests a box only when
    it is valid, as the oracle does, so neither the FP64 nor the FP32 form
    reports it, for boxes or for rays."""
    boxes = _boxes(1000, 11)
    boxes[5] = [0.6, 0.5, 0.5, 0.4, 0.6, 0.6]
    queries = {"boxes": np.array([[0.3, 0.4, 0.4, 0.7, 0.7, 0.7]]),
               "rays": np.array([[0.0, 0.55, 0.55, 1.0, 0.0, 0.0]])}
    for kind, q in queries.items():
        ooff, ocnt, ocand = orc.OracleBVH(boxes, 1.000123).find(kind, q)
        assert 5 not in ocand
        for mode in ("fp64", "fp32"):
            _, cnt, cand = _find(spin.BVH(_t(boxes), fp_mode=mode), kind, q, 1)
            assert np.array_equal(cnt, ocnt) and np.array_equal(cand, ocand), (kind, mode)


@pytest.mark.parametrize("fp_mode", ["fp64", "fp32"])
def test_non_finite_queries(fp_mode):
    """NaN and infinite coordinates, next to a finite query of each kind: a
    NaN point coordinate passes contains() on its axis (an all-NaN point
    holds every box), a NaN box bound overlaps nothing, and rays follow the
    slab test; both traversal forms give the oracle's lists."""
    boxes = synth.workload1(scale=0.2)[0]
    g, o = _build_pair(boxes, 1.000123, fp_mode)
    nan, inf = np.nan, np.inf
    queries = {
        "points": np.array([[nan, 0.5, 0.5], [0.3, nan, 0.7], [0.2, 0.4, nan], [nan, nan, 0.5],
                            [nan, nan, nan], [inf, 0.5, 0.5], [0.5, -inf, 0.5],
                            [0.5, 0.5, 0.5]]),
        "boxes": np.array([[nan, 0.4, 0.4, 0.6, 0.6, 0.6], [0.4, 0.4, 0.4, nan, 0.6, 0.6],
                           [-inf, 0.4, 0.4, inf, 0.6, 0.6], [0.4, 0.4, 0.4, 0.6, 0.6, 0.6]]),
        "rays": np.array([[nan, 0.5, 0.5, 1, 0, 0], [0, 0.5, 0.5, nan, 0, 0],
                          [-inf, 0.5, 0.5, 1, 0, 0], [0, 0.5, 0.5, inf, 0, 0],
                          [0, 0.5, 0.5, 1, 0, 0]]),
    }
    for kind, q in queries.items():
        goff, gcnt, gcand = _find(g, kind, q, q.shape[0])
        ooff, ocnt, ocand = o.find(kind, q)
        assert np.array_equal(gcnt, ocnt), (kind, gcnt, ocnt)
        assert np.array_equal(goff, ooff) and np.array_equal(gcand, ocand), kind
    assert o.find("points", queries["points"][4:5])[1][0] == boxes.shape[0]


@pytest.mark.parametrize("kind", ["points", "boxes", "rays"])
def test_count_pass_visits_match_oracle(kind):
    """The count pass's visit counter (enableStats) equals the oracle walk's
    inner-node visits summed over the queries, for the FP64 tree."""
    boxes = _boxes(40_000, 21, 0.001, 0.03)
    g = spin.BVH(_t(boxes))
    g.enableStats(True)
    rng = np.random.default_rng(22)
    if kind == "points":
        q = rng.random((100_000, 3))
        prims = _t(q)
    elif kind == "boxes":
        c = rng.random((50_000, 3))
        h = rng.uniform(0.001, 0.02, (50_000, 3))
        q = np.concatenate([c - h, c + h], 1)
        prims = _t(q)
    else:
        q = np.concatenate([rng.random((5_000, 3)), rng.normal(size=(5_000, 3))], 1)
        prims = spin.make_rays(_t(q[:, :3]), _t(q[:, 3:]))
        q = prims.cpu().numpy()
    n = q.shape[0]
    off = torch.empty(n, dtype=torch.int32, device=DEV)
    cnt = torch.empty_like(off)
    find = {"points": g.findPoints, "boxes": g.findBoundingBoxes, "rays": g.findRays}[kind]
    find(off, cnt, n, prims)
    assert g.lastStats() == orc.OracleBVH(boxes, 1.000123).visits(kind, q)
