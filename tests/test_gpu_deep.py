"""Walks deeper than the traversal keeps in LDS.

Each lane's first 32 stack entries live in LDS and deeper ones in a per-lane
array. These trees make walks push 37 to 40 entries: Morton codes that differ
in successively lower bits give a chain-shaped Karras tree (each node's far
child is a small subtree, pushed while the walk goes down the chain), ending
in a cluster of identical codes, split by index into a balanced subtree. The
oracle's walk reports the deepest entry it pushed, and the test asserts it is
beyond 32 before comparing the GPU's results with the oracle's bit for bit:
findBoundingBoxes / findPoints / findRays (FP64 and FP32 traversal) and
SignedDistance (one pass, and every walk through the second pass).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from axom_amd import quest, spin  # noqa: E402
from oracle import oracle as orc  # noqa: E402

pytestmark = pytest.mark.gpu

DEV = "cuda"
LDS_ENTRIES = 32


def _t(a):
    return torch.as_tensor(np.ascontiguousarray(a), device=DEV)


def chain_centres(dup):
    """Centres of quantised cell q: dup copies at cell 0, two at each single
    bit of the 30-bit code (axis k % 3, bit k // 3), one at the far corner."""
    def cell(q):
        return [(q[0] + 0.5) / 1024, (q[1] + 0.5) / 1024, (q[2] + 0.5) / 1024]
    cents = [cell((0, 0, 0))] * dup
    for k in range(30):
        q = [0, 0, 0]
        q[k % 3] = 1 << (k // 3)
        cents += [cell(q)] * 2
    cents.append(cell((1023, 1023, 1023)))
    return np.array(cents)


def stretched_boxes(dup, corner, h=1e-6):
    """Boxes from each centre to a common corner point, so that a point query
    at the corner, a box around everything and a ray through the corner all
    pass every box."""
    c = chain_centres(dup)
    return np.concatenate([np.minimum(c, corner) - h, np.maximum(c, corner) + h], 1)


def chain_triangles(dup, e=1e-6):
    """A small triangle per centre whose box is centred there (so the codes
    are the centres'), the copies at one centre differing in shape."""
    c = chain_centres(dup)
    n = c.shape[0]
    k = np.arange(n)
    z = e * ((k % dup) + 1) / dup
    x = -e + 2 * e * ((k * 7919) % 1000) / 1000.0
    one = np.ones(n)
    a = c + np.stack([-e * one, -e * one, -z], 1)
    b = c + np.stack([e * one, -e * one, z], 1)
    cc = c + np.stack([x, e * one, 0 * z], 1)
    v = np.stack([a, b, cc], 1).reshape(-1, 3)
    return v, np.arange(3 * n, dtype=np.int32).reshape(-1, 3)


def _find(g, kind, prims):
    n = prims.shape[0]
    off = torch.empty(n, dtype=torch.int32, device=DEV)
    cnt = torch.empty_like(off)
    fn = {"points": g.findPoints, "boxes": g.findBoundingBoxes, "rays": g.findRays}[kind]
    cand = fn(off, cnt, n, _t(prims))
    return off.cpu().numpy(), cnt.cpu().numpy(), cand.cpu().numpy()


@pytest.mark.parametrize("fp_mode", ["fp64", "fp32"])
@pytest.mark.parametrize("dup", [256, 1024])
def test_deep_find(dup, fp_mode):
    corner = np.full(3, 0.25 / 1024)
    boxes = stretched_boxes(dup, corner)
    o = orc.OracleBVH(boxes)
    g = spin.BVH(_t(boxes), fp_mode=fp_mode)
    rng = np.random.default_rng(dup)
    pts = np.concatenate([np.tile(corner, (70, 1)), rng.random((2000, 3)) * 0.02])
    qb = np.concatenate([np.array([[-1.0, -1.0, -1.0, 2.0, 2.0, 2.0]] * 70),
                         np.concatenate([pts[70:] - 0.001, pts[70:] + 0.001], 1)])
    d = orc.unit_vector([1.0, 1.0, 1.0])
    rays = np.array([[-1.0, -1.0, -1.0, d[0], d[1], d[2]]] * 70 +
                    [list(p - 0.5) + list(orc.unit_vector(p - corner + 1e-3)) for p in pts[70:300]])
    for kind, prims in (("points", pts), ("boxes", qb), ("rays", rays)):
        depth = o.max_depth(kind, prims)
        assert depth > LDS_ENTRIES, (kind, depth)
        ooff, ocnt, ocand = o.find(kind, prims)
        goff, gcnt, gcand = _find(g, kind, prims)
        assert np.array_equal(gcnt, ocnt), kind
        assert np.array_equal(goff, ooff), kind
        assert np.array_equal(gcand, ocand), kind


def _sdf_queries(dup):
    rng = np.random.default_rng(7 + dup)
    near = np.array([[0.1, 0.1, 0.1], [0.3, 0.2, 0.6], [-10.0, -10.0, -10.0]]) / 1024
    return np.concatenate([np.repeat(near, 40, axis=0), rng.random((3000, 3)) * 0.01,
                           rng.random((1000, 3))])


@pytest.mark.parametrize("budget", [0, 1])
@pytest.mark.parametrize("dup", [256, 1024])
def test_deep_signed_distance(monkeypatch, dup, budget):
    """Budget 0: every walk in the first pass; budget 1: every walk suspended
    after one step and finished by the second pass, whose records carry the
    whole stack."""
    v, t = chain_triangles(dup)
    q = _sdf_queries(dup)
    o = orc.OracleSDF(v, t, watertight=False)
    assert o.max_depth(q[:120]) > LDS_ENTRIES
    want = o.eval(q, nthreads=8, want_stats=True)
    monkeypatch.setenv("AXSQ_SDF_BUDGET", str(budget))
    sdf = quest.SignedDistance(_t(v), _t(t), isWatertight=False)
    n = q.shape[0]
    cp = torch.empty((n, 3), dtype=torch.float64, device=DEV)
    elem = torch.empty(n, dtype=torch.int32, device=DEV)
    sdf.enableStats(True)
    d = sdf.computeDistances(n, _t(q), outClosestPts=cp, outElem=elem).cpu().numpy()
    visits, leaves = sdf.lastStats()
    if budget == 1:
        assert sdf.lastSuspended() > n // 2
    assert np.array_equal(d, want["dist"])
    assert np.array_equal(cp.cpu().numpy(), want["cp"])
    assert np.array_equal(elem.cpu().numpy(), want["elem"])
    assert visits == int(want["visits"].sum()) and leaves == int(want["leaves"].sum())


def test_deep_signed_distance_fp32_traversal(monkeypatch):
    """The FP32-traversal form on the deep tree, one pass and two: the same
    distances as the FP64 form (and so the oracle's)."""
    v, t = chain_triangles(1024)
    q = _sdf_queries(1024)
    want = orc.OracleSDF(v, t, watertight=False).eval(q, nthreads=8)
    for budget in (0, 1):
        monkeypatch.setenv("AXSQ_SDF_BUDGET", str(budget))
        sdf = quest.SignedDistance(_t(v), _t(t), isWatertight=False, fp_mode="fp32")
        d = sdf.computeDistances(q.shape[0], _t(q), fp_mode="fp32").cpu().numpy()
        assert np.array_equal(d, want["dist"]), budget
