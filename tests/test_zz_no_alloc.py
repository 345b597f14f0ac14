"""axsq_bvh_find_into_noalloc: the one-call candidate search into caller
storage with its scratch in a fixed per-GPU array instead of allocated
memory. Its results must be those of axsq_bvh_find_into, bit for bit
(offsets, counts, total, ids in traversal order), on calls that follow each
other and reuse the array, the later ones smaller; sizes beyond its maximum
are refused before anything is written."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from axom_amd import _lib as L  # noqa: E402
from axom_amd import spin  # noqa: E402
from oracle import oracle as orc  # noqa: E402

pytestmark = pytest.mark.gpu

DEV = "cuda"
MAX_QUERIES = 1 << 20
GUARD = -7


def kept_ids(nq):
    """Ids per query the count pass keeps in the 32 MiB array (4,096 bytes
    of header, then 8 bytes per query of counts and list)."""
    return min(64, ((32 << 20) - 4096 - 8 * nq) // (4 * nq))


def _boxes(n, seed, lo, hi):
    rng = np.random.default_rng(seed)
    c = rng.random((n, 3))
    h = rng.uniform(lo, hi, (n, 3))
    return np.concatenate([c - h, c + h], 1)


def _queries(kind, nq, seed):
    rng = np.random.default_rng(seed)
    if kind == L.QUERY_POINTS:
        return rng.random((nq, 3))
    if kind == L.QUERY_BOXES:
        c = rng.random((nq, 3))
        h = rng.uniform(0.001, 0.03, (nq, 3))
        return np.concatenate([c - h, c + h], 1)
    o = rng.random((nq, 3))
    d = rng.normal(size=(nq, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d], 1)


def _search(entry, g, kind, q, capacity):
    """Calls `entry` (find_into or find_into_noalloc) on device queries q;
    returns rc, offsets, counts, total and the candidate buffer with
    `capacity` slots and 64 guard words after them."""
    nq = q.shape[0]
    base = q.data_ptr()
    if kind == L.QUERY_POINTS:
        x, y, z, stride = base, base + 8, base + 16, 3
    else:
        x, y, z, stride = base, None, None, 6
    off = torch.full((nq + 64,), GUARD, dtype=torch.int32, device=DEV)
    cnt = torch.full((nq + 64,), GUARD, dtype=torch.int32, device=DEV)
    buf = torch.full((capacity + 64,), GUARD, dtype=torch.int32, device=DEV)
    tot = ctypes.c_int64(-1)
    rc = getattr(L.lib(), entry)(g._h, kind, ctypes.c_void_p(x),
                                 ctypes.c_void_p(y) if y else None,
                                 ctypes.c_void_p(z) if z else None, stride, nq, g._tol,
                                 spin._ptr(off), spin._ptr(cnt), spin._ptr(buf), capacity,
                                 ctypes.byref(tot), None)
    torch.cuda.synchronize()
    return rc, off.cpu().numpy(), cnt.cpu().numpy(), tot.value, buf.cpu().numpy()


def _compare(g, kind, q, capacity=None):
    """noalloc against find_into on the same inputs; returns the total."""
    rc, off, cnt, tot, _ = _search("axsq_bvh_find_into", g, kind, q, 0)
    assert rc == 0, L.lib().axsq_last_error()
    cap = tot if capacity is None else capacity
    want = _search("axsq_bvh_find_into", g, kind, q, cap)
    got = _search("axsq_bvh_find_into_noalloc", g, kind, q, cap)
    assert got[0] == 0, L.lib().axsq_last_error()
    assert want[0] == 0
    assert got[3] == want[3] == tot
    for a, b in zip(got[1:3], want[1:3]):
        assert np.array_equal(a, b)
        assert np.all(a[q.shape[0]:] == GUARD)  # nothing past nq
    assert np.array_equal(got[4], want[4])  # ids, and the guard words after them
    assert np.all(got[4][cap:] == GUARD)
    return tot, got


@pytest.mark.parametrize("fp_mode", ["fp64", "fp32"])
@pytest.mark.parametrize("kind", [L.QUERY_POINTS, L.QUERY_BOXES, L.QUERY_RAYS])
def test_matches_find_into_on_successive_calls(kind, fp_mode):
    """A large call (fewer kept ids per query, many queries with more than
    it keeps), then a smaller one (64 kept ids) that reuses the array, then a
    large one again; the smaller one is also checked against the oracle."""
    boxes = _boxes(30_000, 11, 0.005, 0.06)
    g = spin.BVH(torch.as_tensor(boxes, device=DEV), fp_mode=fp_mode)
    big = torch.as_tensor(_queries(kind, 400_000, 5), device=DEV)
    small_np = _queries(kind, 3_000, 6)
    small = torch.as_tensor(small_np, device=DEV)
    tot_big, got = _compare(g, kind, big)
    counts = got[2][:big.shape[0]]
    assert (counts > kept_ids(big.shape[0])).sum() > 100  # the search-again path runs
    tot_small, got_small = _compare(g, kind, small)
    assert tot_small < tot_big
    name = {L.QUERY_POINTS: "points", L.QUERY_BOXES: "boxes", L.QUERY_RAYS: "rays"}[kind]
    o_off, o_cnt, o_ids = orc.OracleBVH(boxes, 1.000123).find(name, small_np)
    assert np.array_equal(got_small[1][:3000], o_off)
    assert np.array_equal(got_small[2][:3000], o_cnt)
    assert np.array_equal(got_small[4][:tot_small], o_ids)
    _compare(g, kind, big)


def test_maximum_size_and_beyond():
    boxes = _boxes(20_000, 12, 0.001, 0.02)
    g = spin.BVH(torch.as_tensor(boxes, device=DEV))
    q = torch.as_tensor(_queries(L.QUERY_POINTS, MAX_QUERIES + 1, 7), device=DEV)
    # the largest size the array serves
    _compare(g, L.QUERY_POINTS, q[:MAX_QUERIES])
    # one more is refused before any write
    rc, off, cnt, tot, buf = _search("axsq_bvh_find_into_noalloc", g, L.QUERY_POINTS, q, 1000)
    assert rc == L.ERRORS_BY_NAME["AXSQ_ERR_INVALID_ARG"]
    assert b"maximum" in L.lib().axsq_last_error()
    assert tot == 0
    assert np.all(off == GUARD) and np.all(cnt == GUARD) and np.all(buf == GUARD)
    # and a smaller call right after works
    _compare(g, L.QUERY_POINTS, q[:1000])


def test_capacity_too_small_writes_no_id():
    boxes = _boxes(20_000, 13, 0.005, 0.04)
    g = spin.BVH(torch.as_tensor(boxes, device=DEV))
    q = torch.as_tensor(_queries(L.QUERY_POINTS, 50_000, 8), device=DEV)
    rc, _, _, tot, _ = _search("axsq_bvh_find_into", g, L.QUERY_POINTS, q, 0)
    assert rc == 0 and tot > 1
    _compare(g, L.QUERY_POINTS, q, capacity=tot - 1)  # counts, offsets, total; no id
    rc, _, _, tot2, buf = _search("axsq_bvh_find_into_noalloc", g, L.QUERY_POINTS, q, tot - 1)
    assert rc == 0 and tot2 == tot and np.all(buf == GUARD)
    # no queries: nothing to do, total 0
    rc, _, _, tot0, _ = _search("axsq_bvh_find_into_noalloc", g, L.QUERY_POINTS, q[:0], 0)
    assert rc == 0 and tot0 == 0
