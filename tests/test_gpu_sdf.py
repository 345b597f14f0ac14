"""GPU parity of SignedDistance (forward, FP32-traversal form, adjoint)
against the oracle.

Bars: FP64 form -- distances, closest points, signs and winning elements
bit-exact; unit normals within 1e-12 (device acos may differ from glibc by an
ulp). FP32-traversal form -- distances within 1e-12 relative of the FP64
oracle (the FP64 finish makes them equal in practice), no sign mismatch.
Adjoint -- bit-exact against the oracle's identically ordered sums.
"""
import json
import math
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from axom_amd import quest, synth  # noqa: E402
from oracle import oracle as orc  # noqa: E402
from tests import refmeshes  # noqa: E402

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_known_answers.json")))
DEV = "cuda"


def _t(a):
    return torch.as_tensor(np.ascontiguousarray(a), device=DEV)


def _eval(sdf, q, fp_mode=None):
    n = q.shape[0]
    cp = torch.empty((n, 3), dtype=torch.float64, device=DEV)
    nrm = torch.empty((n, 3), dtype=torch.float64, device=DEV)
    elem = torch.empty(n, dtype=torch.int32, device=DEV)
    d = sdf.computeDistances(n, _t(q), outClosestPts=cp, outNormals=nrm, outElem=elem,
                             fp_mode=fp_mode)
    return dict(dist=d.cpu().numpy(), cp=cp.cpu().numpy(), normal=nrm.cpu().numpy(),
                elem=elem.cpu().numpy())


def _compare(g, o, exact_elem=True):
    assert np.array_equal(g["dist"], o["dist"])
    assert np.array_equal(np.signbit(g["dist"]), np.signbit(o["dist"]))
    assert np.array_equal(g["cp"], o["cp"])
    if exact_elem:
        assert np.array_equal(g["elem"], o["elem"])
    np.testing.assert_allclose(g["normal"], o["normal"], rtol=0, atol=1e-12)


@pytest.mark.parametrize("level,nq,watertight", [(3, 20_000, True), (6, 200_000, True),
                                                 (4, 20_000, False)])
def test_sdf_fp64_bit_exact(level, nq, watertight):
    v, t, q = synth.workload2(nq=nq, level=level)
    sdf = quest.SignedDistance(_t(v), _t(t), isWatertight=watertight)
    g = _eval(sdf, q)
    o = orc.OracleSDF(v, t, watertight=watertight).eval(q, nthreads=8)
    _compare(g, o)


def test_sdf_fp64_near_surface_torus():
    v, t = synth.torus(128, 128)
    q = synth.surface_samples(v, t, 5, 0, 100_000, -0.01, 0.01)
    sdf = quest.SignedDistance(_t(v), _t(t))
    _compare(_eval(sdf, q), orc.OracleSDF(v, t).eval(q, nthreads=8))


def test_sdf_without_signs():
    v, t, q = synth.workload2(nq=20_000, level=3)
    sdf = quest.SignedDistance(_t(v), _t(t), computeSign=False)
    g = _eval(sdf, q)
    o = orc.OracleSDF(v, t, compute_sign=False).eval(q)
    _compare(g, o)
    assert np.all(g["dist"] >= 0)


def test_sdf_soa_queries_and_unsorted_path():
    v, t, q = synth.workload2(nq=3000, level=3)  # below the reordering threshold
    sdf = quest.SignedDistance(_t(v), _t(t))
    d = sdf.computeDistances(q.shape[0], tuple(_t(q[:, k]) for k in range(3)))
    o = orc.OracleSDF(v, t).eval(q)
    assert np.array_equal(d.cpu().numpy(), o["dist"])


def test_sdf_single_triangle():
    v = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    t = np.array([[0, 1, 2]], np.int32)
    q = synth.uniform(3, 0, 3000, -1, 2).reshape(-1, 3)
    sdf = quest.SignedDistance(_t(v), _t(t), isWatertight=False)
    _compare(_eval(sdf, q), orc.OracleSDF(v, t, watertight=False).eval(q))


@pytest.mark.parametrize("fp_mode", ["fp64", "fp32"])
def test_non_finite_queries(fp_mode):
    # a NaN or infinite coordinate gives no box distance that passes the
    # pruning test against the initial bound, so the result keeps its initial
    # state (element -1); the traversal counts must agree as well
    v, t, q = synth.workload2(nq=4096, level=4)
    q = q.copy()
    nan, inf = np.nan, np.inf
    rows = [[nan, 0, 0], [0.1, nan, 0.2], [0, 0, nan], [inf, 0, 0], [0, -inf, 0],
            [inf, inf, -inf], [nan, inf, 0], [-inf, 0.3, nan]]
    idx = np.arange(len(rows)) * 509
    q[idx] = rows
    sdf = quest.SignedDistance(_t(v), _t(t), fp_mode=fp_mode)
    sdf.enableStats(True)
    g_out = _eval(sdf, q, fp_mode=fp_mode)
    o = orc.OracleSDF(v, t).eval(q, want_stats=True)
    assert np.all(o["elem"][idx] == -1)
    _compare(g_out, o)
    if fp_mode == "fp64":
        visits, leaves = sdf.lastStats()
        assert visits == int(o["visits"].sum()) and leaves == int(o["leaves"].sum())
    # the adjoint: zero query gradient, no vertex contribution for elem -1
    g = synth.normal_f32(5, 6, q.shape[0])
    g_dist = g_out["dist"]
    gq, gv = sdf.backward(_t(q), _t(g_dist), _t(g_out["elem"]), _t(g))
    rq, rv = orc.OracleSDF(v, t).backward(q, g_dist, g_out["elem"], g)
    assert np.array_equal(gq.cpu().numpy(), rq)
    assert np.array_equal(gv.cpu().numpy(), rv)
    bad = g_out["elem"].copy()
    bad[1] = t.shape[0]
    with pytest.raises(RuntimeError):
        sdf.backward(_t(q), _t(g_dist), _t(bad), _t(g))


def _degenerate_mesh():
    # an icosphere with zero-area triangles on its surface: a repeated corner
    # on an existing edge, a point triangle at a vertex, and a sliver whose
    # third corner is the midpoint of an edge (Triangle::degenerate,
    # Triangle.hpp:327-330, drops them from the pseudo-normal sums)
    v, t = synth.icosphere(3)
    nv = v.shape[0]
    e = t[:40]
    mid = 0.5 * (v[e[:, 0]] + v[e[:, 1]])
    v2 = np.concatenate([v, mid])
    t2 = np.concatenate([t, np.stack([e[:, 0], e[:, 0], e[:, 1]], 1),
                         np.stack([e[:, 2], e[:, 2], e[:, 2]], 1),
                         np.stack([e[:, 0], e[:, 1], nv + np.arange(40)], 1)]).astype(np.int32)
    q = np.concatenate([synth.surface_samples(v, t, 7, 0, 20_000, -0.05, 0.05),
                        v[e.ravel()] * 1.01, mid * 0.99,
                        synth.uniform(7, 1, 30_000, -1.3, 1.3).reshape(-1, 3)])
    return v2, t2, q, t.shape[0]


@pytest.mark.parametrize("fp_mode", ["fp64", "fp32"])
def test_degenerate_triangles(fp_mode):
    v, t, q, nt0 = _degenerate_mesh()
    sdf = quest.SignedDistance(_t(v), _t(t), fp_mode=fp_mode)
    g_out = _eval(sdf, q, fp_mode)
    o = orc.OracleSDF(v, t).eval(q, nthreads=8)
    assert np.sum(o["elem"] >= nt0) > 50  # queries won by a degenerate triangle
    _compare(g_out, o)
    g = synth.normal_f32(7, 2, q.shape[0])
    gq, gv = sdf.backward(_t(q), _t(g_out["dist"]), _t(g_out["elem"]), _t(g))
    rq, rv = orc.OracleSDF(v, t).backward(q, o["dist"], o["elem"], g)
    assert np.array_equal(gq.cpu().numpy(), rq)
    assert np.array_equal(gv.cpu().numpy(), rv)


def test_uv_sphere_pole_rows_bit_exact():
    # the reference's test sphere: its pole rows hold zero-area triangles
    vs, ts = refmeshes.sphere_uv_mesh([0.0, 0.0, 0.0], 1.0, 12, 10)
    rng = np.random.default_rng(31)
    cap = np.concatenate([rng.uniform(-0.05, 0.05, (4000, 2)),
                          rng.choice([-1.0, 1.0], (4000, 1)) * rng.uniform(0.9, 1.1, (4000, 1))], 1)
    q = np.concatenate([cap, rng.uniform(-1.2, 1.2, (8000, 3))])
    sdf = quest.SignedDistance(_t(vs), _t(ts))
    _compare(_eval(sdf, q), orc.OracleSDF(vs, ts).eval(q, nthreads=8))


def test_known_answer_sphere():
    s = GOLD["sdf_sphere"]
    v, t = refmeshes.sphere_uv_mesh(s["center"], s["radius"], s["theta_res"], s["phi_res"])
    sdf = quest.SignedDistance(_t(v), _t(t), s["watertight"], s["compute_signs"])
    q = refmeshes.uniform_grid_nodes(v.min(0) - s["grid_expand"], v.max(0) + s["grid_expand"],
                                     s["grid_nodes"])
    phi = _eval(sdf, q)["dist"]
    diff = phi - (np.linalg.norm(q, axis=1) - s["radius"])
    assert np.all(np.abs(diff) <= s["per_point_tol"])
    assert abs(np.sum(np.abs(diff)) - s["l1"]) <= s["tol"]
    assert abs(math.sqrt(np.sum(diff)) - s["l2_of_sum"]) <= s["tol"]
    assert abs(np.max(np.abs(diff)) - s["linf"]) <= s["tol"]


def test_known_answer_plane_via_c_api(tmp_path):
    p = GOLD["sdf_plane"]
    v, t = refmeshes.plane_soup(p["triangles"])
    from axom_amd import stl
    path = str(tmp_path / "plane.stl")
    stl.write_ascii_stl(path, v, t)
    quest.signed_distance_finalize()
    quest.signed_distance_set_closed_surface(False)
    assert quest.signed_distance_init(path) == 0
    try:
        with pytest.raises(RuntimeError):
            quest.signed_distance_set_compute_signs(True)  # set after init
        q = refmeshes.uniform_grid_nodes(p["grid"]["lo"], p["grid"]["hi"], p["grid"]["nodes"])
        x, y, z = (_t(q[:, k]) for k in range(3))
        phi = quest.signed_distance_evaluate(x, y, z).cpu().numpy()
        ulps = np.abs(phi.view(np.int64) - q[:, 2].view(np.int64))
        assert np.all(ulps <= 4)
        assert quest.signed_distance_evaluate(0.5, 0.5, 2.0) == 2.0
    finally:
        quest.signed_distance_finalize()
        quest.signed_distance_set_closed_surface(True)
    with pytest.raises(RuntimeError):
        quest.signed_distance_evaluate(0.0, 0.0, 0.0)


@pytest.mark.parametrize("wl", [2, 3])
def test_sdf_fp32_traversal_matches_fp64(wl):
    if wl == 2:
        v, t, q = synth.workload2(nq=200_000)
    else:
        v, t = synth.torus(256, 256)
        q = np.concatenate([synth.surface_samples(v, t, 8, 0, 100_000, -0.01, 0.01),
                            synth.uniform(8, 9, 300_000, -1.6, 1.6).reshape(-1, 3)])
    sdf = quest.SignedDistance(_t(v), _t(t), fp_mode="fp32")
    g32 = _eval(sdf, q, "fp32")
    g64 = _eval(sdf, q, "fp64")
    rel = np.abs(g32["dist"] - g64["dist"]) / np.maximum(np.abs(g64["dist"]), 1e-300)
    assert rel.max() <= 1e-12
    assert np.array_equal(np.signbit(g32["dist"]), np.signbit(g64["dist"]))
    # and the FP64 form equals the oracle on a sample
    o = orc.OracleSDF(v, t).eval(q[::20], nthreads=8)
    assert np.array_equal(g64["dist"][::20], o["dist"])


def test_backward_bit_exact():
    v, t, q, g = synth.workload4(nq=200_000, level=5)
    sdf = quest.SignedDistance(_t(v), _t(t))
    qd = _t(q)
    elem = torch.empty(q.shape[0], dtype=torch.int32, device=DEV)
    d = sdf.computeDistances(q.shape[0], qd, outElem=elem)
    gq, gv = sdf.backward(qd, d, elem, _t(g))
    o = orc.OracleSDF(v, t)
    od = o.eval(q, nthreads=8)
    assert np.array_equal(d.cpu().numpy(), od["dist"])
    rq, rv = o.backward(q, od["dist"], od["elem"], g)
    assert np.array_equal(gq.cpu().numpy(), rq)
    assert np.array_equal(gv.cpu().numpy(), rv)
    # deterministic run to run
    gq2, gv2 = sdf.backward(qd, d, elem, _t(g))
    assert torch.equal(gq, gq2) and torch.equal(gv, gv2)


def test_update_vertices_rebuilds():
    v, t, q = synth.workload2(nq=20_000, level=4)
    sdf = quest.SignedDistance(_t(v), _t(t))
    v2 = v * 1.1
    sdf.updateVertices(_t(v2))
    assert np.array_equal(_eval(sdf, q)["dist"], orc.OracleSDF(v2, t).eval(q)["dist"])


def test_traversal_stats_match_oracle():
    v, t, q = synth.workload2(nq=20_000, level=4)
    sdf = quest.SignedDistance(_t(v), _t(t))
    sdf.enableStats(True)
    sdf.computeDistances(q.shape[0], _t(q))
    visits, leaves = sdf.lastStats()
    o = orc.OracleSDF(v, t).eval(q, want_stats=True)
    assert visits == int(o["visits"].sum()) and leaves == int(o["leaves"].sum())


def _sdf_with(monkeypatch, v, t, budget, **kw):
    # the first pass's step budget is read when the SignedDistance is created
    monkeypatch.setenv("AXSQ_SDF_BUDGET", str(budget))
    return quest.SignedDistance(_t(v), _t(t), **kw)


@pytest.mark.parametrize("budget,level,inner", [(1, 4, 0.05), (37, 4, 0.05), (400, 4, 0.05),
                                                (1, 5, 0.3), (37, 5, 0.3)])
def test_suspended_walks_resume_bit_exact(monkeypatch, budget, level, inner):
    """Walks cut off by a small step budget and finished by the second pass
    give the oracle's answers: budget 1 sends every walk to the second pass.
    Queries inside the sphere (radius `inner`) make the long walks."""
    v, t, q = synth.workload2(nq=20_000, level=level)
    q = np.concatenate([q, inner * synth.uniform(11, 0, 9000, -1, 1).reshape(-1, 3)])
    sdf = _sdf_with(monkeypatch, v, t, budget)
    g = _eval(sdf, q)
    assert sdf.lastSuspended() > (q.shape[0] // 2 if budget == 1 else 0)
    _compare(g, orc.OracleSDF(v, t).eval(q, nthreads=8))


@pytest.mark.parametrize("fp_mode", ["fp64", "fp32"])
def test_repeated_calls_walk_the_longest_early(monkeypatch, fp_mode):
    """A later call on the same query arrays walks the previous call's longest
    walks from the start beside its first pass, which skips them. Every call
    gives the same answers as the first: the oracle's, bit for bit with the
    same visit and leaf counts in FP64 (FP32 traversal: within 1e-12 of the
    FP64 oracle, and identical from call to call). So does a call after the
    queries were rewritten in place, whose remembered walks are then ordinary
    ones, a call after they were written back, and calls on two streams in
    turn."""
    v, t, q = synth.workload2(nq=20_000, level=5)
    q = np.concatenate([q, 0.3 * synth.uniform(11, 0, 9000, -1, 1).reshape(-1, 3)])
    q2 = np.concatenate([synth.uniform(12, 0, 60_000, -1.5, 1.5).reshape(-1, 3),
                         0.2 * synth.uniform(13, 0, 9000, -1, 1).reshape(-1, 3)])
    assert q2.shape == q.shape
    mode = None if fp_mode == "fp64" else "fp32"
    monkeypatch.setenv("AXSQ_SDF_EARLY_WALKS", "64")  # on small trees too (read at creation)
    sdf = _sdf_with(monkeypatch, v, t, 37, fp_mode=fp_mode)
    sdf.enableStats(fp_mode == "fp64")
    o = orc.OracleSDF(v, t)
    qt = _t(q)
    n = q.shape[0]

    def run():
        cp = torch.empty((n, 3), dtype=torch.float64, device=DEV)
        nrm = torch.empty((n, 3), dtype=torch.float64, device=DEV)
        elem = torch.empty(n, dtype=torch.int32, device=DEV)
        d = sdf.computeDistances(n, qt, outClosestPts=cp, outNormals=nrm, outElem=elem,
                                 fp_mode=mode)
        return dict(dist=d.cpu().numpy(), cp=cp.cpu().numpy(), normal=nrm.cpu().numpy(),
                    elem=elem.cpu().numpy())

    def check(g, pts, first=None):
        ref = o.eval(pts, nthreads=8, want_stats=True)
        if fp_mode == "fp64":
            _compare(g, ref)
            assert sdf.lastStats() == (int(ref["visits"].sum()), int(ref["leaves"].sum()))
        else:
            np.testing.assert_allclose(g["dist"], ref["dist"], rtol=1e-12, atol=0)
            assert np.array_equal(np.signbit(g["dist"]), np.signbit(ref["dist"]))
            if first is not None:
                for k in ("dist", "cp", "normal", "elem"):
                    assert np.array_equal(g[k], first[k]), k

    g1 = run()
    assert sdf.lastEarly() == 0 and sdf.lastSuspended() > 64
    check(g1, q)
    for _ in range(2):
        g = run()
        assert sdf.lastEarly() == 64
        check(g, q, g1)
    qt.copy_(_t(q2))  # same arrays, new queries
    g = run()
    assert sdf.lastEarly() == 64
    check(g, q2)
    qt.copy_(_t(q))
    g = run()
    assert sdf.lastEarly() == 64
    check(g, q, g1)
    # calls on two streams in turn, with no synchronisation between them
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    outs = []
    for st in (s1, s2, s1):
        with torch.cuda.stream(st):
            outs.append(run())
    torch.cuda.synchronize()
    for g in outs:
        check(g, q, g1)


@pytest.mark.parametrize("tail", ["0", "50", "90"])
def test_wave_order_from_the_previous_call(monkeypatch, tail):
    """A call on the same query arrays as the previous one launches the first
    pass's waves that took fewest steps then last (AXSQ_SDF_WAVE_TAIL percent
    of them). Every call gives the oracle's answers and visit and leaf counts:
    repeated calls, a call after the queries were rewritten in place (the
    remembered steps are then another set's), and a call on fewer queries of
    the same storage."""
    monkeypatch.setenv("AXSQ_SDF_WAVE_TAIL", tail)  # read at creation
    v, t, q = synth.workload2(nq=50_000, level=5)
    q2 = synth.uniform(14, 0, 150_000, -1.2, 1.2).reshape(-1, 3)
    sdf = _sdf_with(monkeypatch, v, t, 400)
    sdf.enableStats(True)
    o = orc.OracleSDF(v, t)
    qt = _t(q)

    def run_check(pts, n):
        cp = torch.empty((n, 3), dtype=torch.float64, device=DEV)
        nrm = torch.empty((n, 3), dtype=torch.float64, device=DEV)
        elem = torch.empty(n, dtype=torch.int32, device=DEV)
        d = sdf.computeDistances(n, qt[:n], outClosestPts=cp, outNormals=nrm, outElem=elem)
        g = dict(dist=d.cpu().numpy(), cp=cp.cpu().numpy(), normal=nrm.cpu().numpy(),
                 elem=elem.cpu().numpy())
        ref = o.eval(pts[:n], nthreads=8, want_stats=True)
        _compare(g, ref)
        assert sdf.lastStats() == (int(ref["visits"].sum()), int(ref["leaves"].sum()))

    n = q.shape[0]
    for _ in range(3):
        run_check(q, n)
    qt.copy_(_t(q2))
    run_check(q2, n)
    run_check(q2, n)
    run_check(q2, 30_000)


def test_early_walks_off_on_small_trees(monkeypatch):
    """By default a tree this small (step budget 2048) runs no early walks."""
    monkeypatch.delenv("AXSQ_SDF_EARLY_WALKS", raising=False)
    monkeypatch.delenv("AXSQ_SDF_BUDGET", raising=False)
    v, t, q = synth.workload2(nq=20_000, level=4)
    q = np.concatenate([q, 0.3 * synth.uniform(11, 0, 9000, -1, 1).reshape(-1, 3)])
    sdf = quest.SignedDistance(_t(v), _t(t))
    qt = _t(q)
    for _ in range(2):
        sdf.computeDistances(q.shape[0], qt)
    assert sdf.lastEarly() == 0


@pytest.mark.parametrize("watertight,compute_sign", [(False, True), (True, False)])
def test_suspended_walks_open_surface_and_unsigned(monkeypatch, watertight, compute_sign):
    v, t = synth.torus(96, 64)
    q = np.concatenate([synth.surface_samples(v, t, 5, 0, 20_000, -0.01, 0.01),
                        synth.uniform(5, 1, 60_000, -1.5, 1.5).reshape(-1, 3)])
    sdf = _sdf_with(monkeypatch, v, t, 1, isWatertight=watertight, computeSign=compute_sign)
    _compare(_eval(sdf, q), orc.OracleSDF(v, t, watertight=watertight,
                                          compute_sign=compute_sign).eval(q, nthreads=8))


def test_suspended_walks_fp32_traversal_unchanged(monkeypatch):
    """The FP32-traversal form visits other nodes than the FP64 form, so its
    check is against itself: cut into two passes it gives the one-pass answer
    bit for bit (distances, points, normals, elements)."""
    v, t, q = synth.workload2(nq=30_000, level=4)
    one = _eval(_sdf_with(monkeypatch, v, t, 0, fp_mode="fp32"), q, "fp32")
    two_sdf = _sdf_with(monkeypatch, v, t, 1, fp_mode="fp32")
    two = _eval(two_sdf, q, "fp32")
    assert two_sdf.lastSuspended() > q.shape[0] // 2
    for k in ("dist", "cp", "normal", "elem"):
        assert np.array_equal(one[k], two[k]), k


@pytest.mark.parametrize("pass2", ["fp64", "fp32"])
def test_fp32_traversal_second_pass_nodes(monkeypatch, pass2):
    """The FP32-traversal form's saved walks resume on the FP64 nodes (the
    default) or on the FP32 nodes (AXSQ_SDF_MIXED_PASS2=fp32). Either way the
    distances equal the FP64 oracle's within 1e-12 relative, with the same
    signs and closest points."""
    v, t, q = synth.workload2(nq=30_000, level=4)
    q = np.concatenate([q, 0.3 * synth.uniform(14, 0, 6000, -1, 1).reshape(-1, 3)])
    if pass2 == "fp32":
        monkeypatch.setenv("AXSQ_SDF_MIXED_PASS2", "fp32")
    else:
        monkeypatch.delenv("AXSQ_SDF_MIXED_PASS2", raising=False)
    sdf = _sdf_with(monkeypatch, v, t, 37, fp_mode="fp32")
    g = _eval(sdf, q, "fp32")
    assert sdf.lastSuspended() > q.shape[0] // 10
    o = orc.OracleSDF(v, t).eval(q, nthreads=8)
    rel = np.abs(g["dist"] - o["dist"]) / np.maximum(np.abs(o["dist"]), 1e-300)
    assert rel.max() <= 1e-12
    assert np.array_equal(np.signbit(g["dist"]), np.signbit(o["dist"]))
    np.testing.assert_allclose(g["cp"], o["cp"], rtol=0, atol=1e-12)


def test_suspended_walks_stats_match_oracle(monkeypatch):
    v, t, q = synth.workload2(nq=20_000, level=4)
    sdf = _sdf_with(monkeypatch, v, t, 1)
    sdf.enableStats(True)
    sdf.computeDistances(q.shape[0], _t(q))
    visits, leaves = sdf.lastStats()
    o = orc.OracleSDF(v, t).eval(q, want_stats=True)
    assert visits == int(o["visits"].sum()) and leaves == int(o["leaves"].sum())


@pytest.mark.parametrize("budget", [0, 1])
def test_many_tied_triangles_at_a_vertex(monkeypatch, budget):
    """Queries whose closest point is a pole of a UV sphere, where 25 triangles
    meet: more pseudo-normal terms than the first pass records, so those walks
    are run again from the start by the second pass. Normals, signs and the
    winning element must still be the oracle's."""
    s = GOLD["sdf_sphere"]
    v, t = refmeshes.sphere_uv_mesh(s["center"], s["radius"], s["theta_res"], s["phi_res"])
    z = np.linspace(0.55, 2.0, 64)
    e = 1e-9 * synth.uniform(4, 0, 256, -1, 1).reshape(-1, 2)
    poles = np.concatenate([np.stack([0 * z, 0 * z, z], 1), np.stack([0 * z, 0 * z, -z], 1),
                            np.stack([e[:64, 0], e[:64, 1], z], 1),
                            np.stack([e[64:, 0], e[64:, 1], -z], 1)])
    q = np.concatenate([poles, synth.uniform(4, 1, 30_000, -1.0, 1.0).reshape(-1, 3)])
    sdf = _sdf_with(monkeypatch, v, t, budget)
    g = _eval(sdf, q)
    assert sdf.lastSuspended() >= 128  # the exact-axis pole queries at least
    o = orc.OracleSDF(v, t).eval(q, nthreads=8)
    _compare(g, o)
    # the pole queries are outside, with the pole as closest point
    assert np.all(g["dist"][:128] > 0)
    assert np.array_equal(g["cp"][:64], np.tile(v[0], (64, 1)))


@pytest.mark.parametrize("cap", [0, 16])
def test_walk_records_cap(monkeypatch, cap):
    """With room for at most `cap` (at least one) saved walks, the walks over
    the step budget beyond it finish in the first pass: the results are the
    oracle's."""
    v, t, q = synth.workload2(nq=20_000, level=4)
    monkeypatch.setenv("AXSQ_SDF_SUSPEND_CAP", str(cap))
    sdf = _sdf_with(monkeypatch, v, t, 37)
    g = _eval(sdf, q)
    _compare(g, orc.OracleSDF(v, t).eval(q, nthreads=8))
    one = _sdf_with(monkeypatch, v, t, 0)  # no budget: one pass, the same answers
    _compare(_eval(one, q), g)


def test_walk_records_grow_from_previous_call(monkeypatch):
    """Budget 1 suspends every walk. The first call has records for an
    eighth of the 200,000 queries (65,536 at least), so the rest finish in the
    first pass; the second call on the handle has room for all of them. Both
    give the oracle's answers."""
    v, t, q = synth.workload2(nq=200_000, level=4)
    sdf = _sdf_with(monkeypatch, v, t, 1)
    want = orc.OracleSDF(v, t).eval(q, nthreads=8)
    first = _eval(sdf, q)
    n_first = sdf.lastSuspended()
    second = _eval(sdf, q)
    n_second = sdf.lastSuspended()
    assert n_first < n_second and n_second >= q.shape[0] - 1000
    _compare(first, want)
    _compare(second, want)
