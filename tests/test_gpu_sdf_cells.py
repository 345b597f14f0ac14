"""GPU parity of SignedDistance over quad and mixed triangle/quad cells
(axsq_sdf_create_cells) against the oracle. Bar as for triangle meshes:
distances, closest points, signs and winning cells bit-exact; unit normals
within 1e-12."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from axom_amd import quest, synth  # noqa: E402
from oracle import oracle as orc  # noqa: E402
from tests.test_gpu_sdf import _compare, _eval, _t  # noqa: E402
from tests.test_oracle_quads import mixed_cells, quad_cells  # noqa: E402

pytestmark = pytest.mark.gpu


def _pair(v, conn, off, **kw):
    g = quest.SignedDistance(_t(v), _t(conn), cell_offsets=_t(off), **kw)
    o = orc.OracleSDF(v, conn, cell_offsets=off, watertight=kw.get("isWatertight", True),
                      compute_sign=kw.get("computeSign", True))
    return g, o


@pytest.mark.parametrize("n,nq", [(4, 20_000), (24, 200_000)])
def test_quad_sphere_bit_exact(n, nq):
    v, quads = synth.cube_sphere(n)
    conn, off = quad_cells(quads)
    q = synth.uniform(41 + n, 0, 3 * nq, -1.5, 1.5).reshape(-1, 3)
    g, o = _pair(v, conn, off)
    _compare(_eval(g, q), o.eval(q, nthreads=8))


def test_mixed_cells_and_planar_cube_bit_exact():
    v, quads = synth.cube_sphere(16)
    conn, off, _ = mixed_cells(quads)
    q = synth.uniform(43, 0, 3 * 100_000, -1.5, 1.5).reshape(-1, 3)
    g, o = _pair(v, conn, off)
    _compare(_eval(g, q), o.eval(q, nthreads=8))
    v, quads = synth.cube_sphere(5, project=False)
    conn, off = quad_cells(quads)
    q = synth.uniform(44, 0, 3 * 50_000, -2.0, 2.0).reshape(-1, 3)
    g, o = _pair(v, conn, off, isWatertight=False)
    _compare(_eval(g, q), o.eval(q, nthreads=8))


def test_quads_near_surface_and_without_signs():
    v, quads = synth.cube_sphere(12)
    conn, off = quad_cells(quads)
    r = 1.0 + synth.uniform(45, 0, 50_000, -0.01, 0.01)
    d = synth.uniform(45, 1, 3 * 50_000, -1.0, 1.0).reshape(-1, 3)
    q = d / np.linalg.norm(d, axis=1, keepdims=True) * r[:, None]
    g, o = _pair(v, conn, off, computeSign=False)
    _compare(_eval(g, q), o.eval(q, nthreads=8))


@pytest.mark.parametrize("budget", [1, 37])
def test_quad_suspended_walks_bit_exact(monkeypatch, budget):
    """Quad-cell walks cut off by a small step budget and finished by the
    wave-cooperative second pass (two triangle records per memo lane) give the
    oracle's answers; budget 1 sends every walk to the second pass, and the
    visit and leaf counts stay the reference traversal's."""
    monkeypatch.setenv("AXSQ_SDF_BUDGET", str(budget))
    v, quads = synth.cube_sphere(12)
    conn, off, _ = mixed_cells(quads)
    q = np.concatenate([synth.uniform(47, 0, 3 * 30_000, -1.5, 1.5).reshape(-1, 3),
                        0.05 * synth.uniform(47, 1, 3 * 3000, -1.0, 1.0).reshape(-1, 3)])
    g, o = _pair(v, conn, off)
    g.enableStats(True)
    got = _eval(g, q)
    assert g.lastSuspended() > (q.shape[0] // 2 if budget == 1 else 0)
    want = o.eval(q, nthreads=8, want_stats=True)
    _compare(got, want)
    visits, leaves = g.lastStats()
    assert visits == int(want["visits"].sum()) and leaves == int(want["leaves"].sum())


def test_triangle_cells_take_the_triangle_path():
    v, t, q = synth.workload2(nq=50_000, level=4)
    off = np.arange(0, 3 * t.shape[0] + 1, 3, dtype=np.int32)
    a = quest.SignedDistance(_t(v), _t(t.reshape(-1)), cell_offsets=_t(off))
    b = quest.SignedDistance(_t(v), _t(t))
    ga, gb = _eval(a, q), _eval(b, q)
    for k in ga:
        assert np.array_equal(ga[k], gb[k]), k


def test_quad_stats_and_adjoint_refusal():
    v, quads = synth.cube_sphere(8)
    conn, off = quad_cells(quads)
    q = synth.uniform(46, 0, 3 * 5000, -1.3, 1.3).reshape(-1, 3)
    g, o = _pair(v, conn, off)
    g.enableStats(True)
    _eval(g, q)
    visits, leaves = g.lastStats()
    r = o.eval(q, want_stats=True)
    assert visits == int(r["visits"].sum()) and leaves == int(r["leaves"].sum())
    n = q.shape[0]
    with pytest.raises(RuntimeError):
        g.backward(_t(q), torch.zeros(n, dtype=torch.float64, device="cuda"),
                   torch.zeros(n, dtype=torch.int32, device="cuda"),
                   torch.ones(n, dtype=torch.float32, device="cuda"))
