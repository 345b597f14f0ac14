"""GPU parity at the full mesh sizes of the three largest workloads, with
query subsets that the oracle finishes in seconds. At these sizes the first
pass's step budget is scaled up (23,2xx steps at 2,097,152 triangles, 5,793
at 327,680, 16,384 at 1,310,720), fewer hot-walk tables fit in their 2 GB,
and the walk records are sized from the previous call.

- torus R=1, r=0.3+0.02 sin(7u) cos(5v), 1024 x 1024 quads = 2,097,152
  triangles; the first 200,000 of the 16M queries' half-box / half-near-
  surface mix: FP64 bit-exact against the oracle; FP32 traversal equal to the
  FP64 distances, no sign mismatch.
- icosphere L7 (327,680 triangles) with radial displacement
  0.1 sin(4 theta) cos(3 phi); 200,000 near-surface queries and an N(0,1)
  upstream gradient: forward and adjoint bit-exact against the oracle.
- icosphere L8 (1,310,720 triangles) fitted to a rounded box: two iterations
  of rebuild, forward, L = mean((d - t)^2), adjoint, V -= 1e-3 dL/dV over the
  first 50,000 of the 8M queries. With the FP64 forward the vertices are
  bit-identical to the same loop driven by the oracle's forward and adjoint
  in NumPy; with the FP32-traversal forward (the workload's form) the loss
  and vertices equal the FP64 loop's.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from axom_amd import quest, shapefit, synth  # noqa: E402
from oracle import oracle as orc  # noqa: E402

pytestmark = pytest.mark.gpu

DEV = "cuda"


def _t(a):
    return torch.as_tensor(np.ascontiguousarray(a), device=DEV)


def _eval(sdf, q, mode):
    n = q.shape[0]
    cp = torch.empty((n, 3), dtype=torch.float64, device=DEV)
    elem = torch.empty(n, dtype=torch.int32, device=DEV)
    d = sdf.computeDistances(n, _t(q), outClosestPts=cp, outElem=elem, fp_mode=mode)
    return d.cpu().numpy(), cp.cpu().numpy(), elem.cpu().numpy()


def test_torus_full_mesh():
    v, t, q = synth.workload3(nq=200_000)
    assert t.shape[0] == 2_097_152
    sdf = quest.SignedDistance(_t(v), _t(t), fp_mode="fp32")
    d64, cp64, e64 = _eval(sdf, q, "fp64")
    o = orc.OracleSDF(v, t).eval(q, nthreads=16, want_normal=False)
    assert np.array_equal(d64, o["dist"])
    assert np.array_equal(np.signbit(d64), np.signbit(o["dist"]))
    assert np.array_equal(cp64, o["cp"])
    assert np.array_equal(e64, o["elem"])
    d32, _, _ = _eval(sdf, q, "fp32")
    rel = np.abs(d32 - d64) / np.maximum(np.abs(d64), 1e-300)
    assert rel.max() <= 1e-12
    assert np.array_equal(np.signbit(d32), np.signbit(d64))


def test_wavy_l7_forward_and_adjoint():
    v, t, q, g = synth.workload4(nq=200_000, level=7)
    assert t.shape[0] == 327_680
    sdf = quest.SignedDistance(_t(v), _t(t))
    qd = _t(q)
    elem = torch.empty(q.shape[0], dtype=torch.int32, device=DEV)
    d = sdf.computeDistances(q.shape[0], qd, outElem=elem)
    gq, gv = sdf.backward(qd, d, elem, _t(g))
    o = orc.OracleSDF(v, t)
    od = o.eval(q, nthreads=16, want_normal=False)
    assert np.array_equal(d.cpu().numpy(), od["dist"])
    assert np.array_equal(elem.cpu().numpy(), od["elem"])
    rq, rv = o.backward(q, od["dist"], od["elem"], g)
    assert np.array_equal(gq.cpu().numpy(), rq)
    assert np.array_equal(gv.cpu().numpy(), rv)


def oracle_shape_fit(v, t, q, target, iterations, lr):
    """The loop of axom_amd.shapefit driven by the oracle, in NumPy."""
    V = v.copy()
    for _ in range(iterations):
        o = orc.OracleSDF(V, t)
        r = o.eval(q, nthreads=16, want_cp=False, want_normal=False)
        g = ((r["dist"] - target) * (2.0 / q.shape[0])).astype(np.float32)
        _, gv = o.backward(q, r["dist"], r["elem"], g)
        V = V - lr * gv
    return V


@pytest.mark.parametrize("level,nq", [(3, 20_000), (8, 50_000)])
def test_shape_fit_matches_oracle_loop(level, nq):
    v, t, q, target = synth.workload5(nq=nq, level=level)
    want = oracle_shape_fit(v, t, q, target, 2, 1e-3)
    got, hist = shapefit.shape_fit(_t(v), _t(t), _t(q), _t(target), iterations=2, lr=1e-3,
                                   fp_mode="fp64")
    assert not np.array_equal(want, v)  # the step moved the vertices
    assert np.array_equal(got.cpu().numpy(), want)
    # the workload's form: FP32 traversal for the forward
    got32, hist32 = shapefit.shape_fit(_t(v), _t(t), _t(q), _t(target), iterations=2, lr=1e-3,
                                       fp_mode="fp32")
    assert [h["loss"] for h in hist32] == [h["loss"] for h in hist]
    assert np.array_equal(got32.cpu().numpy(), want)
    if level == 8:  # a large tree: the second iteration walks the first's longest early
        assert hist[1]["early_walks"] > 0 and hist32[1]["early_walks"] > 0


def test_shape_fit_normalized_step_is_opt_in():
    v, t, q, target = synth.workload5(nq=5000, level=2)
    plain, _ = shapefit.shape_fit(_t(v), _t(t), _t(q), _t(target), iterations=1, fp_mode="fp64")
    norm, _ = shapefit.shape_fit(_t(v), _t(t), _t(q), _t(target), iterations=1, fp_mode="fp64",
                                 step="normalized")
    moved = torch.linalg.vector_norm(norm - _t(v), dim=1)
    assert abs(float(moved.max()) - 1e-3) < 1e-12  # the largest step has length lr
    assert not torch.equal(plain, norm)
    with pytest.raises(ValueError):
        shapefit.shape_fit(_t(v), _t(t), _t(q), _t(target), iterations=1, step="adam")
