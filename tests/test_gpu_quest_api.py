"""GPU checks of the quest interface around SignedDistance and of the BVH's
VTK output:
- STL files (ASCII with coordinates that are not float32-exact, binary of
  the same mesh) through the library's reader, as SignedDistance.fromSTL
  and as the process-wide query of QUEST_signed_distance_init_serial: the
  distances, closest points and mesh bounds are the oracle's on the mesh read
  the reference's way (ASCII as doubles, binary widened from float);
- the single-point evaluations (computeDistance, QUEST_..._evaluate_0/1)
  against the batched kernel;
- option setters after init and evaluation before init are errors;
- BVH::writeVtkFile: the file's text equals one written from the oracle's
  tree by bvh_vtkio.hpp's rules.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from axom_amd import quest, spin, stl, synth  # noqa: E402
from oracle import oracle as orc  # noqa: E402

pytestmark = pytest.mark.gpu

DEV = "cuda"


def _t(a):
    return torch.as_tensor(np.ascontiguousarray(a), device=DEV)


def _mesh():
    v, t = synth.icosphere(3)
    v = v * (1.0 + synth.uniform(5, 1, v.shape[0], -1e-3, 1e-3))[:, None]
    return v, t


@pytest.mark.parametrize("fmt", ["ascii", "binary"])
def test_stl_to_signed_distance(tmp_path, fmt):
    v, t = _mesh()
    path = str(tmp_path / f"sphere_{fmt}.stl")
    (stl.write_ascii_stl if fmt == "ascii" else stl.write_binary_stl)(path, v, t)
    sv, st = stl.read_stl(path)  # the reference's reading (pinned by tests/test_stl.py)
    if fmt == "ascii":
        assert np.array_equal(sv, v[t.reshape(-1)])
    else:
        assert np.array_equal(sv, v[t.reshape(-1)].astype(np.float32).astype(np.float64))
    q = synth.uniform(6, 2, 3 * 20_000, -1.5, 1.5).reshape(-1, 3)
    want = orc.OracleSDF(sv, st).eval(q, nthreads=8)

    sdf = quest.SignedDistance.fromSTL(path)
    n = q.shape[0]
    cp = torch.empty((n, 3), dtype=torch.float64, device=DEV)
    d = sdf.computeDistances(n, _t(q), outClosestPts=cp).cpu().numpy()
    assert np.array_equal(d, want["dist"])
    assert np.array_equal(cp.cpu().numpy(), want["cp"])
    lo, hi = sdf.getMeshBounds()
    assert np.array_equal(lo, sv.min(0)) and np.array_equal(hi, sv.max(0))

    quest.signed_distance_finalize()
    assert quest.signed_distance_init(path) == 0
    try:
        assert quest.signed_distance_initialized()
        x, y, z = (_t(q[:, k]) for k in range(3))
        phi = quest.signed_distance_evaluate(x, y, z).cpu().numpy()
        assert np.array_equal(phi, want["dist"])
        lo, hi = quest.signed_distance_get_mesh_bounds()
        assert np.array_equal(lo, sv.min(0)) and np.array_equal(hi, sv.max(0))
        # single points, host in and out
        for i in (0, 17, 4242):
            assert quest.signed_distance_evaluate(*q[i]) == want["dist"][i]
            d1, c1, n1 = quest.signed_distance_evaluate_cp(*q[i])
            assert d1 == want["dist"][i] and np.array_equal(c1, want["cp"][i])
            np.testing.assert_allclose(n1, want["normal"][i], rtol=0, atol=1e-12)
    finally:
        quest.signed_distance_finalize()
    assert not quest.signed_distance_initialized()


def test_compute_distance_single_point_matches_batch():
    v, t = _mesh()
    sdf = quest.SignedDistance(_t(v), _t(t))
    q = synth.uniform(8, 0, 3 * 64, -1.5, 1.5).reshape(-1, 3)
    n = q.shape[0]
    cp = torch.empty((n, 3), dtype=torch.float64, device=DEV)
    nrm = torch.empty((n, 3), dtype=torch.float64, device=DEV)
    d = sdf.computeDistances(n, _t(q), outClosestPts=cp, outNormals=nrm).cpu().numpy()
    cp, nrm = cp.cpu().numpy(), nrm.cpu().numpy()
    for i in range(n):
        di, ci, ni = sdf.computeDistance(q[i])
        assert di == d[i] and np.array_equal(ci, cp[i]) and np.array_equal(ni, nrm[i])


def test_interface_misuse_raises():
    v, t = _mesh()
    quest.signed_distance_finalize()
    with pytest.raises(RuntimeError):
        quest.signed_distance_evaluate(0.0, 0.0, 0.0)  # before init
    with pytest.raises(RuntimeError):
        quest.signed_distance_get_mesh_bounds()
    with pytest.raises(RuntimeError):
        quest.signed_distance_set_dimension(2)
    assert quest.signed_distance_init("/nonexistent/file.stl") == -1
    assert quest.signed_distance_init((_t(v), _t(t))) == 0
    try:
        for setter, arg in ((quest.signed_distance_set_closed_surface, False),
                            (quest.signed_distance_set_compute_signs, False),
                            (quest.signed_distance_set_verbose, True),
                            (quest.signed_distance_set_allocator, 0),
                            (quest.signed_distance_set_execution_space,
                             quest.SignedDistExec_GPU)):
            with pytest.raises(RuntimeError):
                setter(arg)
        assert quest.signed_distance_init((_t(v), _t(t))) == -1  # already initialized
        q = np.array([[0.1, 0.2, 0.3], [2.0, 0.0, 0.0]])
        want = orc.OracleSDF(v, t).eval(q)["dist"]
        assert [quest.signed_distance_evaluate(*p) for p in q] == want.tolist()
    finally:
        quest.signed_distance_finalize()


def _vtk_text(bounds, nodes, children, dims):
    """bvh_vtkio.hpp's output for a tree in the reference's layout: 2(n-1)
    boxes `nodes` (2*dims doubles each) and `children` (2*child or -(pos+1))."""
    pts, cells, levels = [], [], []

    def g(x):
        return "%g" % x

    def box(lo, hi):
        nn = 4 if dims == 2 else 8
        base = len(pts)
        if dims == 2:
            pts.extend([f"{g(lo[0])} {g(lo[1])} 0.0", f"{g(hi[0])} {g(lo[1])} 0.0",
                        f"{g(hi[0])} {g(hi[1])} 0.0", f"{g(lo[0])} {g(hi[1])} 0.0"])
        else:
            for zc in (lo[2], hi[2]):
                pts.extend([f"{g(lo[0])} {g(lo[1])} {g(zc)}", f"{g(hi[0])} {g(lo[1])} {g(zc)}",
                            f"{g(hi[0])} {g(hi[1])} {g(zc)}", f"{g(lo[0])} {g(hi[1])} {g(zc)}"])
        cells.append(f"{nn} " + "".join(f"{base + k} " for k in range(nn)))

    def rec(node, level):
        for side in (0, 1):
            b = nodes[node + side]
            box(b[:dims], b[dims:])
            levels.append(str(level))
        for side in (0, 1):
            c = int(children[node + side])
            if c > -1:
                rec(c, level + 1)

    box(bounds[:dims], bounds[dims:])
    levels.append("0")
    if len(children):
        rec(0, 1)
    nn = 4 if dims == 2 else 8
    out = ["# vtk DataFile Version 3.0", " BVHTree ", "ASCII", "DATASET UNSTRUCTURED_GRID",
           f"POINTS {len(pts)} double"] + pts + ["", f"CELLS {len(cells)} {len(cells) * (nn + 1)}"]
    out += cells + ["", f"CELL_TYPES {len(cells)}"] + [str(9 if dims == 2 else 12)] * len(cells)
    out += [f"CELL_DATA {len(cells)}", "SCALARS level int", "LOOKUP_TABLE default"] + levels
    return "\n".join(out) + "\n\n\n"


@pytest.mark.parametrize("dims", [3, 2])
def test_write_vtk_file(tmp_path, dims):
    rng = np.random.default_rng(dims)
    n = 300
    c = rng.random((n, dims))
    h = rng.uniform(0.001, 0.05, (n, dims))
    boxes = np.concatenate([c - h, c + h], 1)
    g = spin.BVH(_t(boxes), NDIMS=dims)
    path = tmp_path / "bvh.vtk"
    g.writeVtkFile(str(path))
    ex = orc.OracleBVH(boxes, ndims=dims).export()
    lo, hi = g.getBounds()
    want = _vtk_text(np.concatenate([lo, hi]), ex["nodes"], ex["children"], dims)
    assert path.read_text() == want
    assert np.array_equal(np.concatenate([lo, hi]), ex["bounds"])
