"""GPU parity of the triangle-triangle test and of the BVH mesh
self-intersection test against the oracle.

Bar: bit-exact booleans for every triangle pair; identical intersecting
pairs, in the same sequential order, and identical degenerate indices for
every mesh. Inputs: the reference's known answers
(tests/golden/tri_tri_known_answers.json), seeded random pairs (generic,
vertex- and edge-sharing, coplanar, near-touching), triangle soups and
noisy icospheres.
"""
import json
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from axom_amd import primal, quest, synth  # noqa: E402
from oracle import oracle as orc  # noqa: E402

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
KA = json.load(open(os.path.join(HERE, "golden", "tri_tri_known_answers.json")))
DEV = "cuda"


def _t(a, dtype=None):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=dtype), device=DEV)


def _perms(a, b):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    ap, bp = a[[0, 2, 1]], b[[0, 2, 1]]
    roll = lambda t, i: t[[i % 3, (i + 1) % 3, (i + 2) % 3]]
    return [(roll(x, i), roll(y, j)) for (x, y) in ((a, b), (ap, bp), (b, a), (bp, ap))
            for i in range(3) for j in range(3)]


def _gpu_vs_oracle(t1, t2, inc, eps):
    got = primal.intersect(_t(t1), _t(t2), inc, eps).cpu().numpy()
    want = np.array([orc.tri_tri(a, b, inc, eps) for a, b in zip(t1, t2)])
    return got, want


def test_tri_tri_known_answers_all_permutations():
    t1, t2, want = [], [], []
    for case in KA["tri_tri_3d"]:
        for inc, w in ((True, case["expect_include_boundary"]), (False, case["expect_exclude_boundary"])):
            for a, b in _perms(case["a"], case["b"]):
                t1.append((a, inc))
                t2.append(b)
                want.append(w)
    for case in KA["tri_tri_2d_in_plane"]:
        for plane in ((0, 1), (0, 2), (2, 1)):
            a = np.zeros((3, 3))
            b = np.zeros((3, 3))
            a[:, plane[0]], a[:, plane[1]] = np.array(case["a"])[:, 0], np.array(case["a"])[:, 1]
            b[:, plane[0]], b[:, plane[1]] = np.array(case["b"])[:, 0], np.array(case["b"])[:, 1]
            for inc, w in ((True, case["expect_include_boundary"]),
                           (False, case["expect_exclude_boundary"])):
                for x, y in _perms(a, b):
                    t1.append((x, inc))
                    t2.append(y)
                    want.append(w)
    for inc in (True, False):
        idx = [k for k, (_, i) in enumerate(t1) if i == inc]
        a = np.stack([t1[k][0] for k in idx])
        b = np.stack([t2[k] for k in idx])
        got = primal.intersect(_t(a), _t(b), inc, KA["eps"]).cpu().numpy()
        assert np.array_equal(got, np.array([want[k] for k in idx])), inc


def _random_pairs(n, seed):
    """Generic pairs, pairs sharing a vertex or an edge, coplanar pairs and
    pairs placed on the EPS boundary of touching."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1, 1, (n, 3, 3))
    b = rng.uniform(-1, 1, (n, 3, 3)) * rng.uniform(0.3, 1.2, (n, 1, 1)) + \
        rng.uniform(-0.5, 0.5, (n, 1, 3))
    kind = rng.integers(0, 6, n)
    share_v = kind == 1
    b[share_v, 0] = a[share_v, rng.integers(0, 3)]
    share_e = kind == 2
    b[share_e, 0], b[share_e, 1] = a[share_e, 1], a[share_e, 2]
    cop = kind == 3  # both in z = 0.25, then rotated together
    a[cop, :, 2] = 0.25
    b[cop, :, 2] = 0.25
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    a[cop] = a[cop] @ q
    b[cop] = b[cop] @ q
    touch = kind == 4  # a corner of b on a's plane (within a few ulps)
    n_a = np.cross(a[:, 1] - a[:, 0], a[:, 2] - a[:, 0])
    n_a /= np.linalg.norm(n_a, axis=1, keepdims=True)
    d = np.einsum("ij,ij->i", b[:, 0] - a[:, 2], n_a)
    b[touch, 0] -= (d[touch] * (1 + rng.uniform(-1e-9, 1e-9, touch.sum())))[:, None] * n_a[touch]
    small = kind == 5  # tiny triangles near each other: many near-zero products
    a[small] *= 1e-4
    b[small] = a[small] + rng.normal(0, 1e-4, (small.sum(), 3, 3))
    return a, b


@pytest.mark.parametrize("inc", [False, True])
def test_tri_tri_random_pairs_bit_exact(inc):
    a, b = _random_pairs(60_000, 5 + int(inc))
    for eps in (1e-8, 1e-12):
        got, want = _gpu_vs_oracle(a, b, inc, eps)
        assert np.array_equal(got, want), (eps, np.nonzero(got != want)[0][:5])
        assert 0.1 < want.mean() < 0.9


def _mesh_check(v, t, thresh=1e-8):
    pairs, degen = quest.findTriMeshIntersectionsBVH(_t(v), _t(t, np.int32), thresh)
    want_p, want_d = orc.mesh_self_intersections(v, t, thresh, nthreads=8)
    got_p = pairs.cpu().numpy().reshape(-1, 2)
    assert np.array_equal(got_p, want_p.reshape(-1, 2))
    assert np.array_equal(degen.cpu().numpy(), want_d)
    return got_p, degen.cpu().numpy()


@pytest.mark.parametrize("case", KA["mesh_tester"], ids=lambda c: c["name"])
def test_mesh_known_answers(case):
    pairs, degen = _mesh_check(np.array(case["verts"], float), np.array(case["tris"]))
    assert degen.tolist() == case["degenerate"]
    deg = set(case["degenerate"])
    got = sorted(tuple(p) for p in pairs.tolist() if not (deg & set(p)))
    assert got == sorted(map(tuple, case["pairs"]))


def test_mesh_triangle_soup_bit_exact():
    rng = np.random.default_rng(21)
    n = 40_000
    c = rng.uniform(-1, 1, (n, 3))
    v = (c[:, None, :] + rng.normal(0, 0.02, (n, 3, 3))).reshape(-1, 3)
    t = np.arange(3 * n, dtype=np.int32).reshape(n, 3)
    t[::97] = t[::97, [0, 0, 1]]  # some degenerate (repeated corner) triangles
    pairs, degen = _mesh_check(v, t)
    assert pairs.shape[0] > 1000 and degen.size == len(range(0, n, 97))


def test_mesh_icosphere_clean_and_noisy():
    v, t = synth.icosphere(5)
    pairs, degen = _mesh_check(v, t)
    assert pairs.shape[0] == 0 and degen.size == 0  # a closed convex surface
    rng = np.random.default_rng(4)
    vn = v + rng.normal(0, 0.02, v.shape)
    pairs, _ = _mesh_check(vn, t)
    assert pairs.shape[0] > 1000
    # two overlapping spheres: the pairs along their intersection curve
    v2 = np.concatenate([v, v + np.array([0.5, 0.1, 0.0])])
    t2 = np.concatenate([t, t + v.shape[0]]).astype(np.int32)
    pairs, _ = _mesh_check(v2, t2)
    assert pairs.shape[0] > 500 and np.all(pairs[:, 0] < t.shape[0]) and np.all(pairs[:, 1] >= t.shape[0])


def test_mesh_edge_cases():
    for n in (0, 1, 2):
        rng = np.random.default_rng(n)
        v = rng.uniform(-1, 1, (3 * max(n, 1), 3))
        t = np.arange(3 * n, dtype=np.int32).reshape(n, 3)
        pairs, degen = quest.findTriMeshIntersectionsBVH(_t(v), _t(t, np.int32))
        want_p, want_d = orc.mesh_self_intersections(v, t)
        assert np.array_equal(pairs.cpu().numpy().reshape(-1, 2), want_p.reshape(-1, 2))
        assert np.array_equal(degen.cpu().numpy(), want_d)
    v = np.zeros((3, 3))
    with pytest.raises(RuntimeError):
        quest.findTriMeshIntersectionsBVH(_t(v), _t(np.array([[0, 1, 3]], np.int32)))
