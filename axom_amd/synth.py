"""Seeded synthetic inputs for the spatial-query workloads.

All randomness comes from a counter-based splitmix64 stream, so a given
(seed, stream, index) always yields the same double on every machine.

Workloads (seed = 20251202 + workload number):
  1  100,000 boxes, centres U[0,1]^3, half-extents U[0.001,0.01];
     1,000,000 points U[0,1]^3.
  2  icosphere level 6 (40,962 verts, 81,920 tris), r = 1, each vertex scaled
     by 1 + U[-1e-3, 1e-3]; 4,000,000 queries U[-1.5,1.5]^3.
  3  torus R = 1, tube radius 0.3 + 0.02 sin(7u) cos(5v), 1024 x 1024 vertex
     grid (2,097,152 tris); 16M queries: half U[bbox padded by 0.2], half a
     surface sample (uniform triangle, uniform barycentric) moved along the
     unit face normal by U(-0.01, 0.01).
  4  icosphere level 7, radius 1 + 0.1 sin(4 theta) cos(3 phi) with
     theta = atan2(y, x), phi = acos(z) of the unit vertex; 1M surface samples
     moved along the face normal by U[-0.1, 0.1]; upstream gradient N(0,1) f32.
  5  icosphere level 8, r = 1; target = rounded box b = (0.8,0.6,0.5),
     r = 0.1: q = |p| - b + r, d = |max(q,0)| + min(max(qx,qy,qz),0) - r;
     8M queries U[-1.2,1.2]^3.
"""
import math

import numpy as np

SEED0 = 20251202

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z):
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def splitmix64(seed, stream, n):
    """n uint64 values of stream `stream` under `seed` (counter based)."""
    with np.errstate(over="ignore"):
        key = _mix(np.array([np.uint64(seed) * _GOLD + np.uint64(stream) + _GOLD],
                            dtype=np.uint64))[0]
        idx = np.arange(1, n + 1, dtype=np.uint64)
        return _mix(key + idx * _GOLD)


def uniform(seed, stream, n, lo=0.0, hi=1.0):
    """n doubles U[lo, hi) from 53 random bits each."""
    u = (splitmix64(seed, stream, n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    return lo + (hi - lo) * u


def normal_f32(seed, stream, n):
    """N(0,1) via Box-Muller on two uniform streams, rounded to float32."""
    u1 = uniform(seed, stream, n)
    u2 = uniform(seed, stream + 1, n)
    r = np.sqrt(-2.0 * np.log1p(-u1))
    return (r * np.cos(2.0 * math.pi * u2)).astype(np.float32)


# ----------------------------------------------------------------- meshes ----
def icosphere(level):
    """Unit icosphere; returns (verts (V,3) f64, tris (T,3) i32), outward CCW."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1)[:, None]
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    for _ in range(level):
        nv = v.shape[0]
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        key = e[:, 0] * (nv + 1) + e[:, 1]
        uk, inv = np.unique(key, return_inverse=True)
        a = uk // (nv + 1)
        b = uk % (nv + 1)
        mid = 0.5 * (v[a] + v[b])
        mid /= np.linalg.norm(mid, axis=1)[:, None]
        v = np.concatenate([v, mid])
        nf = f.shape[0]
        ab = nv + inv[:nf]
        bc = nv + inv[nf:2 * nf]
        ca = nv + inv[2 * nf:]
        f = np.concatenate([np.stack([f[:, 0], ab, ca], 1), np.stack([f[:, 1], bc, ab], 1),
                            np.stack([f[:, 2], ca, bc], 1), np.stack([ab, bc, ca], 1)])
    return np.ascontiguousarray(v), np.ascontiguousarray(f.astype(np.int32))


def torus(nu=1024, nv=1024, R=1.0, r0=0.3, amp=0.02):
    """Displaced torus on an nu x nv vertex grid, 2*nu*nv outward triangles."""
    u = 2.0 * math.pi * np.arange(nu) / nu
    w = 2.0 * math.pi * np.arange(nv) / nv
    U, W = np.meshgrid(u, w, indexing="ij")
    r = r0 + amp * np.sin(7.0 * U) * np.cos(5.0 * W)
    x = (R + r * np.cos(W)) * np.cos(U)
    y = (R + r * np.cos(W)) * np.sin(U)
    z = r * np.sin(W)
    verts = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
    i = np.arange(nu)[:, None]
    j = np.arange(nv)[None, :]
    i1 = (i + 1) % nu
    j1 = (j + 1) % nv
    a = (i * nv + j).ravel()
    b = (i1 * nv + j).ravel()
    c = (i1 * nv + j1).ravel()
    d = (i * nv + j1).ravel()
    tris = np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)])
    return np.ascontiguousarray(verts), np.ascontiguousarray(tris.astype(np.int32))


def face_normals(verts, tris):
    a, b, c = verts[tris[:, 0]], verts[tris[:, 1]], verts[tris[:, 2]]
    n = np.cross(b - a, c - a)
    return n / np.linalg.norm(n, axis=1)[:, None]


def surface_samples(verts, tris, seed, stream, n, offset_lo, offset_hi):
    """Uniform triangle, uniform barycentric, moved along the unit face normal."""
    t = np.minimum((uniform(seed, stream, n) * tris.shape[0]).astype(np.int64),
                   tris.shape[0] - 1)
    r1 = uniform(seed, stream + 1, n)
    r2 = uniform(seed, stream + 2, n)
    flip = r1 + r2 > 1.0
    r1 = np.where(flip, 1.0 - r1, r1)
    r2 = np.where(flip, 1.0 - r2, r2)
    a, b, c = verts[tris[t, 0]], verts[tris[t, 1]], verts[tris[t, 2]]
    p = a + r1[:, None] * (b - a) + r2[:, None] * (c - a)
    n = face_normals(verts, tris[t])
    off = uniform(seed, stream + 3, n.shape[0], offset_lo, offset_hi)
    return np.ascontiguousarray(p + off[:, None] * n)


def rounded_box_sdf(p, b=(0.8, 0.6, 0.5), r=0.1):
    q = np.abs(p) - np.asarray(b) + r
    outside = np.linalg.norm(np.maximum(q, 0.0), axis=1)
    inside = np.minimum(np.max(q, axis=1), 0.0)
    return outside + inside - r


# -------------------------------------------------------------- workloads ----
def workload1(scale=1.0):
    """Boxes (n,6 AoS min,max) and query points for the findPoints workload."""
    s = SEED0 + 1
    n = max(2, int(100_000 * scale))
    q = max(1, int(1_000_000 * scale))
    c = uniform(s, 0, 3 * n).reshape(n, 3)
    h = uniform(s, 10, 3 * n, 0.001, 0.01).reshape(n, 3)
    boxes = np.concatenate([c - h, c + h], 1)
    pts = uniform(s, 20, 3 * q).reshape(q, 3)
    return np.ascontiguousarray(boxes), np.ascontiguousarray(pts)


def workload2(nq=4_000_000, level=6):
    s = SEED0 + 2
    v, t = icosphere(level)
    v = v * (1.0 + uniform(s, 0, v.shape[0], -1e-3, 1e-3))[:, None]
    q = uniform(s, 10, 3 * nq, -1.5, 1.5).reshape(nq, 3)
    return np.ascontiguousarray(v), t, np.ascontiguousarray(q)


def workload3(nq=16_000_000, nu=1024, nv=1024):
    s = SEED0 + 3
    v, t = torus(nu, nv)
    lo = v.min(0) - 0.2
    hi = v.max(0) + 0.2
    h = nq // 2
    qa = np.stack([uniform(s, 10 + d, h, lo[d], hi[d]) for d in range(3)], 1)
    qb = surface_samples(v, t, s, 20, nq - h, -0.01, 0.01)
    return v, t, np.ascontiguousarray(np.concatenate([qa, qb]))


def workload4(nq=1_000_000, level=7):
    s = SEED0 + 4
    v, t = icosphere(level)
    theta = np.arctan2(v[:, 1], v[:, 0])
    phi = np.arccos(np.clip(v[:, 2], -1.0, 1.0))
    v = v * (1.0 + 0.1 * np.sin(4.0 * theta) * np.cos(3.0 * phi))[:, None]
    q = surface_samples(v, t, s, 10, nq, -0.1, 0.1)
    g = normal_f32(s, 30, nq)
    return np.ascontiguousarray(v), t, q, g


def workload5(nq=8_000_000, level=8):
    s = SEED0 + 5
    v, t = icosphere(level)
    q = uniform(s, 10, 3 * nq, -1.2, 1.2).reshape(nq, 3)
    return v, t, np.ascontiguousarray(q), rounded_box_sdf(q)


def workload6(level=8):
    """Triangle mesh for the self-intersection workload: an icosphere whose
    vertices carry N(0, (0.25 * mean edge)^2) noise per coordinate, plus a
    second copy shifted by (0.5, 0.1, 0) so that the two surfaces cross."""
    s = SEED0 + 6
    v, t = icosphere(level)
    e = float(np.mean(np.linalg.norm(v[t[:, 0]] - v[t[:, 1]], axis=1)))
    v = v + 0.25 * e * normal_f32(s, 0, v.size).astype(np.float64).reshape(-1, 3)
    v2 = np.concatenate([v, v + np.array([0.5, 0.1, 0.0])])
    t2 = np.concatenate([t, t + v.shape[0]]).astype(np.int32)
    return np.ascontiguousarray(v2), np.ascontiguousarray(t2)


def cube_sphere(n, radius=1.0, project=True):
    """A closed quad mesh: the surface of the cube [-1,1]^3 as 6 n x n grids of
    quads (outward, counter-clockwise seen from outside), its vertices
    projected to the sphere of the given radius when `project`. Returns
    (verts (V,3), quads (6 n^2, 4) int32)."""
    ids = {}
    verts = []
    quads = []

    def vid(p):
        key = tuple(np.round(p, 12))
        if key not in ids:
            ids[key] = len(verts)
            verts.append(p)
        return ids[key]

    g = np.linspace(-1.0, 1.0, n + 1)
    for axis in range(3):
        for side in (-1.0, 1.0):
            u_ax, v_ax = [a for a in range(3) if a != axis]
            for i in range(n):
                for j in range(n):
                    corners = []
                    for (a, b) in ((i, j), (i + 1, j), (i + 1, j + 1), (i, j + 1)):
                        p = np.zeros(3)
                        p[axis] = side
                        p[u_ax] = g[a]
                        p[v_ax] = g[b]
                        corners.append(vid(p))
                    # outward orientation: (u x v) points along +axis for this order
                    nrm = np.zeros(3)
                    nrm[axis] = 1.0
                    uv = np.cross(np.eye(3)[u_ax], np.eye(3)[v_ax])
                    if np.dot(uv, nrm) * side < 0:
                        corners = corners[::-1]
                    quads.append(corners)
    v = np.array(verts, dtype=np.float64)
    if project:
        v = radius * v / np.linalg.norm(v, axis=1, keepdims=True)
    return np.ascontiguousarray(v), np.array(quads, dtype=np.int32)


def hex_grid(n, origin=(0.0, 0.0, 0.0), h=1.0, jitter=0.0, seed=0, angle=0.0):
    """A structured mesh of n^3 hexahedra (nodes in the VTK hexahedron order:
    bottom face counter-clockwise, then the top face), spacing h from origin,
    each vertex moved by U[-jitter, jitter] per axis, the whole grid turned by
    `angle` radians about the z axis through its centre. Returns
    (vertices (V,3) f64, hexes (n^3, 8) i32)."""
    m = n + 1
    i, j, k = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    v = np.stack([i, j, k], -1).reshape(-1, 3).astype(np.float64) * h
    if jitter > 0.0:
        v += np.random.default_rng(seed).uniform(-jitter, jitter, v.shape)
    if angle != 0.0:
        c = 0.5 * n * h
        x, y = v[:, 0] - c, v[:, 1] - c
        ca, sa = np.cos(angle), np.sin(angle)
        v[:, 0], v[:, 1] = c + ca * x - sa * y, c + sa * x + ca * y
    v += np.asarray(origin, dtype=np.float64)
    a, b, d = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    base = ((a * m + b) * m + d).reshape(-1)
    di, dj, dk = m * m, m, 1
    hexes = np.stack([base, base + di, base + di + dj, base + dj,
                      base + dk, base + di + dk, base + di + dj + dk, base + dj + dk], 1)
    return np.ascontiguousarray(v), hexes.astype(np.int32)
