"""ctypes front end of the CPU oracle (oracle/lib/liborc.so).

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, as the checker or the timed CPU
baseline. The product path (axom_amd) never imports it.

The C library restates the reference's sequential path; see axsq_oracle.c
for the file:line each function follows.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "liborc.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_f32p = ctypes.POINTER(ctypes.c_float)


def build():
    """Compile liborc.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_bvh_new.restype = ctypes.c_void_p
        L.orc_bvh_new.argtypes = [_dp, ctypes.c_int32, ctypes.c_double]
        L.orc_bvh_new2d.restype = ctypes.c_void_p
        L.orc_bvh_new2d.argtypes = [_dp, ctypes.c_int32, ctypes.c_double]
        L.orc_bvh_new_f32.restype = ctypes.c_void_p
        L.orc_bvh_new_f32.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_int32,
                                      ctypes.c_int, ctypes.c_float]
        L.orc_ray_box_f32.restype = ctypes.c_int
        L.orc_ray_box_f32.argtypes = [_dp, _dp, _dp, ctypes.c_float]
        L.orc_morton2.restype = ctypes.c_uint32
        L.orc_morton2.argtypes = [ctypes.c_int32] * 2
        L.orc_bvh_free.argtypes = [ctypes.c_void_p]
        L.orc_bvh_size.restype = ctypes.c_int32
        L.orc_bvh_size.argtypes = [ctypes.c_void_p]
        L.orc_bvh_export.argtypes = [ctypes.c_void_p, _dp, _u32p, _i32p, _i32p, _i32p,
                                     _i32p, _dp, _dp, _i32p]
        L.orc_bvh_find.restype = ctypes.c_int64
        L.orc_bvh_find.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_int64,
                                   ctypes.c_double, _i32p, _i32p, _i32p, ctypes.c_int]
        L.orc_bvh_max_depth.restype = ctypes.c_int32
        L.orc_bvh_max_depth.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_int64,
                                        ctypes.c_double]
        L.orc_bvh_visits.restype = ctypes.c_int64
        L.orc_bvh_visits.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_int64,
                                     ctypes.c_double]
        L.orc_sdf_max_depth.restype = ctypes.c_int32
        L.orc_sdf_max_depth.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int64]
        L.orc_bvh_brute.restype = ctypes.c_int64
        L.orc_bvh_brute.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_int64,
                                    ctypes.c_double, _i32p, _i32p, ctypes.c_int64]
        L.orc_sdf_new.restype = ctypes.c_void_p
        L.orc_sdf_new.argtypes = [_dp, _dp, _dp, ctypes.c_int32, _i32p, ctypes.c_int32,
                                  ctypes.c_int, ctypes.c_int]
        L.orc_sdf_free.argtypes = [ctypes.c_void_p]
        L.orc_sdf_new_cells.restype = ctypes.c_void_p
        L.orc_sdf_new_cells.argtypes = [_dp, _dp, _dp, ctypes.c_int32, _i32p, _i32p,
                                        ctypes.c_int32, ctypes.c_int, ctypes.c_int]
        L.orc_sdf_bvh.restype = ctypes.c_void_p
        L.orc_sdf_bvh.argtypes = [ctypes.c_void_p]
        L.orc_sdf_bounds.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_sdf_eval.restype = ctypes.c_int
        L.orc_sdf_eval.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int64, _dp, _dp, _dp,
                                   _i32p, _i32p, _i32p, _i32p, ctypes.c_int]
        L.orc_sdf_brute.restype = ctypes.c_int
        L.orc_sdf_brute.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int64, _dp, _i32p,
                                    ctypes.c_int]
        L.orc_sdf_backward.restype = ctypes.c_int
        L.orc_sdf_backward.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int64, _dp, _i32p,
                                       _f32p, _dp, _dp]
        L.orc_morton3.restype = ctypes.c_uint32
        L.orc_morton3.argtypes = [ctypes.c_int32] * 3
        L.orc_closest_point.restype = ctypes.c_int
        L.orc_closest_point.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_double, _dp]
        L.orc_sqdist_point_box.restype = ctypes.c_double
        L.orc_sqdist_point_box.argtypes = [_dp, _dp]
        L.orc_ray_box.restype = ctypes.c_int
        L.orc_ray_box.argtypes = [_dp, _dp, _dp, ctypes.c_double]
        L.orc_box_scale.argtypes = [_dp, ctypes.c_double]
        L.orc_unit_vector.argtypes = [_dp, _dp]
        L.orc_tri_tri.restype = ctypes.c_int
        L.orc_tri_tri.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_double]
        L.orc_mesh_self_isect.restype = ctypes.c_int64
        L.orc_mesh_self_isect.argtypes = [_dp, _dp, _dp, ctypes.c_int32, _i32p, ctypes.c_int32,
                                          ctypes.c_double, _i32p, _i32p, ctypes.c_int64,
                                          ctypes.POINTER(ctypes.c_uint8), ctypes.c_int]
        _lib = L
    return _lib


def _p(a, t):
    if a is None:
        return None
    return a.ctypes.data_as(t)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ---------------------------------------------------------------- primal ----
def morton3(x, y, z):
    return int(lib().orc_morton3(int(x), int(y), int(z)))


def morton2(x, y):
    return int(lib().orc_morton2(int(x), int(y)))


def closest_point(p, a, b, c, eps=1e-12):
    """Returns (cp, loc) following primal::closest_point."""
    cp = np.zeros(3)
    loc = lib().orc_closest_point(_p(f64(p), _dp), _p(f64(a), _dp), _p(f64(b), _dp),
                                  _p(f64(c), _dp), eps, _p(cp, _dp))
    return cp, loc


def sqdist_point_box(p, box6):
    return lib().orc_sqdist_point_box(_p(f64(p), _dp), _p(f64(box6), _dp))


def ray_box(o, d, box6, eps):
    return bool(lib().orc_ray_box(_p(f64(o), _dp), _p(f64(d), _dp), _p(f64(box6), _dp), eps))


def box_scale(box6, sf):
    b = f64(box6).copy()
    lib().orc_box_scale(_p(b, _dp), sf)
    return b


def unit_vector(v):
    out = np.zeros(3)
    lib().orc_unit_vector(_p(f64(v), _dp), _p(out, _dp))
    return out


def tri_tri(t1, t2, include_boundary=False, eps=1e-8):
    """primal::intersect(Triangle3, Triangle3, includeBoundary, EPS); t1, t2 are
    (3, 3) corner arrays."""
    return bool(lib().orc_tri_tri(_p(f64(t1).reshape(9), _dp), _p(f64(t2).reshape(9), _dp),
                                  int(bool(include_boundary)), float(eps)))


def mesh_self_intersections(verts, tris, threshold=1e-8, nthreads=1):
    """quest::findTriMeshIntersectionsBVH: returns (pairs (k, 2) int32 in the
    sequential order, degenerate triangle indices)."""
    v = f64(verts)
    t = i32(tris)
    x, y, z = (np.ascontiguousarray(v[:, k]) for k in range(3))
    nt = t.shape[0]
    degen = np.zeros(max(nt, 1), dtype=np.uint8)
    cap = 0
    while True:
        first = np.zeros(max(cap, 1), dtype=np.int32)
        second = np.zeros(max(cap, 1), dtype=np.int32)
        n = lib().orc_mesh_self_isect(_p(x, _dp), _p(y, _dp), _p(z, _dp), v.shape[0],
                                      _p(t.reshape(-1), _i32p), nt, float(threshold),
                                      _p(first, _i32p), _p(second, _i32p), cap,
                                      degen.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                      nthreads)
        if n < 0:
            raise RuntimeError("oracle mesh self-intersection failed")
        if n <= cap:
            break
        cap = n
    pairs = np.stack([first[:n], second[:n]], axis=1) if n else np.zeros((0, 2), np.int32)
    return pairs, np.nonzero(degen[:nt])[0].astype(np.int32)


# ------------------------------------------------------------------- BVH ----
KIND = {"points": 0, "boxes": 1, "rays": 2}


class OracleBVH:
    """spin::BVH<NDIMS, SEQ_EXEC, FloatType> restated (build + find*); ndims 3
    or 2, dtype float64 (double) or float32 (float). Boxes are AoS (min, max)
    of 2 * ndims values; 2-D queries are points of 2 values and boxes / rays
    of 4. A float tree's queries are float32 values (passed on as doubles,
    exactly), its tolerance FLT_EPSILON by default."""

    def __init__(self, boxes, scale=1.000123, _handle=None, _owner=None, ndims=3,
                 dtype=np.float64):
        self.ndims = ndims
        self.dtype = np.dtype(dtype)
        if _handle is not None:
            self.h = _handle
            self._owner = _owner
        elif self.dtype == np.float32:
            b = np.ascontiguousarray(boxes, np.float32).reshape(-1, 2 * ndims)
            self._boxes = b
            self.h = lib().orc_bvh_new_f32(b.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                           b.shape[0], ndims, np.float32(scale))
            self._owner = None
        else:
            b = f64(boxes).reshape(-1, 2 * ndims)
            self._boxes = b
            new = lib().orc_bvh_new2d if ndims == 2 else lib().orc_bvh_new
            self.h = new(_p(b, _dp), b.shape[0], scale)
            self._owner = None

    def _tol(self, tol):
        return float(np.finfo(self.dtype).eps) if tol is None else float(tol)

    def __del__(self):
        if getattr(self, "_owner", 1) is None and getattr(self, "h", None):
            lib().orc_bvh_free(self.h)
            self.h = None

    @property
    def n(self):
        return lib().orc_bvh_size(self.h)

    def export(self):
        n = self.n
        inner = n - 1
        out = dict(
            bounds=np.zeros(6), mcodes=np.zeros(n, np.uint32), leafs=np.zeros(n, np.int32),
            lchild=np.zeros(inner, np.int32), rchild=np.zeros(inner, np.int32),
            parents=np.zeros(2 * n - 1, np.int32), inner_aabbs=np.zeros((inner, 6)),
            nodes=np.zeros((2 * inner, 6)), children=np.zeros(2 * inner, np.int32))
        lib().orc_bvh_export(self.h, _p(out["bounds"], _dp), _p(out["mcodes"], _u32p),
                             _p(out["leafs"], _i32p), _p(out["lchild"], _i32p),
                             _p(out["rchild"], _i32p), _p(out["parents"], _i32p),
                             _p(out["inner_aabbs"], _dp), _p(out["nodes"], _dp),
                             _p(out["children"], _i32p))
        if self.ndims == 2:  # BoundingBox<double, 2>: min x y, max x y
            for k in ("bounds", "inner_aabbs", "nodes"):
                out[k] = np.ascontiguousarray(out[k][..., [0, 1, 3, 4]])
        return out

    def _stride(self, k):
        return self.ndims if k == 0 else 2 * self.ndims

    def find(self, kind, prims, tol=None, nthreads=1):
        """Returns (offsets, counts, candidates) like LinearBVH::findCandidatesImpl."""
        tol = self._tol(tol)
        k = KIND[kind]
        pr = f64(prims).reshape(-1, self._stride(k))
        nq = pr.shape[0]
        counts = np.zeros(nq, np.int32)
        offsets = np.zeros(nq, np.int32)
        total = lib().orc_bvh_find(self.h, k, _p(pr, _dp), nq, tol, _p(counts, _i32p),
                                   _p(offsets, _i32p), None, nthreads)
        if total < 0:
            raise RuntimeError("oracle traversal stack overflow")
        cands = np.zeros(max(total, 1), np.int32)
        lib().orc_bvh_find(self.h, k, _p(pr, _dp), nq, tol, _p(counts, _i32p),
                           _p(offsets, _i32p), _p(cands, _i32p), nthreads)
        return offsets, counts, cands[:total]

    def max_depth(self, kind, prims, tol=None):
        """The deepest stack entry the queries' walks push (-1: overflow)."""
        tol = self._tol(tol)
        k = KIND[kind]
        pr = f64(prims).reshape(-1, self._stride(k))
        return int(lib().orc_bvh_max_depth(self.h, k, _p(pr, _dp), pr.shape[0], tol))

    def visits(self, kind, prims, tol=None):
        """Inner-node visits of the queries' walks, summed."""
        tol = self._tol(tol)
        k = KIND[kind]
        pr = f64(prims).reshape(-1, self._stride(k))
        v = int(lib().orc_bvh_visits(self.h, k, _p(pr, _dp), pr.shape[0], tol))
        if v < 0:
            raise RuntimeError("oracle traversal stack overflow")
        return v

    def brute(self, kind, prims, tol=None):
        """O(N*Q) brute force over the scaled leaf boxes: (counts, ids)."""
        tol = self._tol(tol)
        k = KIND[kind]
        pr = f64(prims).reshape(-1, self._stride(k))
        nq = pr.shape[0]
        counts = np.zeros(nq, np.int32)
        total = lib().orc_bvh_brute(self.h, k, _p(pr, _dp), nq, tol, _p(counts, _i32p), None, 0)
        cands = np.zeros(max(total, 1), np.int32)
        lib().orc_bvh_brute(self.h, k, _p(pr, _dp), nq, tol, _p(counts, _i32p),
                            _p(cands, _i32p), total)
        return counts, cands[:total]


# -------------------------------------------------------- SignedDistance ----
class OracleSDF:
    """quest::SignedDistance<3, SEQ_EXEC> restated, plus the adjoint."""

    def __init__(self, verts, tris, watertight=True, compute_sign=True, cell_offsets=None):
        """tris: (T, 3) triangles; or, with cell_offsets (ncells + 1), the flat
        connectivity of cells of 3 or 4 nodes (triangles and quads)."""
        v = f64(verts).reshape(-1, 3)
        self.nverts = v.shape[0]
        x, y, z = f64(v[:, 0]), f64(v[:, 1]), f64(v[:, 2])
        self._xyz = (x, y, z)
        if cell_offsets is None:
            t = i32(tris).reshape(-1, 3)
            self.ntris = t.shape[0]
            self.h = lib().orc_sdf_new(_p(x, _dp), _p(y, _dp), _p(z, _dp), self.nverts,
                                       _p(t, _i32p), self.ntris, int(watertight),
                                       int(compute_sign))
        else:
            t = i32(tris).reshape(-1)
            off = i32(cell_offsets)
            self.ntris = off.shape[0] - 1
            self.h = lib().orc_sdf_new_cells(_p(x, _dp), _p(y, _dp), _p(z, _dp), self.nverts,
                                             _p(t, _i32p), _p(off, _i32p), self.ntris,
                                             int(watertight), int(compute_sign))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_sdf_free(self.h)
            self.h = None

    def bvh(self):
        return OracleBVH(None, _handle=lib().orc_sdf_bvh(self.h), _owner=self)

    def bounds(self):
        lo, hi = np.zeros(3), np.zeros(3)
        lib().orc_sdf_bounds(self.h, _p(lo, _dp), _p(hi, _dp))
        return lo, hi

    def eval(self, q, nthreads=1, want_cp=True, want_normal=True, want_stats=False):
        q = f64(q).reshape(-1, 3)
        nq = q.shape[0]
        dist = np.zeros(nq)
        cp = np.zeros((nq, 3)) if want_cp else None
        nrm = np.zeros((nq, 3)) if want_normal else None
        elem = np.zeros(nq, np.int32)
        loc = np.zeros(nq, np.int32)
        visits = np.zeros(nq, np.int32) if want_stats else None
        leaves = np.zeros(nq, np.int32) if want_stats else None
        rc = lib().orc_sdf_eval(self.h, _p(q, _dp), nq, _p(dist, _dp), _p(cp, _dp),
                                _p(nrm, _dp), _p(elem, _i32p), _p(loc, _i32p),
                                _p(visits, _i32p), _p(leaves, _i32p), nthreads)
        if rc != 0:
            raise RuntimeError("oracle traversal stack overflow")
        out = dict(dist=dist, cp=cp, normal=nrm, elem=elem, loc=loc)
        if want_stats:
            out["visits"] = visits
            out["leaves"] = leaves
        return out

    def max_depth(self, q):
        """The deepest stack entry the queries' walks push (-1: overflow)."""
        q = f64(q).reshape(-1, 3)
        return int(lib().orc_sdf_max_depth(self.h, _p(q, _dp), q.shape[0]))

    def brute(self, q, nthreads=1):
        q = f64(q).reshape(-1, 3)
        nq = q.shape[0]
        sq = np.zeros(nq)
        elem = np.zeros(nq, np.int32)
        lib().orc_sdf_brute(self.h, _p(q, _dp), nq, _p(sq, _dp), _p(elem, _i32p), nthreads)
        return sq, elem

    def backward(self, q, dist, elem, gout):
        q = f64(q).reshape(-1, 3)
        nq = q.shape[0]
        gq = np.zeros((nq, 3))
        gv = np.zeros((self.nverts, 3))
        rc = lib().orc_sdf_backward(self.h, _p(q, _dp), nq, _p(f64(dist), _dp),
                                    _p(i32(elem), _i32p),
                                    _p(np.ascontiguousarray(gout, np.float32), _f32p),
                                    _p(gq, _dp), _p(gv, _dp))
        if rc != 0:
            raise ValueError("adjoint: a cell mesh, or an element id outside [-1, ntris)")
        return gq, gv


# ------------------------------------------------- hex candidate search ----
def cell_boxes(vertices, cells):
    """primal::compute_bounding_box of each cell's nodes (BoundingBox.hpp:
    466-482: a coordinate replaces the bound it is strictly below / above),
    AoS (min xyz, max xyz)."""
    v = f64(vertices).reshape(-1, 3)
    c = np.asarray(cells, dtype=np.int64).reshape(len(cells), -1)
    big = np.finfo(np.float64).max
    lo = np.full((c.shape[0], 3), big)
    hi = np.full((c.shape[0], 3), -big)
    for k in range(c.shape[1]):
        p = v[c[:, k]]
        lo = np.where(p < lo, p, lo)
        hi = np.where(p > hi, p, hi)
    return np.concatenate([lo, hi], 1)


def mesh_candidates(insert_vertices, insert_cells, query_vertices, query_cells, nthreads=1):
    """findCandidatesBVH (quest/examples/quest_candidates_example.cpp:397-495):
    (query cell, insert cell) pairs from findBoundingBoxes over a BVH of the
    insert cells' boxes (default scale), in candidate-array order."""
    ib = cell_boxes(insert_vertices, insert_cells)
    qb = cell_boxes(query_vertices, query_cells)
    off, cnt, cand = OracleBVH(ib).find("boxes", qb, nthreads=nthreads)
    first = np.repeat(np.arange(qb.shape[0], dtype=np.int32), cnt)
    return np.stack([first, cand.astype(np.int32)], 1)
