"""Host mirror of axom::quest::SignedDistance<3, MI355X_EXEC> and of the
quest C-style signed-distance interface, over libaxsq.

SignedDistance follows src/axom/quest/SignedDistance.hpp:147-397 (ctor,
computeDistances, computeDistance); the module-level signed_distance_*
functions follow src/axom/quest/interface/signed_distance.hpp:88-319 (options
must be set before init; evaluate before init is an error). Misuse raises
RuntimeError where the reference aborts through SLIC_ERROR.
"""
import ctypes

import numpy as np
import torch

from . import _lib as L
from .spin import FP_MODES, _dev, _ptr, point_args, stream_ptr


def _soa(vertices):
    if isinstance(vertices, (tuple, list)):
        return tuple(_dev(c, torch.float64, "vertices") for c in vertices)
    v = _dev(vertices, torch.float64, "vertices", 3)
    return tuple(v[:, d].contiguous() for d in range(3))


class SignedDistance:
    """quest::SignedDistance<3, MI355X_EXEC> (SignedDistance.hpp:147-397).

    vertices: (V,3) float64 CUDA tensor or (x, y, z) SoA tuple (mint node
    coordinates); triangles: (T,3) int32 CUDA tensor (cell connectivity).
    fp_mode "fp32" additionally builds FP32 traversal nodes so that
    computeDistances(..., fp_mode="fp32") can run the mixed form.
    """

    def __init__(self, vertices, triangles, isWatertight=True, computeSign=True,
                 fp_mode="fp64", cell_offsets=None):
        """cell_offsets: for a mesh of triangle and quad cells, the (ncells + 1,)
        int32 offsets into the flat int32 connectivity `triangles` (a mint
        MIXED_SHAPE layout; a QUAD mesh has offsets 0, 4, 8, ...)."""
        self._h = None
        x, y, z = _soa(vertices)
        self.nverts = x.shape[0]
        self._fp = FP_MODES[fp_mode]
        h = ctypes.c_void_p()
        if cell_offsets is None:
            t = _dev(triangles, torch.int32, "triangles", 3)
            self.ntris = t.shape[0]
            L.check(L.lib().axsq_sdf_create(ctypes.byref(h), _ptr(x), _ptr(y), _ptr(z),
                                            self.nverts, _ptr(t), self.ntris,
                                            int(bool(isWatertight)), int(bool(computeSign)),
                                            self._fp, stream_ptr()))
        else:
            t = _dev(triangles, torch.int32, "connectivity").reshape(-1)
            off = _dev(cell_offsets, torch.int32, "cell_offsets").reshape(-1)
            self.ntris = off.shape[0] - 1
            L.check(L.lib().axsq_sdf_create_cells(ctypes.byref(h), _ptr(x), _ptr(y), _ptr(z),
                                                  self.nverts, _ptr(t), _ptr(off), self.ntris,
                                                  int(bool(isWatertight)),
                                                  int(bool(computeSign)), self._fp,
                                                  stream_ptr()))
        self._h = h

    @classmethod
    def fromSTL(cls, path, isWatertight=True, computeSign=True, fp_mode="fp64"):
        """The query over the triangle soup of an STL file, read as
        quest::STLReader reads it (axsq_sdf_create_from_stl)."""
        self = cls.__new__(cls)
        self._h = None
        self._fp = FP_MODES[fp_mode]
        h = ctypes.c_void_p()
        L.check(L.lib().axsq_sdf_create_from_stl(ctypes.byref(h), str(path).encode(),
                                                 int(bool(isWatertight)), int(bool(computeSign)),
                                                 self._fp, stream_ptr()))
        self._h = h
        nodes, nn, nf = ctypes.c_void_p(), ctypes.c_int64(), ctypes.c_int64()
        L.check(L.lib().axsq_stl_read(str(path).encode(), ctypes.byref(nodes), ctypes.byref(nn),
                                      ctypes.byref(nf)))
        L.lib().axsq_host_free(nodes)
        self.nverts, self.ntris = nn.value, nf.value
        return self

    def __del__(self):
        # at interpreter exit the module globals may already be gone; the
        # process then releases the device memory itself
        if getattr(self, "_h", None) and L is not None:
            L.lib().axsq_sdf_destroy(self._h)
            self._h = None

    def getMeshBounds(self):
        lo, hi = L._dbl3(), L._dbl3()
        L.check(L.lib().axsq_sdf_mesh_bounds(self._h, lo, hi))
        return np.array(lo[:]), np.array(hi[:])

    def updateVertices(self, vertices):
        """New coordinates for the same connectivity; rebuilds the BVH."""
        x, y, z = _soa(vertices)
        if x.shape[0] != self.nverts:
            raise ValueError("vertex count changed")
        L.check(L.lib().axsq_sdf_update_vertices(self._h, _ptr(x), _ptr(y), _ptr(z),
                                                 stream_ptr()))

    # -- SignedDistance.hpp:529-605
    def computeDistances(self, npts, queryPts, outSgnDist=None, outClosestPts=None,
                         outNormals=None, outElem=None, fp_mode=None):
        x, y, z, stride, n, keep = point_args(queryPts)
        npts = int(npts)
        if npts > n:
            raise ValueError("npts exceeds the number of query points")
        dev = keep[0].device
        dist = outSgnDist if outSgnDist is not None else torch.empty(npts, dtype=torch.float64,
                                                                     device=dev)
        for t, name, dt, shape in ((dist, "outSgnDist", torch.float64, (npts,)),
                                   (outClosestPts, "outClosestPts", torch.float64, (npts, 3)),
                                   (outNormals, "outNormals", torch.float64, (npts, 3)),
                                   (outElem, "outElem", torch.int32, (npts,))):
            if t is not None:
                _dev(t, dt, name)
                if tuple(t.shape) != shape or not t.is_contiguous():
                    raise ValueError(f"{name} must be a contiguous tensor of shape {shape}")
        mode = self._fp if fp_mode is None else FP_MODES[fp_mode]
        L.check(L.lib().axsq_sdf_eval(self._h, x, y, z, stride, npts, mode, _ptr(dist),
                                      _ptr(outClosestPts), _ptr(outNormals), _ptr(outElem),
                                      stream_ptr()))
        return dist

    # -- SignedDistance.hpp:508-524 (one host point; batched through the same kernel)
    def computeDistance(self, pt, fp_mode=None):
        """(distance, closest point (3,), unit normal (3,))."""
        q = L._dbl3(*[float(c) for c in np.asarray(pt, np.float64).reshape(3)])
        d, cp, nrm = ctypes.c_double(), L._dbl3(), L._dbl3()
        mode = self._fp if fp_mode is None else FP_MODES[fp_mode]
        L.check(L.lib().axsq_sdf_eval_point(self._h, q, mode, ctypes.byref(d), cp, nrm))
        return d.value, np.array(cp[:]), np.array(nrm[:])

    def backward(self, queryPts, dist, elem, gout, want_query=True, want_vertices=True):
        """Adjoint: (grad_q (n,3) or None, grad_v (V,3) or None), float64."""
        x, y, z, stride, n, keep = point_args(queryPts)
        dev = keep[0].device
        d = _dev(dist, torch.float64, "dist")
        e = _dev(elem, torch.int32, "elem")
        g = _dev(gout, torch.float32, "gout")
        if d.numel() != n or e.numel() != n or g.numel() != n:
            raise ValueError("dist/elem/gout must have one entry per query")
        gq = torch.empty((n, 3), dtype=torch.float64, device=dev) if want_query else None
        gv = torch.empty((self.nverts, 3), dtype=torch.float64, device=dev) if want_vertices \
            else None
        L.check(L.lib().axsq_sdf_backward(self._h, x, y, z, stride, n, _ptr(d), _ptr(e), _ptr(g),
                                          _ptr(gq), _ptr(gv), stream_ptr()))
        return gq, gv

    def enableStats(self, on=True):
        L.check(L.lib().axsq_sdf_stats_enable(self._h, int(bool(on))))

    def lastStats(self):
        v, lf = ctypes.c_int64(), ctypes.c_int64()
        L.check(L.lib().axsq_sdf_stats_get(self._h, ctypes.byref(v), ctypes.byref(lf)))
        return v.value, lf.value

    def lastSuspended(self):
        """Walks of the last computeDistances finished by the second pass."""
        n = ctypes.c_int64()
        L.check(L.lib().axsq_sdf_last_suspended(self._h, ctypes.byref(n)))
        return n.value

    def lastEarly(self):
        """Walks the last computeDistances ran from the start beside its first
        pass (the previous call's longest, on the same query arrays)."""
        n = ctypes.c_int64()
        L.check(L.lib().axsq_sdf_last_early(self._h, ctypes.byref(n)))
        return n.value


# ---- quest C-style interface (quest/interface/signed_distance.{hpp,cpp}) ----
# These call the library's process-wide query through its C names
# (include/axsq_quest.h), so Python, C and C++ callers share one query and
# one set of options. Misuse raises RuntimeError where the reference aborts.
def _qcheck():
    err = L.last_error()
    if err:
        raise RuntimeError(err)


def signed_distance_initialized():
    return bool(L.lib().QUEST_signed_distance_initialized())


def signed_distance_set_dimension(dim):
    L.lib().QUEST_signed_distance_set_dimension(int(dim))
    _qcheck()


def signed_distance_set_closed_surface(status):
    L.lib().QUEST_signed_distance_set_closed_surface(bool(status))
    _qcheck()


def signed_distance_set_compute_signs(compute_sign):
    L.lib().QUEST_signed_distance_set_compute_signs(bool(compute_sign))
    _qcheck()


def signed_distance_set_allocator(allocator_id):
    L.lib().QUEST_signed_distance_set_allocator(int(allocator_id))
    _qcheck()


def signed_distance_set_verbose(status):
    L.lib().QUEST_signed_distance_set_verbose(bool(status))
    _qcheck()


def signed_distance_use_shared_memory(status):
    L.lib().QUEST_signed_distance_use_shared_memory(bool(status))
    _qcheck()


SignedDistExec_CPU, SignedDistExec_OpenMP, SignedDistExec_GPU = 0, 1, 2


def signed_distance_set_execution_space(exec_space):
    L.lib().QUEST_signed_distance_set_execution_space(int(exec_space))
    _qcheck()


def signed_distance_set_precision(mode):
    """Selects "fp64" (plain) or "fp32" (FP32 traversal, FP64 finish)."""
    if mode not in ("fp64", "fp32"):
        raise RuntimeError(f"unknown precision {mode!r}")
    L.lib().axsq_signed_distance_set_precision(FP_MODES[mode])
    _qcheck()


def signed_distance_init(mesh_or_file):
    """Init from an STL path or a (vertices, triangles) pair of CUDA tensors.
    Returns 0 on success, -1 on failure (signed_distance.cpp:29-30, 109-227)."""
    if isinstance(mesh_or_file, str):
        return int(L.lib().QUEST_signed_distance_init_serial(mesh_or_file.encode()))
    verts, tris = mesh_or_file
    x, y, z = _soa(verts)
    t = _dev(tris, torch.int32, "triangles", 3)
    torch.cuda.current_stream().synchronize()
    return int(L.lib().axsq_signed_distance_init_mesh(_ptr(x), _ptr(y), _ptr(z), x.shape[0],
                                                      _ptr(t), t.shape[0]))


def signed_distance_evaluate(x, y=None, z=None, phi=None, npoints=None):
    """Scalar form evaluate(x, y, z) -> float, or the batched SoA form
    evaluate(x, y, z, phi) with (n,) float64 CUDA tensors
    (signed_distance.cpp:346-465)."""
    if isinstance(x, torch.Tensor):
        xs, ys, zs = (_dev(c, torch.float64, "coordinates").contiguous() for c in (x, y, z))
        n = xs.shape[0] if npoints is None else int(npoints)
        out = phi if phi is not None else torch.empty(n, dtype=torch.float64, device=xs.device)
        _dev(out, torch.float64, "phi")
        torch.cuda.current_stream().synchronize()
        L.check(L.lib().axsq_signed_distance_evaluate_batch(_ptr(xs), _ptr(ys), _ptr(zs), n,
                                                            _ptr(out)))
        return out
    d = L.lib().QUEST_signed_distance_evaluate_0(float(x), float(y), float(z))
    _qcheck()
    return d


def signed_distance_evaluate_cp(x, y, z):
    """evaluate(x, y, z, cp_x, ..., n_z) (signed_distance.hpp:266-274):
    (distance, closest point (3,), unit normal (3,))."""
    out = [ctypes.c_double() for _ in range(6)]
    d = L.lib().QUEST_signed_distance_evaluate_1(float(x), float(y), float(z),
                                                 *[ctypes.byref(o) for o in out])
    _qcheck()
    return d, np.array([o.value for o in out[:3]]), np.array([o.value for o in out[3:]])


def signed_distance_get_mesh_bounds():
    lo, hi = L._dbl3(), L._dbl3()
    L.lib().QUEST_signed_distance_get_mesh_bounds(lo, hi)
    _qcheck()
    return np.array(lo[:]), np.array(hi[:])


def signed_distance_finalize():
    L.lib().QUEST_signed_distance_finalize()


# ---- mesh tester ---------------------------------------------------------------
class _DeviceArray:
    """A library-owned device int32 array seen through __cuda_array_interface__."""

    def __init__(self, ptr, n):
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": "<i4",
                                         "data": (ptr, False), "version": 2}


def _take_int32(ptr, n, device):
    """Copies a library-allocated device int32 array into a torch tensor and
    releases the library's copy."""
    try:
        if n == 0:
            return torch.empty(0, dtype=torch.int32, device=device)
        view = torch.as_tensor(_DeviceArray(ptr, n), device=device)
        out = view.clone()
        torch.cuda.current_stream(device).synchronize()
        return out
    finally:
        if ptr:
            L.lib().axsq_free(ctypes.c_void_p(ptr))


def findTriMeshIntersectionsBVH(vertices, triangles, intersectionThreshold=1e-8):
    """quest::findTriMeshIntersectionsBVH<MI355X_EXEC, double>
    (src/axom/quest/MeshTester.hpp:66-87).

    vertices: (V,3) float64 CUDA tensor or (x, y, z); triangles: (T,3) int32
    CUDA tensor. Returns (intersections, degenerateIndices): a (K,2) int32 CUDA
    tensor of intersecting triangle pairs (first < second) in the reference's
    sequential order, and the degenerate triangles' indices, ascending."""
    x, y, z = _soa(vertices)
    t = _dev(triangles, torch.int32, "triangles", 3)
    pf, ps, pd = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    npairs, ndeg = ctypes.c_int64(), ctypes.c_int64()
    L.check(L.lib().axsq_mesh_self_intersections(
        _ptr(x), _ptr(y), _ptr(z), x.shape[0], _ptr(t), t.shape[0], float(intersectionThreshold),
        ctypes.byref(pf), ctypes.byref(ps), ctypes.byref(npairs), ctypes.byref(pd),
        ctypes.byref(ndeg), stream_ptr()))
    first = _take_int32(pf.value, npairs.value, x.device)
    second = _take_int32(ps.value, npairs.value, x.device)
    degen = _take_int32(pd.value, ndeg.value, x.device)
    return torch.stack([first, second], dim=1), degen



def findCandidatesBVH(insertVertices, insertCells, queryVertices, queryCells):
    """findCandidatesBVH of src/axom/quest/examples/quest_candidates_example.cpp:
    397-495: candidate pairs between two meshes of cells with the same number of
    nodes (hexahedra: 8).

    Vertices: (V,3) float64 CUDA tensors or (x, y, z); cells: (C,k) int32 CUDA
    tensors, k in 1..8. Returns a (K,2) int32 CUDA tensor of (query cell, insert
    cell) pairs whose bounding boxes overlap (the insert boxes scaled by the
    BVH's default 1.000123), grouped by query cell ascending and, within one
    query cell, in BVH traversal order."""
    ix, iy, iz = _soa(insertVertices)
    qx, qy, qz = _soa(queryVertices)
    ic = _dev(insertCells, torch.int32, "insertCells")
    qc = _dev(queryCells, torch.int32, "queryCells")
    if ic.dim() != 2 or qc.dim() != 2 or ic.shape[1] != qc.shape[1]:
        raise ValueError("cells must be (C,k) tensors with the same k for both meshes")
    pf, ps = ctypes.c_void_p(), ctypes.c_void_p()
    npairs = ctypes.c_int64()
    L.check(L.lib().axsq_mesh_candidates(
        _ptr(ix), _ptr(iy), _ptr(iz), ix.shape[0], _ptr(ic), ic.shape[0],
        _ptr(qx), _ptr(qy), _ptr(qz), qx.shape[0], _ptr(qc), qc.shape[0], int(ic.shape[1]),
        ctypes.byref(pf), ctypes.byref(ps), ctypes.byref(npairs), stream_ptr()))
    first = _take_int32(pf.value, npairs.value, ix.device)
    second = _take_int32(ps.value, npairs.value, ix.device)
    return torch.stack([first, second], dim=1)
