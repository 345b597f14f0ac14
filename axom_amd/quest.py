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

    # -- SignedDistance.hpp:508-524 (one point; batched through the same kernel)
    def computeDistance(self, pt):
        q = torch.as_tensor(np.asarray(pt, np.float64).reshape(1, 3), device="cuda")
        cp = torch.empty((1, 3), dtype=torch.float64, device="cuda")
        nrm = torch.empty((1, 3), dtype=torch.float64, device="cuda")
        d = self.computeDistances(1, q, outClosestPts=cp, outNormals=nrm)
        return float(d[0]), cp[0].cpu().numpy(), nrm[0].cpu().numpy()

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


# ---- quest C-style interface (quest/interface/signed_distance.{hpp,cpp}) ----
class _Params:
    def __init__(self):
        self.dimension = 3
        self.verbose = False
        self.is_closed_surface = True
        self.compute_sign = True
        self.fp_mode = "fp64"


_params = _Params()
_query = None


def _must_not_be_initialized(name):
    if _query is not None:
        raise RuntimeError(f"{name}: signed distance has already been initialized; "
                           "call signed_distance_finalize() first")


def signed_distance_initialized():
    return _query is not None


def signed_distance_set_dimension(dim):
    _must_not_be_initialized("signed_distance_set_dimension")
    if dim != 3:
        raise RuntimeError("only 3-D triangle surfaces are supported")
    _params.dimension = dim


def signed_distance_set_closed_surface(status):
    _must_not_be_initialized("signed_distance_set_closed_surface")
    _params.is_closed_surface = bool(status)


def signed_distance_set_compute_signs(compute_sign):
    _must_not_be_initialized("signed_distance_set_compute_signs")
    _params.compute_sign = bool(compute_sign)


def signed_distance_set_verbose(status):
    _must_not_be_initialized("signed_distance_set_verbose")
    _params.verbose = bool(status)


def signed_distance_set_precision(mode):
    """Selects "fp64" (plain) or "fp32" (FP32 traversal, FP64 finish)."""
    _must_not_be_initialized("signed_distance_set_precision")
    if mode not in ("fp64", "fp32"):
        raise RuntimeError(f"unknown precision {mode!r}")
    _params.fp_mode = mode


def signed_distance_init(mesh_or_file):
    """Init from an STL path or a (vertices, triangles) pair. Returns 0 on
    success, -1 on failure (signed_distance.cpp:29-30, 109-227)."""
    global _query
    _must_not_be_initialized("signed_distance_init")
    try:
        if isinstance(mesh_or_file, str):
            from .stl import read_stl
            v, t = read_stl(mesh_or_file)
            verts = torch.as_tensor(v, device="cuda")
            tris = torch.as_tensor(t, device="cuda")
        else:
            verts, tris = mesh_or_file
        _query = SignedDistance(verts, tris, _params.is_closed_surface, _params.compute_sign,
                                fp_mode=_params.fp_mode)
    except (OSError, ValueError) as e:
        if _params.verbose:
            print(f"signed_distance_init failed: {e}")
        return -1
    return 0


def signed_distance_evaluate(x, y=None, z=None, phi=None, npoints=None):
    """Scalar form evaluate(x, y, z) -> float, or the batched SoA form
    evaluate(x, y, z, phi) with (n,) CUDA tensors (signed_distance.cpp:346-465)."""
    if _query is None:
        raise RuntimeError("signed_distance_evaluate: signed distance is not initialized")
    if isinstance(x, torch.Tensor):
        n = x.shape[0] if npoints is None else int(npoints)
        out = _query.computeDistances(n, (x, y, z), outSgnDist=phi)
        return out
    return _query.computeDistance((x, y, z))[0]


def signed_distance_get_mesh_bounds():
    if _query is None:
        raise RuntimeError("signed_distance_get_mesh_bounds: not initialized")
    return _query.getMeshBounds()


def signed_distance_finalize():
    global _query
    _query = None


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
