"""Host mirror of axom::spin::BVH<NDIMS, MI355X_EXEC, FloatType> (NDIMS 3 or 2,
FloatType double or float) over libaxsq.

Method names, argument meaning and defaults follow src/axom/spin/BVH.hpp:129-419
so that code and tests written against the reference read the same. Data
lives in torch CUDA tensors (device memory); the compute is the HIP library.

Layouts: boxes are (n, 6) float64 {min xyz, max xyz} (BoundingBox<double,3>);
points are (n, 3) float64 (Point<double,3>) or an (x, y, z) tuple of (n,)
float64 tensors (ZipPoint); rays are (n, 6) float64 {origin, unit direction}
as built by make_rays (Ray's constructor normalises, Ray.hpp:122-128).
With NDIMS=2: boxes (n, 4) {min xy, max xy}, points (n, 2) or (x, y), rays
(n, 4) {origin xy, unit direction xy} (make_rays2d).
With FloatType=torch.float32 (BVH<NDIMS, Exec, float>) every box, point and
ray tensor is float32 instead (make_rays / make_rays2d with
dtype=torch.float32 normalise as Ray<float> does).
"""
import ctypes

import numpy as np
import torch

from . import _lib as L

BVH_BUILD_OK = 0      # spin/BVH.hpp:39-43
BVH_BUILD_FAILED = -1

FP_MODES = {"fp64": L.AXSQ_FP64, "fp32": L.AXSQ_FP32_TRAVERSAL,
            L.AXSQ_FP64: L.AXSQ_FP64, L.AXSQ_FP32_TRAVERSAL: L.AXSQ_FP32_TRAVERSAL}


_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def stream_ptr():
    """The current CUDA (HIP) stream of the current device, as a pointer."""
    if _raw_stream is not None:
        return ctypes.c_void_p(_raw_stream(torch.cuda.current_device()))
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _dev(t, dtype, name, cols=None):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{name} must be a CUDA tensor")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if cols is not None and (t.dim() != 2 or t.shape[1] != cols):
        raise ValueError(f"{name} must have shape (n, {cols})")
    return t.contiguous()


def point_args(points, dtype=torch.float64):
    """(x, y, z, stride, n, keepalive) for AoS (n,3) or SoA (x,y,z) points."""
    if isinstance(points, (tuple, list)):
        x, y, z = (_dev(c, dtype, "points") for c in points)
        if not (x.shape == y.shape == z.shape) or x.dim() != 1:
            raise ValueError("SoA points need three (n,) tensors of equal length")
        return _ptr(x), _ptr(y), _ptr(z), 1, x.shape[0], (x, y, z)
    p = _dev(points, dtype, "points", 3)
    base, w = p.data_ptr(), p.element_size()
    return (ctypes.c_void_p(base), ctypes.c_void_p(base + w), ctypes.c_void_p(base + 2 * w), 3,
            p.shape[0], (p,))


def point_args2d(points, dtype=torch.float64):
    """(x, y, None, stride, n, keepalive) for AoS (n,2) or SoA (x,y) 2-D points."""
    if isinstance(points, (tuple, list)):
        x, y = (_dev(c, dtype, "points") for c in points)
        if x.shape != y.shape or x.dim() != 1:
            raise ValueError("SoA 2-D points need two (n,) tensors of equal length")
        return _ptr(x), _ptr(y), None, 1, x.shape[0], (x, y)
    p = _dev(points, dtype, "points", 2)
    base = p.data_ptr()
    return ctypes.c_void_p(base), ctypes.c_void_p(base + p.element_size()), None, 2, \
        p.shape[0], (p,)


def _unit_f32(d):
    """Vector<float>::unitVector (Vector.hpp:477-490): the squared norm summed in
    float, v / float(sqrt(double norm)) as v * (1 / double(that))
    (core/NumericArray.hpp:498-514), each component rounded to float; (1, 0..)
    for a zero vector."""
    len_sq = torch.zeros(d.shape[0], dtype=torch.float32, device=d.device)
    for k in range(d.shape[1]):
        len_sq = len_sq + d[:, k] * d[:, k]
    s = torch.sqrt(len_sq.double()).float()
    u = (d.double() * (1.0 / s.double())[:, None]).float()
    tiny = len_sq.double() < 1e-50
    if bool(tiny.any()):
        e = torch.zeros(d.shape[1], dtype=torch.float32, device=d.device)
        e[0] = 1.0
        u[tiny] = e
    return u


def make_rays2d(origins, directions, dtype=torch.float64):
    """2-D rays {origin xy, unit direction xy}, the direction normalised as
    primal::Ray<double, 2> does (1/sqrt of (0 + x*x) + y*y), or as
    Ray<float, 2> does for dtype float32."""
    o = _dev(origins, dtype, "origins", 2)
    d = _dev(directions, dtype, "directions", 2)
    if dtype == torch.float32:
        return torch.cat([o, _unit_f32(d)], 1).contiguous()
    len_sq = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]
    inv = 1.0 / torch.sqrt(len_sq)
    u = d * inv[:, None]
    tiny = len_sq < 1e-50
    if bool(tiny.any()):
        u[tiny] = torch.tensor([1.0, 0.0], dtype=torch.float64, device=u.device)
    return torch.cat([o, u], 1).contiguous()


def make_rays(origins, directions, dtype=torch.float64):
    """Rays with directions normalised as primal::Ray does (Vector::unitVector,
    Vector.hpp:477-490: multiply by 1/sqrt of ((0 + x*x) + y*y) + z*z), or as
    Ray<float> does for dtype float32."""
    o = _dev(origins, dtype, "origins", 3)
    d = _dev(directions, dtype, "directions", 3)
    if dtype == torch.float32:
        return torch.cat([o, _unit_f32(d)], 1).contiguous()
    len_sq = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    inv = 1.0 / torch.sqrt(len_sq)
    u = d * inv[:, None]
    tiny = len_sq < 1e-50
    if bool(tiny.any()):
        u[tiny] = torch.tensor([1.0, 0.0, 0.0], dtype=torch.float64, device=u.device)
    return torch.cat([o, u], 1).contiguous()


class BVH:
    """axom::spin::BVH<NDIMS, MI355X_EXEC, FloatType> (spin/BVH.hpp:129-419), NDIMS
    3 (default) or 2, FloatType torch.float64 (default, double) or torch.float32
    (float: boxes, queries, scale factor and tolerance are float; fp_mode does
    not apply)."""

    DEFAULT_SCALE_FACTOR = 1.000123                         # BVH.hpp:410
    DEFAULT_TOLERANCE = float(np.finfo(np.float64).eps)     # BVH.hpp:411-412

    def __init__(self, boxes=None, numItems=None, tolerance=None,
                 scaleFactor=DEFAULT_SCALE_FACTOR, fp_mode="fp64", NDIMS=3,
                 FloatType=torch.float64):
        if NDIMS not in (2, 3):
            raise ValueError("NDIMS must be 2 or 3")
        FloatType = {"double": torch.float64, "float": torch.float32}.get(FloatType, FloatType)
        if FloatType not in (torch.float64, torch.float32):
            raise ValueError("FloatType must be torch.float64 or torch.float32")
        self.NDIMS = NDIMS
        self.FloatType = FloatType
        self._f32 = FloatType == torch.float32
        self._h = None
        eps = np.finfo(np.float32 if self._f32 else np.float64).eps
        self._tol = float(eps if tolerance is None else tolerance)
        self._scale = float(np.float32(scaleFactor)) if self._f32 else float(scaleFactor)
        self._fp = FP_MODES[fp_mode]
        if boxes is not None:
            self.initialize(boxes, numItems)

    def __del__(self):
        # at interpreter exit the module globals may already be gone; the
        # process then releases the device memory itself
        if getattr(self, "_h", None) and L is not None:
            L.lib().axsq_bvh_destroy(self._h)
            self._h = None

    # -- BVH.hpp:246-247, 426-477
    def initialize(self, boxes, numItems=None):
        b = _dev(boxes, self.FloatType, "boxes", 2 * self.NDIMS)
        n = b.shape[0] if numItems is None else int(numItems)
        if n > b.shape[0]:
            raise ValueError("numItems exceeds the number of boxes")
        if self._h:  # build again in place, keeping the device arrays
            if self._f32:
                rc = L.lib().axsq_bvh_rebuild_f32(self._h, _ptr(b), n, self._scale, stream_ptr())
            else:
                rc = L.lib().axsq_bvh_rebuild(self._h, _ptr(b), n, self._scale, self._fp,
                                              stream_ptr())
            if rc != 0:
                L.lib().axsq_bvh_destroy(self._h)
                self._h = None
                L.check(rc)
            return BVH_BUILD_OK
        h = ctypes.c_void_p()
        if self._f32:
            L.check(L.lib().axsq_bvh_build_f32(ctypes.byref(h), _ptr(b), n, self.NDIMS,
                                               self._scale, stream_ptr()))
        else:
            build = L.lib().axsq_bvh_build2d if self.NDIMS == 2 else L.lib().axsq_bvh_build
            L.check(build(ctypes.byref(h), _ptr(b), n, self._scale, self._fp, stream_ptr()))
        self._h = h
        if getattr(self, "_alloc", 0):
            L.check(L.lib().axsq_bvh_set_allocator(self._h, self._alloc))
        if getattr(self, "_stats", False):
            L.check(L.lib().axsq_bvh_stats_enable(self._h, 1))
        return BVH_BUILD_OK

    def isInitialized(self):
        return bool(self._h)

    def setScaleFactor(self, s):
        self._scale = float(np.float32(s)) if self._f32 else float(s)

    def getScaleFactor(self):
        return self._scale

    def setTolerance(self, eps):
        self._tol = float(np.float32(eps)) if self._f32 else float(eps)

    def getTolerance(self):
        return self._tol

    def getBounds(self):
        """(lo, hi) of the root box; an invalid box before initialize()."""
        d = self.NDIMS
        ft = np.float32 if self._f32 else np.float64
        if not self._h:
            return (np.full(d, np.finfo(ft).max, ft), np.full(d, -np.finfo(ft).max, ft))
        lo, hi = L._dbl3(), L._dbl3()
        L.check(L.lib().axsq_bvh_bounds(self._h, lo, hi))
        return np.array(lo[:d]).astype(ft), np.array(hi[:d]).astype(ft)

    def enableStats(self, on=True):
        """Counts inner-node visits in the candidate searches' count passes
        (no reference counterpart)."""
        self._stats = bool(on)
        if self._h:
            L.check(L.lib().axsq_bvh_stats_enable(self._h, int(self._stats)))

    def lastStats(self):
        """Inner-node visits of the last count pass, summed over its queries."""
        v = ctypes.c_int64()
        L.check(L.lib().axsq_bvh_stats_get(self._h, ctypes.byref(v)))
        return v.value

    def numLeaves(self):
        return int(L.lib().axsq_bvh_num_leaves(self._h)) if self._h else 0

    def _find(self, kind, offsets, counts, n, x, y, z, stride):
        """One search into a candidate tensor sized from the previous call's
        total; when that is too small, a fill of the exact size follows."""
        if not self._h:
            raise RuntimeError("BVH is not initialized")
        off = _dev(offsets, torch.int32, "offsets")
        cnt = _dev(counts, torch.int32, "counts")
        # LinearBVH.hpp:284-285
        if off.numel() != n or cnt.numel() != n:
            raise ValueError("offsets/counts length not equal to the number of queries")
        total = ctypes.c_int64()
        s = stream_ptr()
        cap = getattr(self, "_cap_hint", 0)
        cand = torch.empty(max(cap, 1), dtype=torch.int32, device=off.device)
        sfx = "_f32" if self._f32 else ""
        find_into = getattr(L.lib(), "axsq_bvh_find_into" + sfx)
        find_fill = getattr(L.lib(), "axsq_bvh_find_fill" + sfx)
        L.check(find_into(self._h, kind, x, y, z, stride, n, self._tol, _ptr(off), _ptr(cnt),
                          _ptr(cand), cap, ctypes.byref(total), s))
        t = max(total.value, 0)
        if t > cap:
            cand = torch.empty(t, dtype=torch.int32, device=off.device)
            L.check(find_fill(self._h, kind, x, y, z, stride, n, self._tol, _ptr(off),
                              _ptr(cnt), _ptr(cand), s))
            self._cap_hint = t + t // 8 + 1024
        if off.data_ptr() != offsets.data_ptr():
            offsets.copy_(off)
        if cnt.data_ptr() != counts.data_ptr():
            counts.copy_(cnt)
        return cand[:t]

    def setAllocatorID(self, allocator_id):
        """BVH::setAllocatorID (BVH.hpp:251-261): the memory of the candidate
        arrays of the C-ABI's one-call search (0 device, 1 managed, 2 pinned
        host). This mirror returns torch tensors, which live in device memory."""
        if allocator_id not in (0, 1, 2):
            raise ValueError(f"unknown allocator id {allocator_id}")
        self._alloc = int(allocator_id)
        if self._h:
            L.check(L.lib().axsq_bvh_set_allocator(self._h, self._alloc))

    def getAllocatorID(self):
        return getattr(self, "_alloc", 0)

    def writeVtkFile(self, fileName):
        """BVH::writeVtkFile (BVH.hpp:405,572-577)."""
        if not self._h:
            raise RuntimeError("BVH is not initialized")
        L.check(L.lib().axsq_bvh_write_vtk(self._h, str(fileName).encode()))

    # -- BVH.hpp:341-346, 482-505
    def findPoints(self, offsets, counts, numPts, points):
        x, y, z, stride, n, keep = (point_args2d if self.NDIMS == 2 else point_args)(
            points, self.FloatType)
        return self._find(L.QUERY_POINTS, offsets, counts, int(numPts), x, y, z, stride)

    # -- BVH.hpp:367-372, 510-536 (rays from make_rays)
    def findRays(self, offsets, counts, numRays, rays):
        w = 2 * self.NDIMS
        r = _dev(rays, self.FloatType, "rays", w)
        return self._find(L.QUERY_RAYS, offsets, counts, int(numRays), _ptr(r), None, None, w)

    # -- BVH.hpp:393-398, 541-568
    def findBoundingBoxes(self, offsets, counts, numBoxes, boxes):
        w = 2 * self.NDIMS
        b = _dev(boxes, self.FloatType, "boxes", w)
        return self._find(L.QUERY_BOXES, offsets, counts, int(numBoxes), _ptr(b), None, None, w)

    def export(self):
        """Tree arrays in the reference's LinearBVH layout (device tensors)."""
        n = self.numLeaves()
        dev = torch.device("cuda", torch.cuda.current_device())
        inner = torch.empty((2 * (n - 1), 2 * self.NDIMS), dtype=torch.float64, device=dev)
        children = torch.empty(2 * (n - 1), dtype=torch.int32, device=dev)
        leafs = torch.empty(n, dtype=torch.int32, device=dev)
        mcodes = torch.empty(n, dtype=torch.int32, device=dev)
        L.check(L.lib().axsq_bvh_export(self._h, _ptr(inner), _ptr(children), _ptr(leafs),
                                        _ptr(mcodes), stream_ptr()))
        return dict(nodes=inner, children=children, leafs=leafs, mcodes=mcodes)
