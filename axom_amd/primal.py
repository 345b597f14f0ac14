"""Host mirror of the primal operator this package runs on the device:
primal::intersect(Triangle3, Triangle3, includeBoundary, EPS)
(src/axom/primal/operators/intersect.hpp:65-71), batched over pairs.
"""
import torch

from . import _lib as L
from .spin import _dev, _ptr, stream_ptr


def intersect(t1, t2, includeBoundary=False, EPS=1e-8):
    """t1, t2: (n, 3, 3) float64 CUDA tensors of triangle corners (A, B, C).
    Returns an (n,) bool CUDA tensor: pair k intersects."""
    a = _dev(t1, torch.float64, "t1").reshape(-1, 9)
    b = _dev(t2, torch.float64, "t2").reshape(-1, 9)
    if a.shape != b.shape:
        raise ValueError("t1 and t2 must hold the same number of triangles")
    out = torch.empty(a.shape[0], dtype=torch.uint8, device=a.device)
    L.check(L.lib().axsq_tri_tri_intersect(_ptr(a), _ptr(b), a.shape[0], int(bool(includeBoundary)),
                                           float(EPS), _ptr(out), stream_ptr()))
    return out.bool()
