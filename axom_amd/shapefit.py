"""Shape-fitting loop over the signed-distance forward and adjoint.

Each iteration: rebuild the BVH on the moved vertices (Morton sort + refit),
signed-distance forward at fixed query points, loss L = sum((d - target)^2)
(the squared L2 norm of the residual), adjoint to the vertices with upstream
gradient dL/dd = 2 (d - target) (FP32), and a vertex step of length lr: every
vertex moves along -dL/dV, scaled so that the vertex with the largest gradient
moves by lr (a plain gradient step is either negligible or unstable here,
since a vertex that is the closest feature of many queries collects a large
gradient). Per-iteration build / forward / backward times are measured with
events on the current stream; the reported loss is the root-mean-square
residual."""
import time

import torch

from .quest import SignedDistance


def shape_fit(vertices, triangles, queries, target, iterations=20, lr=1e-3, fp_mode="fp32"):
    """vertices (V,3) f64, triangles (T,3) i32, queries (Q,3) f64, target (Q,) f64,
    all CUDA tensors. Returns (final vertices, list of per-iteration dicts)."""
    V = vertices.clone()
    q = queries.contiguous()
    nq = q.shape[0]
    sdf = SignedDistance(V, triangles, isWatertight=True, computeSign=True, fp_mode=fp_mode)
    dist = torch.empty(nq, dtype=torch.float64, device=q.device)
    elem = torch.empty(nq, dtype=torch.int32, device=q.device)
    history = []
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    for it in range(iterations):
        ev[0].record()
        if it > 0:
            sdf.updateVertices(V)
        ev[1].record()
        sdf.computeDistances(nq, q, outSgnDist=dist, outElem=elem, fp_mode=fp_mode)
        ev[2].record()
        r = dist - target
        g = (2.0 * r).to(torch.float32)
        _, gv = sdf.backward(q, dist, elem, g, want_query=False, want_vertices=True)
        gmax = float(torch.linalg.vector_norm(gv, dim=1).max())
        if gmax > 0.0:
            V -= (lr / gmax) * gv
        ev[3].record()
        torch.cuda.synchronize()
        rms = float(torch.sqrt(torch.mean(r * r)))
        history.append({"iteration": it, "rms_residual": rms,
                        "suspended_walks": sdf.lastSuspended(),
                        "build_ms": ev[0].elapsed_time(ev[1]),
                        "forward_ms": ev[1].elapsed_time(ev[2]),
                        "backward_ms": ev[2].elapsed_time(ev[3])})
    return V, history
