"""Shape-fitting loop over the signed-distance forward and adjoint.

Each iteration: rebuild the BVH on the moved vertices (Morton sort + refit),
signed-distance forward at fixed query points, loss L = mean((d - target)^2),
adjoint to the vertices with upstream gradient dL/dd = 2 (d - target) / Q (FP32),
and a gradient step V -= lr * dL/dV. Per-iteration build / forward / backward
times are measured with events on the current stream.
"""
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
        loss = float(torch.mean(r * r))
        g = (2.0 / nq * r).to(torch.float32)
        _, gv = sdf.backward(q, dist, elem, g, want_query=False, want_vertices=True)
        V -= lr * gv
        ev[3].record()
        torch.cuda.synchronize()
        history.append({"iteration": it, "loss": loss,
                        "build_ms": ev[0].elapsed_time(ev[1]),
                        "forward_ms": ev[1].elapsed_time(ev[2]),
                        "backward_ms": ev[2].elapsed_time(ev[3])})
    return V, history
