"""Shape-fitting loop over the signed-distance forward and adjoint.

Each iteration, on a triangle mesh whose vertices move:
  1. rebuild the BVH on the current vertices (Morton sort + refit);
  2. signed-distance forward d(q) at the fixed query points (FP32 traversal,
     FP64 finish, by default);
  3. loss L = mean((d - target)^2);
  4. adjoint to the vertices with the upstream gradient
     dL/dd = 2 (d - target) / Q, passed to the kernel in float32;
  5. vertex step V -= lr * dL/dV.

step="normalized" replaces step 5 by a step of length lr for the vertex with
the largest gradient (every vertex moves along -dL/dV, scaled by
lr / max_k |dL/dV_k|); it is not the plain gradient step and is only used
when asked for.

Per-iteration rebuild / forward / backward times are measured with events on
the current stream.
"""
import torch

from .quest import SignedDistance


def _upstream(dist, target):
    """dL/dd of L = mean((d - target)^2), in float32 (the adjoint's input type)."""
    r = dist - target
    return (r * (2.0 / r.numel())).to(torch.float32), r


def shape_fit(vertices, triangles, queries, target, iterations=20, lr=1e-3, fp_mode="fp32",
              step="plain", sdf=None):
    """vertices (V,3) f64, triangles (T,3) i32, queries (Q,3) f64, target (Q,) f64,
    all CUDA tensors. Returns (final vertices, list of per-iteration dicts).
    `sdf` may be a quest.SignedDistance already built on `vertices` (created
    with fp_mode "fp32" when fp_mode is "fp32")."""
    if step not in ("plain", "normalized"):
        raise ValueError(f"unknown step {step!r}")
    V = vertices.clone()
    q = queries.contiguous()
    nq = q.shape[0]
    if sdf is None:
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
        g, r = _upstream(dist, target)
        _, gv = sdf.backward(q, dist, elem, g, want_query=False, want_vertices=True)
        if step == "plain":
            V -= lr * gv
        else:
            gmax = float(torch.linalg.vector_norm(gv, dim=1).max())
            if gmax > 0.0:
                V -= (lr / gmax) * gv
        ev[3].record()
        torch.cuda.synchronize()
        history.append({"iteration": it, "loss": float(torch.mean(r * r)),
                        "suspended_walks": sdf.lastSuspended(),
                        "early_walks": sdf.lastEarly(),
                        "build_ms": ev[0].elapsed_time(ev[1]),
                        "forward_ms": ev[1].elapsed_time(ev[2]),
                        "backward_ms": ev[2].elapsed_time(ev[3])})
    return V, history
