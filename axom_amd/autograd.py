"""torch.autograd.Function harness over the signed-distance adjoint.

signed_distance(queries, vertices, sdf) returns the signed distances of the
(Q,3) float64 queries to the mesh held by `sdf` (a quest.SignedDistance whose
current vertices are `vertices`), and routes gradients to `queries` and
`vertices` through axsq_sdf_backward:

  d sd / d q   = sgn * (q - cp) / |q - cp|
  d sd / d V_k = -w_k * d sd / d q      (w = barycentric weights of cp)

The upstream gradient is passed to the kernel in float32 (the adjoint's
input type); the gradients come back in float64 with deterministic sums.
"""
import torch

from .quest import SignedDistance


class SignedDistanceFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, queries, vertices, sdf, fp_mode):
        if not isinstance(sdf, SignedDistance):
            raise TypeError("sdf must be an axom_amd.quest.SignedDistance")
        n = queries.shape[0]
        q = queries.detach().contiguous()
        elem = torch.empty(n, dtype=torch.int32, device=q.device)
        dist = sdf.computeDistances(n, q, outElem=elem, fp_mode=fp_mode)
        ctx.sdf = sdf
        ctx.save_for_backward(q, dist, elem)
        return dist

    @staticmethod
    def backward(ctx, gout):
        q, dist, elem = ctx.saved_tensors
        need_q, need_v = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        gq, gv = ctx.sdf.backward(q, dist, elem, gout.detach().to(torch.float32).contiguous(),
                                  want_query=need_q, want_vertices=need_v)
        return gq, gv, None, None


def signed_distance(queries, vertices, sdf, fp_mode="fp64"):
    """Differentiable signed distance (see module docstring)."""
    return SignedDistanceFn.apply(queries, vertices, sdf, fp_mode)
