"""MI355X (gfx950) spatial-query path of Axom: spin::BVH and
quest::SignedDistance over hand-written HIP kernels (libaxsq.so).

Submodules: spin (BVH mirror), quest (SignedDistance mirror + C-style API),
autograd (torch.autograd.Function harness), synth (seeded workloads),
stl (STL reader), build (compiles libaxsq.so).
"""
__version__ = "0.1.0"


def load():
    """Loads libaxsq.so (raises if it has not been built)."""
    from . import _lib
    return _lib.lib()
