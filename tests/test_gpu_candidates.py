"""GPU parity of the hex candidate search (findCandidatesBVH of
quest/examples/quest_candidates_example.cpp:397-495) against the oracle.

Bar: the pair list bit-exact, order included (query cell ascending, then the
insert cells in BVH traversal order), for hexahedra and for 4-node cells;
the closed form of a grid against itself; empty meshes; bad connectivity
refused."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from axom_amd import quest, synth  # noqa: E402
from axom_amd._lib import AxsqError  # noqa: E402
from oracle import oracle as orc  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _t(a):
    return torch.as_tensor(np.ascontiguousarray(a), device=DEV)


@pytest.mark.parametrize("ni,nq", [(3, 3), (20, 15), (48, 40)])
def test_pairs_bit_exact(ni, nq):
    iv, ih = synth.hex_grid(ni, jitter=0.15, seed=ni)
    qv, qh = synth.hex_grid(nq, origin=(0.7, 0.3, 0.5), h=ni / nq * 1.01, jitter=0.1,
                            seed=nq, angle=0.5)
    got = quest.findCandidatesBVH(_t(iv), _t(ih), _t(qv), _t(qh)).cpu().numpy()
    want = orc.mesh_candidates(iv, ih, qv, qh, nthreads=8)
    assert np.array_equal(got, want)
    assert got.shape[0] > 0


def test_grid_against_itself():
    v, h = synth.hex_grid(3)
    got = quest.findCandidatesBVH(_t(v), _t(h), _t(v), _t(h)).cpu().numpy()
    cnt = np.bincount(got[:, 0], minlength=27)
    assert cnt[13] == 27 and cnt[0] == 8 and got.shape[0] == 8 * 8 + 12 * 12 + 6 * 18 + 27


def test_four_node_cells_and_soa_vertices():
    iv, ih = synth.hex_grid(10, jitter=0.1, seed=3)
    qv, qh = synth.hex_grid(7, origin=(1.0, 1.0, 1.0), h=1.1, seed=5, angle=0.2)
    ic, qc = ih[:, [0, 1, 3, 4]], qh[:, [0, 2, 5, 7]]  # tetrahedra from the hexes' corners
    got = quest.findCandidatesBVH(tuple(_t(iv[:, d]) for d in range(3)), _t(ic),
                                  _t(qv), _t(qc)).cpu().numpy()
    assert np.array_equal(got, orc.mesh_candidates(iv, ic, qv, qc))


def test_empty_and_bad_connectivity():
    v, h = synth.hex_grid(2)
    empty = torch.zeros((0, 8), dtype=torch.int32, device=DEV)
    assert quest.findCandidatesBVH(_t(v), empty, _t(v), _t(h)).shape == (0, 2)
    assert quest.findCandidatesBVH(_t(v), _t(h), _t(v), empty).shape == (0, 2)
    bad = h.copy()
    bad[3, 5] = v.shape[0]
    with pytest.raises(AxsqError):
        quest.findCandidatesBVH(_t(v), _t(bad), _t(v), _t(h))
