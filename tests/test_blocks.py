"""Block tables of the CBFM: SetBlockCtrs (Division_blocks.f90:1616-1660) and
Extend_blocks (Extend_blocks.f90:1-162), host routines of the C-ABI library
(midas_block_frames, midas_extend_blocks).

Extend_blocks.f90 uses the MPI module, which cannot be built here, so there is
no reference-compiled vector for it.  Parity is pinned by (1) the oracle's
independent numpy restatement on a DLCA partition, exact integer lists; (2) a
lattice with a closed-form answer (the extension of a block is exactly the
nearest nc_extended layers of its neighbour); (3) the bounding-sphere
adjacency filter on a case where a cell inside the zone is excluded; (4) the
reference's edge cases (one block, no layers) and its Nbext_max stop."""
import numpy as np
import pytest

from midas_amd import api, geometry as g


def _cells(p):
    n = len(p)
    return api.CellTable(p[:, 0], p[:, 1], p[:, 2], np.ones(n), np.full(n, 1.5 + 0j))


def test_extend_blocks_matches_oracle(oracle):
    s = g.dlca_aggregate(nmono=6, cells_per_mono=40, seed=7)
    p = np.stack([s.x, s.y, s.z], 1)
    fr = api.block_frames(_cells(p), s.blocks)
    fo = oracle.block_frames(s.x, s.y, s.z, s.blocks)
    assert np.array_equal(fr, fo)
    for nc in (1, 2, 3):
        ext, adj = api.extend_blocks(_cells(p), s.blocks, float(s.sc[0]), nc)
        eo, ao = oracle.extend_blocks(s.x, s.y, s.z, s.blocks, float(s.sc[0]), nc)
        for b in range(len(s.blocks)):
            assert np.array_equal(ext[b], eo[b]), (nc, b)
            assert np.array_equal(adj[b], ao[b]), (nc, b)
        assert sum(len(e) for e in ext) > 0


@pytest.mark.parametrize("nc", [1, 2])
def test_extend_blocks_lattice_layers(nc):
    """Two face-touching 4x4x4 lattice cubes (spacing sc) and a far third one:
    block 0's extension is the nc layers of block 1 nearest to it (16 nc
    cells, in cell order), the far block adds nothing."""
    sc = 1.0
    ax = np.arange(4.0)
    cube = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
    p = np.concatenate([cube, cube + [4.0, 0, 0], cube + [40.0, 0, 0]])
    blocks = [np.arange(0, 64), np.arange(64, 128), np.arange(128, 192)]
    ext, adj = api.extend_blocks(_cells(p), blocks, sc, nc)
    want = np.nonzero((p[:, 0] >= 4.0) & (p[:, 0] <= 3.0 + nc))[0]
    assert np.array_equal(ext[0], want)
    assert len(ext[0]) == 16 * nc
    assert list(adj[0]) == [1]
    assert len(ext[2]) == 0 and len(adj[2]) == 0


def test_extend_blocks_sphere_filter():
    """A cell inside block 0's zone is not taken when its block's bounding
    sphere is farther than R_0 + R_1 + nc sc from block 0's (Extend_blocks.f90:
    59-79): centres 5.38 apart, threshold 1 + 3 + 1 = 5."""
    p = np.array([[0.0, 0, 0], [2.0, 0, 0], [3.0, 1.4, 1.4], [9.0, 1.4, 1.4]])
    blocks = [np.arange(0, 2), np.arange(2, 4)]
    ext, adj = api.extend_blocks(_cells(p), blocks, 1.0, 1)
    assert len(ext[0]) == 0 and len(adj[0]) == 0
    # with a longer extension the blocks become adjacent and the cell is taken
    ext, adj = api.extend_blocks(_cells(p), blocks, 1.0, 2)
    assert list(ext[0]) == [2] and list(adj[0]) == [1]


def test_extend_blocks_edge_cases():
    p = np.random.default_rng(1).normal(size=(30, 3))
    ext, _ = api.extend_blocks(_cells(p), [np.arange(30)], 0.5, 2)
    assert len(ext[0]) == 0
    ext, _ = api.extend_blocks(_cells(p), [np.arange(15), np.arange(15, 30)], 0.5, 0)
    assert all(len(e) == 0 for e in ext)


def test_extend_blocks_nbext_max_stop():
    """Twelve one-cell blocks at one point: each extension would hold the 11
    other cells, above 10 x the largest block (1 cell), where the reference
    stops (Extend_blocks.f90:112-119); the library returns an error."""
    p = np.zeros((12, 3))
    blocks = [np.arange(i, i + 1) for i in range(12)]
    with pytest.raises(api.MidasError):
        api.extend_blocks(_cells(p), blocks, 1.0, 1)
