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


# ---------------------------------------------------------------------------
# Division into blocks: Division_blocks (Division_blocks.f90), SPH and CSH.
# The file uses the MPI module, so there is no reference-compiled vector.
# Parity is pinned by (1) the oracle's independent restatement on several
# geometries, exact integer orders and sizes; (2) lattices whose division has
# a closed form; (3) invariants (a permutation, contiguous blocks, hBlock =
# 2 x the largest block radius); (4) the reference's failure case.
# ---------------------------------------------------------------------------

def _lattice(nx, ny, nz, sc=1.0):
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ijk = np.stack([i.ravel(), j.ravel(), k.ravel()], 1)
    return ijk, (ijk + 0.5) * sc


@pytest.mark.parametrize("geom", ["sphere", "dlca", "hull"])
@pytest.mark.parametrize("navg", [40, 150, 600])
def test_partition_matches_oracle(oracle, geom, navg):
    s = {"sphere": lambda: g.sphere(2.0, 14), "dlca": lambda: g.dlca_aggregate(nmono=8, cells_per_mono=120, seed=3),
         "hull": lambda: g.convex_hull_particle(5.0, target_cells=2500)}[geom]()
    c = _cells(np.stack([s.x, s.y, s.z], 1))
    for best in (0, 1):
        o1, n1, h1 = api.partition_sph(c, navg, best)
        o2, n2, h2 = oracle.partition_sph(s.x, s.y, s.z, navg, best)
        assert np.array_equal(o1, o2) and np.array_equal(n1, n2) and h1 == h2, (geom, navg, best)
    for type_s in (2, 3):
        o1, n1, h1 = api.partition_csh(c, float(s.sc[0]), navg, type_s)
        o2, n2, h2 = oracle.partition_csh(s.x, s.y, s.z, float(s.sc[0]), navg, type_s)
        assert np.array_equal(o1, o2) and np.array_equal(n1, n2) and h1 == h2, (geom, navg, type_s)


def _invariants(p, order, sizes, hblock, input_order=True):
    n = len(p)
    assert np.array_equal(np.sort(order), np.arange(n))
    assert sizes.sum() == n and np.all(sizes > 0)
    starts = np.concatenate([[0], np.cumsum(sizes)])
    q = p[order]
    fr = api.block_frames(_cells(q), [np.arange(starts[b], starts[b + 1]) for b in range(len(sizes))])
    assert hblock == 2.0 * fr[:, 12].max()
    # SPH: cells keep their input order inside a block (CSH: the order the
    # previous trial division left)
    for b in range(len(sizes) if input_order else 0):
        assert np.all(np.diff(order[starts[b]:starts[b + 1]]) > 0)


@pytest.mark.parametrize("navg,nblocks", [(512, 8), (64, 64), (600, 8), (4096, 8)])
def test_partition_sph_lattice_cube(navg, nblocks):
    """16^3 lattice cube: the octree cuts at the lattice midplanes.  N = 4096
    splits (|4096 - navg| > |512 - navg|) into 8 octants of 512; an octant
    splits again only when |512 - navg| > |64 - navg|.  Blocks are ordered by
    octant 1 + ix + 2 iy + 4 iz, then sub-octant, then input index."""
    ijk, p = _lattice(16, 16, 16)
    order, sizes, hb = api.partition_sph(_cells(p), navg)
    top = (ijk[:, 0] >= 8) + 2 * (ijk[:, 1] >= 8) + 4 * (ijk[:, 2] >= 8)
    sub = ((ijk % 8)[:, 0] >= 4) + 2 * ((ijk % 8)[:, 1] >= 4) + 4 * ((ijk % 8)[:, 2] >= 4)
    key = top * 8 + (sub if nblocks == 64 else 0)
    want = np.lexsort((np.arange(len(p)), key))
    assert len(sizes) == nblocks and np.all(sizes == 4096 // nblocks)
    assert np.array_equal(order, want)
    _invariants(p, order, sizes, hb)


def test_partition_csh_box():
    """40 x 20 x 10 lattice box (sc = 1): with navg = 4000 (at most 2 navg
    cells) the box is cut once across its longest side into two 20 x 20 x 10
    halves; with navg = 2000 each half (4000 > 2000, |2000 - 4000| >
    |2000 - 2000|) is cut across its next longest side, y, at y = 10 (the
    balancing search keeps the first, exact cut): 4 blocks of 2000.  type_s
    = 3 (cylinder) keeps the slabs only, so the trials go on: hBlock 40, 20,
    10 give 2, 3 and 5 slabs (widths nint(40 / div) = 20, 13, 8) and mean
    sizes 4000, 2666, 1600; the last is closest to 2000: 5 slabs of 8."""
    ijk, p = _lattice(40, 20, 10)
    idx = np.arange(len(p))
    order, sizes, hb = api.partition_csh(_cells(p), 1.0, 4000)
    assert list(sizes) == [4000, 4000]
    assert np.array_equal(order, np.lexsort((idx, ijk[:, 0] >= 20)))
    _invariants(p, order, sizes, hb)
    order, sizes, hb = api.partition_csh(_cells(p), 1.0, 2000)
    assert list(sizes) == [2000] * 4
    assert np.array_equal(order, np.lexsort((idx, ijk[:, 1] >= 10, ijk[:, 0] >= 20)))
    _invariants(p, order, sizes, hb)
    order, sizes, _ = api.partition_csh(_cells(p), 1.0, 2000, type_s=3)
    assert list(sizes) == [1600] * 5
    assert np.array_equal(order, np.lexsort((idx, ijk[:, 0] // 8)))


@pytest.mark.parametrize("navg", [30, 200])
def test_partition_invariants_dlca(navg):
    s = g.dlca_aggregate(nmono=6, cells_per_mono=60, seed=9)
    p = np.stack([s.x, s.y, s.z], 1)
    _invariants(p, *api.partition_sph(_cells(p), navg))
    _invariants(p, *api.partition_sph(_cells(p), navg, 1))
    _invariants(p, *api.partition_csh(_cells(p), float(s.sc[0]), navg), input_order=False)


def test_partition_csh_single_cell_fails(oracle):
    """One cell: the box gives two slabs, one empty, so the reference's
    smallest block size is 0 and its one-cell block has no valid cut (its
    balancing search starts from 0/0): the division fails."""
    p = np.array([[0.5, 0.5, 0.5]])
    with pytest.raises(api.MidasError):
        api.partition_csh(_cells(p), 1.0, 10)
    with pytest.raises(ValueError):
        oracle.partition_csh(p[:, 0], p[:, 1], p[:, 2], 1.0, 10)
    order, sizes, hb = api.partition_sph(_cells(p), 10)
    assert list(order) == [0] and list(sizes) == [1] and hb == 0.0
