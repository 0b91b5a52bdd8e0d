"""Readers and writers of the reference's files (midas_amd/shape_io.py).

Pinned by golden vectors of the Fortran runtime itself: tests/golden/
make_format_golden.py compiles a program applying the reference's edit
descriptors (Read_ShapeFile, Read_CellsFile, Write_geometry_files,
Write_txt_Sfiles) with flang and records what it reads and writes; and by the
reference's own input deck examples/Run_Sims_p0004 (shape.dat, parsed
independently into tests/golden/deck_p0004.npz)."""
import json
import os

import numpy as np
import pytest

from midas_amd import api, geometry as g, shape_io as io

GOLD = os.path.join(os.path.dirname(__file__), "golden")
REF_DECKS = "/root/reference/examples"


@pytest.fixture(scope="module")
def fv():
    with open(os.path.join(GOLD, "format_vectors.json")) as f:
        d = json.load(f)
    by = {}
    for line in d["lines"]:
        tag, _, rest = line.partition("|")
        by.setdefault(tag, []).append(rest)
    d["by"] = by
    return d


@pytest.fixture(scope="module")
def deck():
    return np.load(os.path.join(GOLD, "deck_p0004.npz"), allow_pickle=False)


def test_edit_descriptor_output(fv):
    by = fv["by"]
    for v, f12, f10 in zip(fv["F_VALUES"], by["F12.6"], by["F10.4"]):
        assert io.fmt_f(v, 12, 6) == f12, v
        assert io.fmt_f(v, 10, 4) == f10, v
    for v, e in zip(fv["E_VALUES"], by["E12.4"]):
        assert io.fmt_e(v, 12, 4) == e, v
    for v, i8, i6, i5, i4 in zip(fv["I_VALUES"], by["I8"], by["I6"], by["I5"], by["I4"]):
        assert (io.fmt_i(v, 8), io.fmt_i(v, 6), io.fmt_i(v, 5), io.fmt_i(v, 4)) == (i8, i6, i5, i4), v
    assert (io.fmt_f(1.36, 4, 2), io.fmt_f(13.6, 5, 2), io.fmt_f(136.0, 6, 2)) == \
        (by["F4.2"][0], by["F5.2"][0], by["F6.2"][0])


def test_record_writers(fv):
    by, F, E = fv["by"], fv["F_VALUES"], fv["E_VALUES"]
    assert io.cells_file_lines([F[0]], [F[1]], [F[12]], [F[13]], [1], [3]) == [by["CELLS"][0]]
    S4 = np.array([[E[0] + 1j * E[1], E[2] + 1j * E[3], E[4] + 1j * E[5], E[6] + 1j * E[7]]])
    lines = io.s_table_lines(12.5, 270.0, 181, [90.5], [0.0], S4)
    assert lines[0] == by["SHEAD"][0]
    assert lines[2] == by["SROW"][0]
    blk = io.blocks_file_lines([375] * 7, [[]] * 6 + [[4, 9]])
    assert blk[-7:-3] == [by["BLK"][0], by["BLK"][1], "Nbc_ext =     2", "Num_cells_ext ="]
    assert blk[-3:] == ["       5", "      10", " "]


def test_record_readers(fv):
    by = fv["by"]
    for rec, want in zip(fv["SHAPE7"], by["R7"]):
        assert io.read_fixed(rec, io._SHAPE_FMT[7]) == [int(t) for t in want.split()], rec
    for rec, want in zip(fv["SHAPE4"], by["R4"]):
        assert io.read_fixed(rec, io._SHAPE_FMT[4]) == [int(t) for t in want.split()], rec
    for rec, want in zip(fv["CELLS"], by["RC"]):
        got = io.read_fixed(rec, io._CELLS_READ_FMT)
        w = want.split()
        assert got[:4] == [float(t.replace("E", "e")) for t in w[:4]], rec
        assert got[4:] == [int(t) for t in w[4:]], rec


def test_s_table_name():
    assert io.s_table_name(13.6e9, 9, "GHz", 1, "CBFM-E  ") == "Smtable_13.60GHz_kt0001_CBFM-E.dat"
    assert io.s_table_name(1.36e9, 9, "GHz", 12, "MoM     ") == "Smtable_1.36GHz_kt0012_MoM.dat"
    assert io.s_table_name(136e6, 6, "MHz", 3, "RGE", nfreq=2, num_freq=2) == "Sim2_Smtable_136.00MHz_kt0003_RGE.dat"


def test_shape_file_deck(tmp_path, deck):
    """The reader on the reference's p0004 deck: header inspection (7 header
    lines, 7 columns), the count from line 2, and every record equal to the
    whitespace-split parse.  The deck is rebuilt from the fixture (header
    lines kept, records in the DDSCAT (i7, 6 i5) layout the file uses)."""
    lat = deck["lattice"].astype(np.int64)
    path = tmp_path / "shape.dat"
    with open(path, "w") as f:
        f.write("\n".join(str(h) for h in deck["header"]) + "\n")
        for i, r in enumerate(lat):
            f.write(f"{i + 1:7d}" + "".join(f"{v:5d}" for v in r[:3]) + "".join(f"{v:4d}" for v in r[3:]) + "\n")
    assert io.inspect_shape_file(path) == (7, 7)
    n, got = io.read_shape_file(path)
    assert n == int(deck["count"]) == 11508
    assert np.array_equal(got, lat)


@pytest.mark.skipif(not os.path.isdir(REF_DECKS), reason="reference decks not present")
@pytest.mark.parametrize("name", ["Run_Sims_p0004", "Run_Sims_a0006"])
def test_shape_file_reference_decks(name):
    path = os.path.join(REF_DECKS, name, "inputs", "shape.dat")
    n, lat = io.read_shape_file(path)
    lines = open(path).read().splitlines()
    want = np.array([[int(t) for t in ln.split()[1:7]] for ln in lines[7:7 + n]])
    assert lat.shape == (n, 6) and np.array_equal(lat, want)


def test_cells_from_shape(deck):
    """Sc = (4 pi / (3 N))**(1/3) a: N Sc^3 is the equivalent sphere's volume
    (exactly without the quirks, to the single-precision exponent with them);
    the p0004 deck's Sc is the deck's nominal 50 um."""
    lat = deck["lattice"]
    a = float(deck["a_eff_m"])
    x, y, z, sc, m = io.cells_from_shape(lat, a, quirks=0)
    n = len(lat)
    assert abs(n * sc[0] ** 3 - 4 * np.pi / 3 * a ** 3) < 1e-14 * a ** 3
    assert np.allclose(x, sc[0] * (lat[:, 0] + 0.5), rtol=1e-15, atol=0)
    assert np.array_equal(m, lat[:, 3])
    xq, _, _, scq, _ = io.cells_from_shape(lat, a)
    assert abs(scq[0] / sc[0] - 1) < 1e-7 and scq[0] != sc[0]
    assert abs(sc[0] / float(deck["sc_nominal_m"]) - 1) < 1e-3


def test_cells_file_round_trip(tmp_path):
    """Cells.dat written by the reference's writer and read back by its
    reader: the writer puts no separators, the reader skips one character
    after each field, so every field is read one to four columns late.  For
    positions and sizes under 100 (the leading blanks absorb the shift) the
    values come back to the 1e-6 grid; a block number of 100 or more runs into
    the material index (here 1 -> 11), as the reference reads it."""
    rng = np.random.default_rng(3)
    n = 50
    x, y, z = rng.uniform(-20, 20, (3, n)).round(6)
    sc = np.full(n, 0.05)
    mi = rng.integers(1, 4, n)
    blk = np.where(np.arange(n) < 40, rng.integers(1, 99, n), 123)
    path = tmp_path / "Cells.dat"
    io.write_cells_file(path, x, y, z, sc, mi, blk)
    rx, ry, rz, rsc, rm, sc_min = io.read_cells_file(path, lamb_mag=0)
    assert np.allclose(rx, x, atol=1e-12) and np.allclose(ry, y, atol=1e-12) and np.allclose(rz, z, atol=1e-12)
    assert np.allclose(rsc, sc, atol=1e-12) and sc_min == rsc.min()
    assert np.array_equal(rm[:40], mi[:40])
    assert np.array_equal(rm[40:], mi[40:] * 10 + 1)
    # millimetre files: the single-precision 10.**(-3) scale
    rx3 = io.read_cells_file(path, lamb_mag=3)[0]
    assert np.array_equal(rx3, float(np.float32(1e-3)) * rx)


def test_deck_partition_and_extension(deck):
    """p0004 through the reference's pipeline on the host: cells from the
    lattice, CSH division with the deck's NBlock (avgNcells) = 2000 (library =
    oracle, exactly), then the extension of every block."""
    from oracle import oracle as o
    x, y, z, sc, m = io.cells_from_shape(deck["lattice"], float(deck["a_eff_m"]))
    cells = api.CellTable(x, y, z, sc, np.full(len(x), complex(deck["m"]) ** 2))
    navg = int(deck["navg_cells"])
    order, sizes, hb = api.partition_csh(cells, float(sc[0]), navg)
    oo, os_, ohb = o.partition_csh(x, y, z, float(sc[0]), navg)
    assert np.array_equal(order, oo) and np.array_equal(sizes, os_) and hb == ohb
    assert sizes.sum() == len(x) and len(sizes) >= 4
    s = g.Scatterer(x, y, z, sc, np.full(len(x), complex(deck["m"]) ** 2), 2 * np.pi * float(deck["freq_hz"]) / 299792458.0)
    t, hb2 = g.divide_blocks(s, "CSH", navg)
    assert hb2 == hb and [len(b) for b in t.blocks] == list(sizes)
    assert all(len(e) > 0 for e in t.ext)
