"""Readers and writers of the reference's geometry and result files (host code).

  read_shape_file    Read_ShapeFile (InOut_ReadWrite_Functions.f90:439-493) with
                     Inspect_ShapeFile (DiverseUtil.f90:273-358): DDSCAT shape.dat
  cells_from_shape   cell centres and size from the lattice (Discretization.f90:355-375)
  read_cells_file    Read_CellsFile (InOut_ReadWrite_Functions.f90:495-532) and the
                     'cells' route of Discretization (Discretization.f90:77-97)
  write_cells_file   Cells.dat of Write_geometry_files (InOut_ReadWrite_Functions.f90:593-607)
  write_blocks_file  Blocks.dat of Write_geometry_files (InOut_ReadWrite_Functions.f90:668-681)
  s_table_name,      Smtable_<f><unit>_kt<k>_<method>.dat of Write_txt_Sfiles
  write_s_table      (InOut_ReadWrite_Functions.f90:950-1047)

Records are read and written with the reference's Fortran edit descriptors:
fixed-width fields in which blanks are ignored and an all-blank field is
zero (formatted input, BLANK='NULL', short records padded), F-fields without
a decimal point scaled by 10**-d, and Fw.d / Ew.d / Iw output with the
asterisk fill on overflow.
"""
from __future__ import annotations

import re

import numpy as np

from . import api

# ---------------------------------------------------------------------------
# Fortran edit descriptors
# ---------------------------------------------------------------------------

_REAL = re.compile(r"^([+-]?)(\d*)(?:\.(\d*))?(?:[EeDdQq]([+-]?\d+)|([+-]\d+))?$")


def _field(line, pos, width):
    return line[pos:pos + width]


def f_int(field):
    """Iw input: blanks ignored, all blank = 0."""
    s = field.replace(" ", "")
    if not s:
        return 0
    return int(s)


def f_real(field, d):
    """Fw.d input: blanks ignored, all blank = 0; without a decimal point the
    digits carry d implied decimals; an exponent may follow with or without
    its letter."""
    s = field.replace(" ", "")
    if not s:
        return 0.0
    m = _REAL.match(s)
    if not m or (not m.group(2) and not m.group(3)):
        raise ValueError(f"bad real field {field!r}")
    sign, ip, fp, e1, e2 = m.groups()
    exp = int(e1 or e2 or 0)
    if fp is None:   # no decimal point: d implied decimals
        v = float(f"{ip or '0'}e{exp - d}")
    else:
        v = float(f"{ip or '0'}.{fp or '0'}e{exp}")
    return -v if sign == "-" else v


def read_fixed(line, fmt):
    """Read one record with a list of ('i', w), ('f', w, d) or ('a', w)
    descriptors; returns the values of the I and F fields."""
    out, pos = [], 0
    for f in fmt:
        w = f[1]
        s = _field(line, pos, w).ljust(w)
        if f[0] == "i":
            out.append(f_int(s))
        elif f[0] == "f":
            out.append(f_real(s, f[2]))
        pos += w
    return out


def fmt_i(v, w):
    s = f"{int(v):d}"
    return s.rjust(w) if len(s) <= w else "*" * w


def fmt_f(v, w, d):
    s = f"{float(v):.{d}f}"
    if len(s) > w and s.startswith(("0.", "-0.")):   # the optional leading zero goes first
        s = s.replace("0.", ".", 1)
    return s.rjust(w) if len(s) <= w else "*" * w


def fmt_e(v, w, d):
    """Ew.d output: 0.ddddE+xx (0.dddd+xxx beyond two exponent digits)."""
    v = float(v)
    if v == 0.0:
        mant, exp, neg = "0" * d, 0, np.signbit(v)
    else:
        neg = v < 0
        digits, e = f"{abs(v):.{d - 1}e}".split("e")
        mant = digits.replace(".", "")
        exp = int(e) + 1
    ex = f"E{'+' if exp >= 0 else '-'}{abs(exp):02d}" if abs(exp) <= 99 else f"{'+' if exp >= 0 else '-'}{abs(exp):03d}"
    s = f"{'-' if neg else ''}0.{mant}{ex}"
    if len(s) > w:   # the optional leading zero goes first
        s = s.replace("0.", ".", 1)
    return s.rjust(w) if len(s) <= w else "*" * w


# ---------------------------------------------------------------------------
# shape.dat (DDSCAT)
# ---------------------------------------------------------------------------

def _is_header(line):
    """IsHeader (DiverseUtil.f90:330-358): blank, or holding an ASCII letter."""
    t = line.rstrip(" ")
    return len(t) == 0 or any(("A" <= ch <= "Z") or ("a" <= ch <= "z") for ch in t)


def _count_columns(line):
    """CountColumns (DiverseUtil.f90:305-328): blank-separated fields."""
    n, inside = 0, False
    for ch in line.rstrip(" "):
        if ch != " " and not inside:
            n, inside = n + 1, True
        elif ch == " ":
            inside = False
    return n


def inspect_shape_file(path):
    """Inspect_ShapeFile: (header lines before the first data line, columns of
    that line)."""
    nheader, ncols = 0, 0
    with open(path, "r", errors="replace") as f:
        for raw in f:
            line = raw.rstrip("\n").rstrip("\r")[:512]
            if _is_header(line):
                nheader += 1
            else:
                ncols = _count_columns(line)
                break
    return nheader, ncols


_SHAPE_FMT = {
    7: [("i", 7)] + [("i", 5)] * 6,   # num, i, j, k, m_i, m_j, m_k
    5: [("i", 7)] + [("i", 5)] * 4,   # num, i, j, k, m
    4: [("i", 5)] * 4,                # i, j, k, m
}


def read_shape_file(path):
    """Read_ShapeFile: (number of cells, lattice) with lattice (N, 6) for the
    7-column layout (i, j, k, m_x, m_y, m_z) or (N, 4) for the 5- and
    4-column ones (i, j, k, m).  The count is the first integer of line 2
    (list-directed); the records follow the header lines."""
    nheader, ncols = inspect_shape_file(path)
    if ncols not in _SHAPE_FMT:
        raise ValueError(f"{path}: {ncols} columns on the first data line; the reader takes 7, 5 or 4")
    with open(path, "r", errors="replace") as f:
        lines = f.read().splitlines()
    count = int(lines[1].replace(",", " ").split()[0])
    first = max(nheader, 2)
    recs = lines[first:first + count]
    if len(recs) < count:
        raise ValueError(f"{path}: {count} cells announced, {len(recs)} records")
    fmt = _SHAPE_FMT[ncols]
    vals = np.array([read_fixed(r, fmt) for r in recs], dtype=np.int64)
    lat = vals[:, 1:] if ncols in (7, 5) else vals
    return count, lat


def cells_from_shape(lattice, a_eff, quirks=api.Q_ALL):
    """Cells of a shape.dat lattice (Discretization.f90:355-375): Sc =
    (4 pi / (3 N))**(1/3) a_eff (a_eff: volume-equivalent radius in metres),
    centres Sc * (i, j, k) + Sc/2, material index from the 4th column.  With
    Q_PI the reference's single-precision pi, with Q_LITERALS its
    single-precision 1./3. exponent.  Returns (x, y, z, sc, m_index)."""
    lat = np.asarray(lattice)
    nbc = len(lat)
    pi = float(np.float32(np.arccos(np.float32(-1.0)))) if quirks & api.Q_PI else float(np.pi)
    third = float(np.float32(1.0) / np.float32(3.0)) if quirks & api.Q_LITERALS else 1.0 / 3.0
    sc = ((4 * pi) / (3 * nbc)) ** third * a_eff
    x = sc * lat[:, 0] + sc / 2.0
    y = sc * lat[:, 1] + sc / 2.0
    z = sc * lat[:, 2] + sc / 2.0
    return x, y, z, np.full(nbc, sc), lat[:, 3].astype(np.int64)


# ---------------------------------------------------------------------------
# Cells.dat, Blocks.dat
# ---------------------------------------------------------------------------

# Read_CellsFile's record: x, y, z, Sc (f12.6) and the material and block
# indices (i8, i6), each field followed by a one-character separator
_CELLS_READ_FMT = [("f", 12, 6), ("a", 1), ("f", 12, 6), ("a", 1), ("f", 12, 6), ("a", 1), ("f", 12, 6), ("a", 1),
                   ("i", 8), ("a", 1), ("i", 6)]


def read_cells_file(path, lamb_mag=0, ref32=True):
    """Read_CellsFile and the 'cells' route of Discretization: every record
    (blank ones included) is a cell; positions and sizes are multiplied by
    10.**(-lamb_mag) (a single-precision power with ref32, as the reference
    writes it).  Returns (x, y, z, sc, m_index, SimScatterer%Sc = min sc).
    The separators the reader expects are absent from the files
    write_cells_file (the reference's writer) produces; the fields are then
    read across the column boundaries exactly as the reference reads them."""
    with open(path, "r", errors="replace") as f:
        lines = f.read().splitlines()
    vals = [read_fixed(r, _CELLS_READ_FMT) for r in lines]
    a = np.array([v[:4] for v in vals], dtype=np.float64).reshape(-1, 4)
    m = np.array([v[4] for v in vals], dtype=np.int64)
    scale = float(np.float32(10.0) ** np.float32(-lamb_mag)) if ref32 else 10.0 ** (-lamb_mag)
    x, y, z, sc = (scale * a[:, i] for i in range(4))
    return x, y, z, sc, m, float(sc.min()) if len(sc) else 0.0


def cells_file_lines(x, y, z, sc, m_index, block_index):
    """Cells.dat records, '(f12.6,f12.6,f12.6,f12.6,i8,i6)': centre, size,
    material index and 1-based block number of every cell."""
    return [fmt_f(x[i], 12, 6) + fmt_f(y[i], 12, 6) + fmt_f(z[i], 12, 6) + fmt_f(sc[i], 12, 6) +
            fmt_i(m_index[i], 8) + fmt_i(block_index[i], 6) for i in range(len(x))]


def write_cells_file(path, x, y, z, sc, m_index, block_index):
    with open(path, "w") as f:
        for line in cells_file_lines(x, y, z, sc, m_index, block_index):
            f.write(line + "\n")


def blocks_file_lines(block_sizes, ext):
    """Blocks.dat records: per block its number, size, extension size and the
    extension cells as 1-based indices (one '(i8)' line each)."""
    out = []
    for b, n in enumerate(block_sizes):
        out.append("****** Block " + fmt_i(b + 1, 4))
        out.append("Nbc = " + fmt_i(n, 5))
        out.append("Nbc_ext = " + fmt_i(len(ext[b]), 5))
        out.append("Num_cells_ext =")
        out.extend(fmt_i(int(c) + 1, 8) for c in ext[b])
        out.append(" ")
    return out


def write_blocks_file(path, block_sizes, ext):
    with open(path, "w") as f:
        for line in blocks_file_lines(block_sizes, ext):
            f.write(line + "\n")


# ---------------------------------------------------------------------------
# S tables
# ---------------------------------------------------------------------------

def s_table_name(freq, freq_mag, freq_unit, kt, method, nfreq=1, num_freq=1, eqsph=False):
    """File name of Write_txt_Sfiles for transmitter kt (1-based): the
    frequency freq / 10.**freq_mag (single-precision power) as f4.2 / f5.2 /
    f6.2 by its rounded size, the 'kt%04d' tag and the method name ('CBFM-E',
    'MoM' or 'RGE'); 'Sim<k>_' in front when several frequencies run."""
    scaled = freq / float(np.float32(10.0) ** np.float32(freq_mag))
    a = int(np.floor(abs(scaled) + 0.5))
    w = 4 if a < 10 else (5 if a < 100 else 6)
    prefix = "" if nfreq == 1 else f"Sim{num_freq}_"
    meth = method.strip()[:6] if method.strip().startswith("CBFM") else method.strip()[:3]
    return f"{prefix}{'SmtableES_' if eqsph else 'Smtable_'}{fmt_f(scaled, w, 2)}{freq_unit}_kt{kt:04d}_{meth}.dat"


_S_COLS = ("      theta       phi       Re(Svv)        Im(Svv)         Re(Svh)       Im(Svh) "
           "        Re(Shv)       Im(Shv)        Re(Shh)        Im(Shh) ")


def s_table_lines(tr_theta, tr_phi, nrx, rx_theta, rx_phi, S4):
    """One transmitter's table: the header '(a,f10.4,a,f10.4,a,f10.4,a,i8)'
    (beta 0, NRX = nrx), the column titles, then per receiver
    '(f10.4,f10.4,8e12.4)' of theta, phi and Re/Im of Svv, Svh, Shv, Shh.
    S4: (NRx_tot, 4) complex, the transmitter's four columns of S_total."""
    out = ["THETA =" + fmt_f(tr_theta, 10, 4) + "; PHI =" + fmt_f(tr_phi, 10, 4) + "; BETA =" + fmt_f(0.0, 10, 4) +
           "; NRX =" + fmt_i(nrx, 8), _S_COLS]
    S4 = np.asarray(S4)
    for r in range(S4.shape[0]):
        vals = []
        for c in range(4):
            vals += [S4[r, c].real, S4[r, c].imag]
        out.append(fmt_f(rx_theta[r], 10, 4) + fmt_f(rx_phi[r], 10, 4) + "".join(fmt_e(v, 12, 4) for v in vals))
    return out


def write_s_table(path, tr_theta, tr_phi, nrx, rx_theta, rx_phi, S4):
    with open(path, "w") as f:
        for line in s_table_lines(tr_theta, tr_phi, nrx, rx_theta, rx_phi, S4):
            f.write(line + "\n")
