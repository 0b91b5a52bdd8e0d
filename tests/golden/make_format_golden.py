"""Golden vectors for the reference's Fortran record formats (shape.dat,
Cells.dat, Blocks.dat and S-table records).

Compiles a small Fortran program (written here, not the reference's) that
applies the reference's own edit descriptors -- the READ formats of
Read_ShapeFile (InOut_ReadWrite_Functions.f90:474-490) and Read_CellsFile
(:527-528), and the WRITE formats of Write_geometry_files (:604, :673-679) and
Write_txt_Sfiles (:1034-1042, :982-990) -- with flang, runs it on fixed
inputs and stores what the Fortran runtime reads and writes in
format_vectors.json.  Run from the repository root:

    python tests/golden/make_format_golden.py
"""
import json
import os
import shutil
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))

F_VALUES = [0.0015, -0.0015, 123.456789123, -1e-9, 0.0, 12345.6789, 99999.9999999, 1e6, -1234.5678915, 0.5,
            -0.0000004, 3.14159265, 2.5e-4, 7.0e-5, -42.42424242, 0.99999949, 1.0000005e-3]
E_VALUES = F_VALUES + [1.23456e123, -9.99996e-3, 6.02214076e23, -1.602176634e-19, 1e-100, 0.099999996]
I_VALUES = [0, 1, -1, 42, 9999, 10000, 99999, -9999, 123456, 1234567, 99999999, 100000000]

SHAPE7 = ["      1    1   43    0    1   1   1", "  11508  104   77   35    2   2   2", "     12   -3    4    5    1",
          "      7    1 2 3    1    1", "      3   12   13"]
SHAPE4 = ["    1   43    0    1", "   -5   10  100    3", "  1 2    3    4    5", "    7"]
CELLS = [
    "    0.001500;   -0.001500;    0.000500;    0.000050;       1;     3",
    "    0.001500   -0.001500    0.000500    0.000050       1     3",
    "    0.001500   -0.001500    0.000500    0.000050       1   123",
    "   12.500000;    1.5E-03 ;     1.5-3  ;    1.5D+2  ;      17;    42",
    "        1500;        -250;           0;          50;       2;     1",
    "    0.1",
    "",
]


def fortran_program():
    nf, ne, ni = len(F_VALUES), len(E_VALUES), len(I_VALUES)
    fv = ", ".join(f"{v:.17e}".replace("e", "d") for v in F_VALUES)
    ev = ", ".join(f"{v:.17e}".replace("e", "d") for v in E_VALUES)
    iv = ", ".join(str(v) for v in I_VALUES)
    return f"""
program fmt_golden
  implicit none
  real(8) :: fv({nf}), ev({ne}), r(4)
  integer :: iv({ni}), i, k(7), ios
  character(len=256) :: line
  character(len=200) :: s
  character :: cc
  fv = [{fv}]
  ev = [{ev}]
  iv = [{iv}]
  open(10, file='out.txt')
  do i = 1, {nf}
    write(s, '(f12.6)') fv(i); write(10, '(a)') 'F12.6|' // s(1:12)
    write(s, '(f10.4)') fv(i); write(10, '(a)') 'F10.4|' // s(1:10)
  end do
  do i = 1, {ne}
    write(s, '(e12.4)') ev(i); write(10, '(a)') 'E12.4|' // s(1:12)
  end do
  do i = 1, {ni}
    write(s, '(i8)') iv(i); write(10, '(a)') 'I8|' // s(1:8)
    write(s, '(i6)') iv(i); write(10, '(a)') 'I6|' // s(1:6)
    write(s, '(i5)') iv(i); write(10, '(a)') 'I5|' // s(1:5)
    write(s, '(i4)') iv(i); write(10, '(a)') 'I4|' // s(1:4)
  end do
  write(s, '(f12.6,f12.6,f12.6,f12.6,i8,i6)') fv(1), fv(2), fv(13), fv(14), 1, 3
  write(10, '(a)') 'CELLS|' // s(1:62)
  write(s, '(a,f10.4,a,f10.4,a,f10.4,a,i8)') 'THETA =', 12.5d0, '; PHI =', 270.0d0, '; BETA =', 0.0d0, '; NRX =', 181
  write(10, '(a)') 'SHEAD|' // trim(s)
  write(s, '(f10.4,f10.4,e12.4,e12.4,e12.4,e12.4,e12.4,e12.4,e12.4,e12.4)') 90.5d0, 0.0d0, ev(1), ev(2), ev(3), &
      ev(4), ev(5), ev(6), ev(7), ev(8)
  write(10, '(a)') 'SROW|' // s(1:116)
  write(s, '(f4.2)') 1.36d0; write(10, '(a)') 'F4.2|' // s(1:4)
  write(s, '(f5.2)') 13.6d0; write(10, '(a)') 'F5.2|' // s(1:5)
  write(s, '(f6.2)') 136.0d0; write(10, '(a)') 'F6.2|' // s(1:6)
  write(s, '(a,i4)') '****** Block ', 7; write(10, '(a)') 'BLK|' // trim(s)
  write(s, '(a,i5)') 'Nbc = ', 375; write(10, '(a)') 'BLK|' // trim(s)
  open(11, file='shape7.txt')
  do
    read(11, '(a)', iostat=ios) line
    if (ios /= 0) exit
    k = 0
    read(line, '(i7,i5,i5,i5,i5,i5,i5)') k(1), k(2), k(3), k(4), k(5), k(6), k(7)
    write(10, '(a,7i12)') 'R7|', k
  end do
  close(11)
  open(11, file='shape4.txt')
  do
    read(11, '(a)', iostat=ios) line
    if (ios /= 0) exit
    k = 0
    read(line, '(i5,i5,i5,i5)') k(1), k(2), k(3), k(4)
    write(10, '(a,4i12)') 'R4|', k(1:4)
  end do
  close(11)
  open(11, file='cells.txt')
  do
    read(11, '(a)', iostat=ios) line
    if (ios /= 0) exit
    k = 0
    read(line, '(f12.6,a,f12.6,a,f12.6,a,f12.6,a,i8,a,i6)') r(1), cc, r(2), cc, r(3), cc, r(4), cc, k(1), cc, k(2)
    write(10, '(a,4es26.17e3,2i12)') 'RC|', r, k(1:2)
  end do
  close(11)
  close(10)
end program fmt_golden
"""


def main():
    fc = shutil.which("flang") or "/opt/rocm/llvm/bin/flang"
    d = tempfile.mkdtemp()
    with open(os.path.join(d, "g.f90"), "w") as f:
        f.write(fortran_program())
    for name, recs in (("shape7.txt", SHAPE7), ("shape4.txt", SHAPE4), ("cells.txt", CELLS)):
        with open(os.path.join(d, name), "w") as f:
            f.write("\n".join(recs) + "\n")
    subprocess.run([fc, "-o", "g", "g.f90"], cwd=d, check=True)
    subprocess.run(["./g"], cwd=d, check=True)
    out = {"F_VALUES": F_VALUES, "E_VALUES": E_VALUES, "I_VALUES": I_VALUES, "SHAPE7": SHAPE7, "SHAPE4": SHAPE4,
           "CELLS": CELLS, "lines": open(os.path.join(d, "out.txt")).read().splitlines(),
           "compiler": subprocess.run([fc, "--version"], capture_output=True, text=True).stdout.splitlines()[0]}
    with open(os.path.join(HERE, "format_vectors.json"), "w") as f:
        json.dump(out, f, indent=0)
    shutil.rmtree(d)
    print("wrote", os.path.join(HERE, "format_vectors.json"), len(out["lines"]), "lines")


if __name__ == "__main__":
    main()
