"""The Fortran drop-in run as the reference's own caller would run it: a flang
program (tests/dropin/dropin_driver.f90, built against the reference's
Initialization / common_variables modules) sets the module variables and the
Cell / Dipole / CBFM_Block tables, calls Compute_EFields_MoM,
Compute_EFields_CBFME, Compute_EFields_CBFME_ACA and Calcul_MatrixZij_ACA
through midas_amd/fortran/midas_dropin.f90, and writes what they return; this
test compares that with the oracle.  The CBFM cases play rank 0 of two MPI
ranks that owns blocks 3 and 1 (in that order), so the rank-local layout of
E_total, K_patchs and C_job_patchs (Compute_EFields_CBFME.f90:154-163,
443-444, 990-1006) is checked too."""
import os
import subprocess

import numpy as np
import pytest

from midas_amd import geometry as g

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "dropin", "_build", "dropin_driver")


def _case(name):
    print(f"\n[case] {name}", flush=True)


def _rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300)


@pytest.fixture(scope="module")
def driver():
    if not os.path.exists(DRIVER):
        pytest.skip("drop-in test program not built (tests/dropin/Makefile needs flang and oracle/_ref)")
    return DRIVER


def _run(driver, tmp_path, icase, s, oc, tr, rx=None, blocks=None, ext=None, niws=0, mine=(), counts=(), sr=0,
         kmax=50, eps_aca=1e-4, pair=None):
    f_in, f_out = tmp_path / f"in{icase}.bin", tmp_path / f"out{icase}.bin"
    n = s.n
    nrx = 0 if rx is None else len(rx[0])
    nb = 0 if blocks is None else len(blocks)
    next_ = 0 if ext is None else max(1, max(len(e) for e in ext))
    hdr = np.array([icase, n, len(tr[0]), nrx, nb, next_, niws, sr, kmax, max(len(counts), 1), 0, len(mine)],
                   np.int32)
    with open(f_in, "wb") as f:
        f.write(hdr.tobytes())
        f.write(np.array([s.k0, eps_aca, 0.01], np.float64).tobytes())
        for a in (s.x, s.y, s.z, s.sc, oc.kappa):
            f.write(np.ascontiguousarray(a, np.float64).tobytes())
        for a in (np.sqrt(s.eps), s.eps, oc.che, oc.znnpp):
            f.write(np.ascontiguousarray(a, np.complex128).tobytes())
        f.write(np.arange(1, n + 1, dtype=np.int32).tobytes())
        for a in tr:
            f.write(np.ascontiguousarray(a, np.float64).tobytes())
        if nrx:
            for a in rx:
                f.write(np.ascontiguousarray(a, np.float64).tobytes())
        if nb:
            f.write(np.array([len(b) for b in blocks], np.int32).tobytes())
            f.write(np.array([len(e) for e in ext], np.int32).tobytes())
            tab = np.zeros((nb, next_), np.int32)
            for b, e in enumerate(ext):
                tab[b, :len(e)] = np.asarray(e) + 1
            f.write(np.asfortranarray(tab).tobytes(order="F"))
            f.write(np.asarray(mine, np.int32).tobytes())
            f.write(np.asarray(counts, np.int32).tobytes())
        if pair is not None:
            f.write(np.asarray(pair, np.int32).tobytes())
    print("running", driver, flush=True)
    r = subprocess.run([driver, str(f_in), str(f_out)], capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    return np.fromfile(f_out, np.uint8)


def _take(buf, off, dtype, count):
    a = np.frombuffer(buf[off:off + np.dtype(dtype).itemsize * count].tobytes(), dtype)
    return a, off + a.nbytes


def test_dropin_compute_efields_mom(driver, oracle, tmp_path):
    s = g.sphere(size_param=2.0, nd=6)
    oc = oracle.Cells(s.x, s.y, s.z, s.sc, s.eps, s.k0)
    tth, tph = np.array([0.0, 40.0]), np.array([0.0, 10.0])
    rth, rph = np.arange(0.0, 181.0, 20.0), np.zeros(10)
    _case(f"drop-in Compute_EFields_MoM n={3 * s.n}")
    buf = _run(driver, tmp_path, 1, s, oc, (tth, tph), (rth, rph))
    flag, off = _take(buf, 0, np.int32, 1)
    S, off = _take(buf, off, np.complex128, len(rth) * 8)
    ce, off = _take(buf, off, np.complex128, 2)
    ca, off = _take(buf, off, np.complex128, 2)
    S = S.reshape((len(rth), 8), order="F")
    E = np.linalg.solve(oracle.mom_fill(oc), oracle.mom_rhs(oc, tth, tph))
    So = oracle.scattering_matrix(oc, E, rth, rph)
    ceo, cao = oracle.cext_cabs(oc, E, tth, tph)
    print("GPU context", flag, "S err", _rel(S, So), "Cext err", _rel(ce, ceo), "Cabs err", _rel(ca, cao))
    assert flag[0] == 1
    assert _rel(S, So) < 1e-10 and _rel(ce, ceo) < 1e-10 and _rel(ca, cao) < 1e-10


@pytest.fixture(scope="module")
def dlca():
    return g.dlca_aggregate(nmono=4, cells_per_mono=60, seed=3)


@pytest.mark.parametrize("icase", [2, 3], ids=["CBFME", "CBFME_ACA"])
def test_dropin_compute_efields_cbfme(driver, oracle, dlca, tmp_path, icase):
    s = dlca
    oc = oracle.Cells(s.x, s.y, s.z, s.sc, s.eps, s.k0)
    tth, tph = np.array([0.0, 35.0]), np.array([0.0, 80.0])
    niws = 40
    mine, counts = [3, 1], [2, 2]           # rank 0 of 2 owns blocks 3 and 1
    _case(f"drop-in Compute_EFields_{'CBFME' if icase == 2 else 'CBFME_ACA'} n={3 * s.n} blocks={len(s.blocks)}")
    buf = _run(driver, tmp_path, icase, s, oc, (tth, tph), blocks=s.blocks, ext=s.ext, niws=niws, mine=mine,
               counts=counts)
    rows = np.concatenate([(3 * np.asarray(s.blocks[b - 1])[:, None] + np.arange(3)).ravel() for b in mine])
    flag, off = _take(buf, 0, np.int32, 1)
    E, off = _take(buf, off, np.complex128, len(rows) * 4)
    E = E.reshape((len(rows), 4), order="F")
    th_c, ph_c = oracle.transmitters_un(niws)
    Eo, Ko, Cs = oracle.cbfm(oc, s.blocks, s.ext, th_c, ph_c, tth, tph, use_aca=icase == 3)
    print("GPU context", flag, "E err (rank-local rows)", _rel(E, Eo[rows]))
    assert flag[0] == 1
    assert _rel(E, Eo[rows]) < 1e-9
    if icase == 2:
        K, off = _take(buf, off, np.int32, len(mine))
        kc, off = _take(buf, off, np.int32, 1)
        Cj, off = _take(buf, off, np.complex128, len(rows) * int(kc[0]))
        Cj = Cj.reshape((len(rows), int(kc[0])), order="F")
        print("K_patchs", list(K), "oracle", [Ko[b - 1] for b in mine])
        assert list(K) == [Ko[b - 1] for b in mine]
        r0 = 0
        for b, k in zip(mine, K):
            nr = 3 * len(s.blocks[b - 1])
            Cg = Cj[r0:r0 + nr, :k]
            coef, *_ = np.linalg.lstsq(Cg, Cs[b - 1], rcond=None)
            assert _rel(Cg @ coef, Cs[b - 1]) < 1e-9
            r0 += nr


def test_dropin_calcul_matrixzij_aca(driver, oracle, dlca, tmp_path):
    s = dlca
    oc = oracle.Cells(s.x, s.y, s.z, s.sc, s.eps, s.k0)
    i, j = 0, 2
    bi, bj = s.blocks[i], s.blocks[j]
    kmax = 50
    _case(f"drop-in Calcul_MatrixZij_ACA blocks {i + 1},{j + 1} ({len(bi)} x {len(bj)} cells)")
    pair = [i + 1, j + 1, 3 * int(bi[0]) + 1, 3 * int(bj[0]) + 1, 3 * len(bi), 3 * len(bj)]
    buf = _run(driver, tmp_path, 4, s, oc, (np.zeros(1), np.zeros(1)), blocks=s.blocks, ext=s.ext, mine=[1],
               counts=[1], kmax=kmax, pair=pair)
    flag, off = _take(buf, 0, np.int32, 1)
    k, off = _take(buf, off, np.int32, 1)
    U, off = _take(buf, off, np.complex128, 3 * len(bi) * kmax)
    V, off = _take(buf, off, np.complex128, kmax * 3 * len(bj))
    U = U.reshape((3 * len(bi), kmax), order="F")
    V = V.reshape((kmax, 3 * len(bj)), order="F")
    ko, Uo, Vo, _, _ = oracle.aca(oc, int(bi[0]), len(bi), int(bj[0]), len(bj), kmax=kmax, eps=1e-4)
    kk = min(int(k[0]), kmax)
    print("GPU context", flag, "nb_iter_out", k, "oracle", ko, "UV err", _rel(U[:, :kk] @ V[:kk], Uo @ Vo))
    assert flag[0] == 1
    assert int(k[0]) == ko
    assert _rel(U[:, :kk] @ V[:kk], Uo @ Vo) < 1e-9
