"""Host-side mirror of the MIDAS hot-path interfaces over libmidas_gpu.so.

Names follow the reference routines they replace:
  compute_efields_mom      Compute_EFields_MoM        (Compute_EFields_MoM.f90:1-19)
  compute_efields_cbfme    Compute_EFields_CBFME(_ACA) (Compute_EFields_CBFME.f90:1-21,
                                                        Compute_EFields_CBFME_ACA.f90:1-18)
  calcul_matrixzij_aca     Calcul_MatrixZij_ACA       (Calcul_ACA_matrices.f90:6-20)
  calcul_genmatrix_aca     Calcul_GenMatrix_ACA       (Calcul_ACA_matrices.f90:222-236)
  set_cells_params         SetCellsParams             (SetCellsParams.f90:1-50)
  green_s_tr_partial       Green_s_tr_partial         (Green_functions.f90:96-173)
  incident_field           Incident_Field             (Incident_Field.f90:1-63)
plus the building blocks (mom_fill, solve, scattering_matrix, cext_cabs).

Every call runs on the GPU through the C-ABI declared in include/midas_gpu.h;
there is no CPU fallback: if the library is missing or no GPU is present the
calls raise.  Arrays are numpy, complex128 where the reference uses
COMPLEX(real64), matrices Fortran-ordered.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libmidas_gpu.so")

Q_PI, Q_LITERALS, Q_CEXT_SUM, Q_FP32_BLOCKS = 1, 2, 4, 8
Q_ALL = 15

SOLVE_FP64, SOLVE_C32_IR, SOLVE_F16_IR, SOLVE_F16X3_IR, SOLVE_BF16_IR, SOLVE_BF16X3_IR = range(6)
# refinement outcome (midas_solve_info_t.status) and the no_fallback error code
STATUS_CONVERGED, STATUS_NOT_CONVERGED, STATUS_FP64_FALLBACK = 0, 1, 2
ERR_NOT_CONVERGED = -4
MODE_NAMES = {SOLVE_FP64: "fp64", SOLVE_C32_IR: "c32+ir", SOLVE_F16_IR: "f16+ir", SOLVE_F16X3_IR: "f16x3+ir",
              SOLVE_BF16_IR: "bf16+ir", SOLVE_BF16X3_IR: "bf16x3+ir"}

EXPORTS = (
    "midas_init", "midas_finalize", "midas_last_error", "midas_stream", "midas_cell_params", "midas_mom_fill",
    "midas_block_fill", "midas_plane_wave", "midas_spherical_wave", "midas_solve", "midas_scattering_matrix", "midas_cext_cabs",
    "midas_compute_efields_mom", "midas_aca_block", "midas_aca_matrix", "midas_compute_efields_cbfm",
    "midas_dev_mom_fill", "midas_dev_mom_solve", "midas_dev_compute_efields_cbfm", "midas_profile", "midas_profile_query", "midas_block_frames",
    "midas_extend_blocks", "midas_scattering_quantities", "midas_near_field", "midas_partition_sph",
    "midas_partition_csh", "midas_scattering_quantities_grid", "midas_quad_adaptive", "midas_quad_trap2d",
    "midas_quad_gauss_legendre", "midas_release_workspace", "midas_set_cbfm_budgets", "midas_dev_mom_solve_noalloc",
    "midas_tune_set", "midas_tune_get",
)

RX_SD, RX_RF, RX_LB = 0, 1, 2
INT_SD, INT_LB, INT_UN, INT_GL, INT_TR, INT_SM = range(6)   # Compute_Scattering_Quantities_1 rules


class MidasError(RuntimeError):
    pass


class NotConverged(MidasError):
    """A mixed-precision solve with no_fallback missed its backward-error
    target; .result holds what the call returned (the best iterate)."""

    def __init__(self, msg, result):
        super().__init__(msg)
        self.result = result


class cells_t(C.Structure):
    _fields_ = [("n", C.c_int), ("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p), ("sc", C.c_void_p),
                ("eps", C.c_void_p), ("kappa", C.c_void_p), ("che", C.c_void_p), ("znnpp", C.c_void_p),
                ("id", C.c_void_p)]


class solve_opts_t(C.Structure):
    _fields_ = [("max_iter", C.c_int), ("tol", C.c_double), ("nb", C.c_int), ("no_fallback", C.c_int)]


class solve_info_t(C.Structure):
    _fields_ = [("info", C.c_int), ("iterations", C.c_int), ("berr", C.c_double), ("anorm", C.c_double),
                ("t_factor_ms", C.c_double), ("t_solve_ms", C.c_double), ("t_refine_ms", C.c_double),
                ("status", C.c_int), ("pivot_search", C.c_int), ("max_l21", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class blocks_t(C.Structure):
    _fields_ = [("nblocks", C.c_int), ("block_start", C.c_void_p), ("ext_start", C.c_void_p),
                ("ext_cells", C.c_void_p)]


_LIB = None


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise MidasError(f"{LIBPATH} not built: run `python -m midas_amd.build`")
        lib = C.CDLL(LIBPATH)
        lib.midas_last_error.restype = C.c_char_p
        lib.midas_stream.restype = C.c_void_p
        _LIB = lib
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


TUNE_NAMES = ("lu_nb_split", "lu_two_panel_min_n", "lu_nb_mixed", "lu_nb_fp64", "panel_leaf", "trail_group_kib",
              "refine_graph", "zgemm_tile", "fp64_gemm_library", "cbfm_block_mode", "cgemm_tile", "cbfm_graph",
              "lu_pivot_search", "mom_matrix_free", "u12_split",
              "l21_split", "lu_lookahead")


def tune_get(name):
    """A schedule parameter of the library (process-wide)."""
    v = C.c_int(0)
    _check(load().midas_tune_get(name.encode(), C.byref(v)), "tune_get")
    return v.value


def tune_set(name, value):
    """Set a schedule parameter (process-wide); returns the previous value."""
    old = tune_get(name)
    _check(load().midas_tune_set(name.encode(), C.c_int(int(value))), "tune_set")
    return old


def _check(r, what):
    if r < 0:
        raise MidasError(f"{what} failed ({r}): {load().midas_last_error().decode()}")
    return r


class CellTable:
    """SoA cell table (reference `type Cell` fields the hot path reads)."""

    def __init__(self, x, y, z, sc, eps, kappa=None, che=None, znnpp=None, ids=None):
        f = lambda a: np.ascontiguousarray(a, np.float64)
        self.x, self.y, self.z, self.sc = f(x), f(y), f(z), f(sc)
        self.eps = np.ascontiguousarray(eps, np.complex128)
        self.kappa = None if kappa is None else f(kappa)
        self.che = None if che is None else np.ascontiguousarray(che, np.complex128)
        self.znnpp = None if znnpp is None else np.ascontiguousarray(znnpp, np.complex128)
        self.ids = None if ids is None else np.ascontiguousarray(ids, np.int32)
        self.n = len(self.x)

    @classmethod
    def from_scatterer(cls, s):
        return cls(s.x, s.y, s.z, s.sc, s.eps)

    def subset(self, idx):
        idx = np.asarray(idx)
        g = lambda a: None if a is None else a[idx]
        ids = (idx + 1) if self.ids is None else self.ids[idx]
        return CellTable(self.x[idx], self.y[idx], self.z[idx], self.sc[idx], self.eps[idx], g(self.kappa),
                         g(self.che), g(self.znnpp), ids)

    def struct(self):
        s = cells_t(self.n, _p(self.x), _p(self.y), _p(self.z), _p(self.sc), _p(self.eps), _p(self.kappa),
                    _p(self.che), _p(self.znnpp), _p(self.ids))
        s._keep = self
        return s


def _cells(c):
    return c if isinstance(c, CellTable) else CellTable.from_scatterer(c)


def _block_starts(blocks, n):
    starts = np.zeros(len(blocks) + 1, np.int32)
    for b, blk in enumerate(blocks):
        blk = np.asarray(blk)
        if len(blk) == 0 or np.any(np.diff(blk) != 1) or blk[0] != starts[b]:
            raise MidasError("blocks must be non-empty contiguous cell ranges tiling the cells in order")
        starts[b + 1] = blk[-1] + 1
    if starts[-1] != n:
        raise MidasError("blocks must cover every cell")
    return starts


def block_frames(cells, blocks):
    """SetBlockCtrs for every block: (nblocks x 13) box min, max, extent,
    sphere centre and radius.  Host routine, no GPU context needed."""
    cells = _cells(cells)
    starts = _block_starts(blocks, cells.n)
    fr = np.empty((len(blocks), 13), np.float64)
    cs = cells.struct()
    _check(load().midas_block_frames(C.byref(cs), C.c_int(len(blocks)), _p(starts), _p(fr)), "block_frames")
    return fr


def extend_blocks(cells, blocks, sc, nc_extended=1):
    """Extend_blocks: (extension cells per block, adjacent blocks per block),
    0-based.  Host routine, no GPU context needed."""
    cells = _cells(cells)
    lib = load()
    nb = len(blocks)
    starts = _block_starts(blocks, cells.n)
    es = np.zeros(nb + 1, np.int32)
    cs = cells.struct()
    _check(lib.midas_extend_blocks(C.byref(cs), C.c_int(nb), _p(starts), C.c_double(sc), C.c_int(nc_extended),
                                   _p(es), None, C.c_long(0), None, None, C.c_long(0)), "extend_blocks")
    ec = np.zeros(max(int(es[-1]), 1), np.int32)
    ad = np.zeros(max(nb * nb, 1), np.int32)
    ast = np.zeros(nb + 1, np.int32)
    _check(lib.midas_extend_blocks(C.byref(cs), C.c_int(nb), _p(starts), C.c_double(sc), C.c_int(nc_extended),
                                   _p(es), _p(ec), C.c_long(len(ec)), _p(ast), _p(ad), C.c_long(len(ad))),
           "extend_blocks")
    ext = [ec[es[b]:es[b + 1]].astype(np.int64) for b in range(nb)]
    adj = [ad[ast[b]:ast[b + 1]].astype(np.int64) for b in range(nb)]
    return ext, adj


def _partition(fn, cells, *args):
    cells = _cells(cells)
    order = np.zeros(cells.n, np.int32)
    sizes = np.zeros(cells.n, np.int32)
    nb = C.c_int(0)
    hb = C.c_double(0.0)
    cs = cells.struct()
    _check(fn(C.byref(cs), *args, _p(order), _p(sizes), C.c_int(cells.n), C.byref(nb), C.byref(hb)), "partition")
    return order.astype(np.int64), sizes[:nb.value].astype(np.int64), hb.value


def partition_sph(cells, navg_cells, best_combin=0):
    """Division_blocks, div_type SPH (octree): (order, block sizes, hBlock).
    order[i] = input index of the cell at position i of the reordered cells.
    Host routine, no GPU context needed."""
    return _partition(load().midas_partition_sph, cells, C.c_int(navg_cells), C.c_int(best_combin))


def partition_csh(cells, sc, navg_cells, type_s=2):
    """Division_blocks, div_type CSH (slabs along the longest side, then
    balanced halvings): (order, block sizes, hBlock).  Host routine."""
    return _partition(load().midas_partition_csh, cells, C.c_double(sc), C.c_int(navg_cells), C.c_int(type_s))


def scattering_quantities_grid(S, nrx, rx_dirs, rx_type, tr_dirs, tr_type, k0, dm, cext, cabs, rx_w=None,
                               cext_intfields=False, quirks=Q_ALL):
    """Compute_Scattering_Quantities_1 on the host: S (NRx_tot x 4 ntr),
    rx_dirs = (theta, phi) of all NRx_tot receivers (the first nrx are the
    scattering grid or list), tr_dirs of the ntr transmitters (a theta x phi
    grid, theta fastest).  Returns (Q (ntr x 5): Q_ext, Q_abs, Q_sca, Q_bks,
    g; Q_av (5))."""
    S = np.asfortranarray(S, np.complex128)
    rth, rph = (np.ascontiguousarray(a, np.float64) for a in rx_dirs)
    tth, tph = (np.ascontiguousarray(a, np.float64) for a in tr_dirs)
    ntr = len(tth)
    ce = np.ascontiguousarray(cext, np.complex128)
    ca = np.ascontiguousarray(cabs, np.complex128)
    w = None if rx_w is None else np.ascontiguousarray(rx_w, np.float64)
    Q = np.zeros((ntr, 5))
    Qav = np.zeros(5)
    _check(load().midas_scattering_quantities_grid(_p(S), C.c_long(S.shape[0]), C.c_int(len(rth)), _p(rth), _p(rph),
                                                   C.c_int(nrx), C.c_int(rx_type), _p(w), C.c_int(ntr), _p(tth),
                                                   _p(tph), C.c_int(tr_type), C.c_double(k0), C.c_double(dm),
                                                   C.c_int(quirks), _p(ce), _p(ca), C.c_int(int(cext_intfields)),
                                                   _p(Q), _p(Qav)), "scattering_quantities_grid")
    return Q, Qav


def quad_adaptive(f, nth, nph, ph, jj, a0, b0):
    """getAdaptiveQuadPts2 of the library (column jj, 1-based, of f)."""
    f = np.asfortranarray(f, np.float64)
    out = C.c_double(0)
    _check(load().midas_quad_adaptive(C.c_int(nth), C.c_int(nph), C.c_int(f.shape[0]), C.c_int(f.shape[1]), _p(f),
                                      C.c_int(1 if ph else 0), C.c_int(jj), C.c_double(a0), C.c_double(b0),
                                      C.byref(out)), "quad_adaptive")
    return out.value


def quad_trap2d(f, a, b, c, d):
    f = np.asfortranarray(f, np.float64)
    out = C.c_double(0)
    _check(load().midas_quad_trap2d(C.c_int(f.shape[0]), C.c_int(f.shape[1]), _p(f), C.c_double(a), C.c_double(b),
                                    C.c_double(c), C.c_double(d), C.byref(out)), "quad_trap2d")
    return out.value


def quad_gauss_legendre(n, a, b):
    x, w = np.zeros(n), np.zeros(n)
    _check(load().midas_quad_gauss_legendre(C.c_int(n), C.c_double(a), C.c_double(b), _p(x), _p(w)),
           "quad_gauss_legendre")
    return x, w


class Midas:
    """One context (one GPU, one stream)."""

    def __init__(self, device=0):
        self.lib = load()
        self.ctx = C.c_void_p()
        _check(self.lib.midas_init(C.c_int(device), C.byref(self.ctx)), "midas_init")

    def close(self):
        if self.ctx:
            self.lib.midas_finalize(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self):
        return self.lib.midas_stream(self.ctx)

    def release_workspace(self):
        """Free the solver workspace the context keeps between calls."""
        _check(self.lib.midas_release_workspace(self.ctx), "release_workspace")

    def set_cbfm_budgets(self, lu=0.0, strip=0.0, dense=0.0, aca=0.0):
        """Batching budgets of the CBFM path in bytes (0 = default)."""
        _check(self.lib.midas_set_cbfm_budgets(self.ctx, C.c_double(lu), C.c_double(strip), C.c_double(dense),
                                               C.c_double(aca)), "set_cbfm_budgets")

    # -- a1 ---------------------------------------------------------------
    def set_cells_params(self, sc, eps, k0, quirks=Q_ALL):
        sc = np.ascontiguousarray(sc, np.float64)
        eps = np.ascontiguousarray(eps, np.complex128)
        n = len(sc)
        a, kap = np.empty(n), np.empty(n)
        che, zn = np.empty(n, np.complex128), np.empty(n, np.complex128)
        _check(self.lib.midas_cell_params(self.ctx, C.c_int(n), _p(sc), _p(eps), C.c_double(k0), C.c_int(quirks),
                                          _p(a), _p(che), _p(zn), _p(kap)), "cell_params")
        return a, che, zn, kap

    # -- a2 / a7 ------------------------------------------------------------
    def mom_fill(self, cells, k0, quirks=Q_ALL):
        cells = _cells(cells)
        n3 = 3 * cells.n
        Z = np.empty((n3, n3), np.complex128, order="F")
        cs = cells.struct()
        _check(self.lib.midas_mom_fill(self.ctx, C.byref(cs), C.c_double(k0), C.c_int(quirks), _p(Z), C.c_long(n3)),
               "mom_fill")
        return Z

    def green_s_tr_partial(self, c1, c2, k0, quirks=Q_ALL):
        c1, c2 = _cells(c1), _cells(c2)
        Z = np.empty((3 * c1.n, 3 * c2.n), np.complex128, order="F")
        s1, s2 = c1.struct(), c2.struct()
        _check(self.lib.midas_block_fill(self.ctx, C.byref(s1), C.byref(s2), C.c_double(k0), C.c_int(quirks), _p(Z),
                                         C.c_long(3 * c1.n)), "block_fill")
        return Z

    # -- a3 / a10 -----------------------------------------------------------
    def plane_wave(self, cells, k0, th, ph, ref32=False, quirks=Q_ALL):
        cells = _cells(cells)
        th, ph = np.ascontiguousarray(th, np.float64), np.ascontiguousarray(ph, np.float64)
        V = np.empty((3 * cells.n, 2 * len(th)), np.complex128, order="F")
        cs = cells.struct()
        _check(self.lib.midas_plane_wave(self.ctx, C.byref(cs), C.c_double(k0), C.c_int(quirks), C.c_int(len(th)),
                                         _p(th), _p(ph), C.c_int(int(ref32)), _p(V), C.c_long(3 * cells.n)),
               "plane_wave")
        return V

    def incident_field(self, cells, k0, th, ph, quirks=Q_ALL):
        return self.plane_wave(cells, k0, th, ph, ref32=True, quirks=quirks)

    # -- a10: spherical-wave CBF excitation (Incident_Field_Spherical) ------
    def spherical_wave(self, cells, k0, th, ph, ref32=True, quirks=Q_ALL):
        cells = _cells(cells)
        th, ph = np.ascontiguousarray(th, np.float64), np.ascontiguousarray(ph, np.float64)
        V = np.empty((3 * cells.n, 2 * len(th)), np.complex128, order="F")
        cs = cells.struct()
        _check(self.lib.midas_spherical_wave(self.ctx, C.byref(cs), C.c_double(k0), C.c_int(quirks),
                                             C.c_int(len(th)), _p(th), _p(ph), C.c_int(int(ref32)), _p(V),
                                             C.c_long(3 * cells.n)), "spherical_wave")
        return V

    # -- (f) near field (FFA = 0) --------------------------------------------
    def near_field(self, cells, k0, E, rth, rph, rso, quirks=Q_ALL):
        cells = _cells(cells)
        E = np.asfortranarray(E, np.complex128)
        ntr = E.shape[1] // 2
        rth, rph = np.ascontiguousarray(rth, np.float64), np.ascontiguousarray(rph, np.float64)
        Es = np.empty((len(rth), 4 * ntr), np.complex128, order="F")
        cs = cells.struct()
        _check(self.lib.midas_near_field(self.ctx, C.byref(cs), C.c_double(k0), C.c_int(quirks), _p(E),
                                         C.c_long(E.shape[0]), C.c_int(ntr), C.c_int(len(rth)), _p(rth), _p(rph),
                                         C.c_double(rso), _p(Es), C.c_long(len(rth))), "near_field")
        return Es

    # -- (f) orientation-averaged quantities ---------------------------------
    def scattering_quantities(self, S, nrx, ntr, rx_dirs, tr_dirs, k0, dm, cext, cabs, rx_type=RX_SD, rx_w=None,
                              tr_type=RX_SD, tr_w=None, cext_intfields=False, quirks=Q_ALL):
        """Q (ntr x 5: Q_ext, Q_abs, Q_sca, Q_bks, g) and their averages."""
        S = np.asfortranarray(S, np.complex128)
        f = lambda a: None if a is None else np.ascontiguousarray(a, np.float64)
        rth, rph = (f(a) for a in rx_dirs)
        tth, tph = (f(a) for a in tr_dirs)
        rw, tw = f(rx_w), f(tr_w)
        ce, ca = np.ascontiguousarray(cext, np.complex128), np.ascontiguousarray(cabs, np.complex128)
        Q = np.zeros((ntr, 5), np.float64)
        av = np.zeros(5, np.float64)
        _check(self.lib.midas_scattering_quantities(self.ctx, _p(S), C.c_long(S.shape[0]), C.c_int(S.shape[0]),
                                                    C.c_int(nrx), C.c_int(ntr), _p(rth), _p(rph), C.c_int(rx_type),
                                                    _p(rw), _p(tth), _p(tph), C.c_int(tr_type), _p(tw),
                                                    C.c_double(k0), C.c_double(dm), C.c_int(quirks), _p(ce), _p(ca),
                                                    C.c_int(int(cext_intfields)), _p(Q), _p(av)),
               "scattering_quantities")
        return Q, av

    # -- a4 -----------------------------------------------------------------
    def solve(self, Z, B, mode=SOLVE_FP64, max_iter=0, tol=0.0, nb=0, no_fallback=False):
        """X = Z^-1 B and the solve report.  With no_fallback a refinement that
        misses tol raises NotConverged (its .result = (X, report))."""
        Z = np.asfortranarray(Z, np.complex128)
        X = np.array(B, np.complex128, order="F", copy=True)
        if X.ndim == 1:
            X = X.reshape(-1, 1, order="F")
        n, nrhs = X.shape
        opts = solve_opts_t(max_iter, tol, nb, int(no_fallback))
        info = solve_info_t()
        r = self.lib.midas_solve(self.ctx, C.c_int(n), C.c_int(nrhs), _p(Z), C.c_long(Z.shape[0]), _p(X),
                                 C.c_long(n), C.c_int(mode), C.byref(opts), C.byref(info))
        d = info.as_dict()
        d["info"] = r
        if r == ERR_NOT_CONVERGED:
            raise NotConverged(self.lib.midas_last_error().decode(), (X, d))
        _check(r, "solve")
        return X, d

    # -- a5 / a6 ------------------------------------------------------------
    def scattering_matrix(self, cells, k0, E, rth, rph, beta=(0.0,), quirks=Q_ALL):
        cells = _cells(cells)
        E = np.asfortranarray(E, np.complex128)
        ntr = E.shape[1] // 2
        rth, rph = np.ascontiguousarray(rth, np.float64), np.ascontiguousarray(rph, np.float64)
        beta = np.ascontiguousarray(beta, np.float64)
        S = np.empty((len(rth), 4 * ntr * len(beta)), np.complex128, order="F")
        cs = cells.struct()
        _check(self.lib.midas_scattering_matrix(self.ctx, C.byref(cs), C.c_double(k0), C.c_int(quirks), _p(E),
                                                C.c_long(E.shape[0]), C.c_int(ntr), C.c_int(len(rth)), _p(rth),
                                                _p(rph), C.c_int(len(beta)), _p(beta), _p(S), C.c_long(len(rth))),
               "scattering_matrix")
        return S

    def cext_cabs(self, cells, k0, E, th, ph, quirks=Q_ALL):
        cells = _cells(cells)
        E = np.asfortranarray(E, np.complex128)
        th, ph = np.ascontiguousarray(th, np.float64), np.ascontiguousarray(ph, np.float64)
        ce = np.empty(len(th), np.complex128)
        ca = np.empty(len(th), np.complex128)
        cs = cells.struct()
        _check(self.lib.midas_cext_cabs(self.ctx, C.byref(cs), C.c_double(k0), C.c_int(quirks), _p(E),
                                        C.c_long(E.shape[0]), C.c_int(len(th)), _p(th), _p(ph), _p(ce), _p(ca)),
               "cext_cabs")
        return ce, ca

    # -- Compute_EFields_MoM -------------------------------------------------
    def compute_efields_mom(self, cells, k0, transmitters, receivers, mode=SOLVE_FP64, quirks=Q_ALL, want_E=True,
                            max_iter=0, tol=0.0, nb=0, no_fallback=False):
        cells = _cells(cells)
        tth, tph = (np.ascontiguousarray(a, np.float64) for a in transmitters)
        rth, rph = (np.ascontiguousarray(a, np.float64) for a in receivers)
        ntr, nrx = len(tth), len(rth)
        S = np.empty((nrx, 4 * ntr), np.complex128, order="F")
        ce = np.empty(ntr, np.complex128)
        ca = np.empty(ntr, np.complex128)
        E = np.empty((3 * cells.n, 2 * ntr), np.complex128, order="F") if want_E else None
        opts = solve_opts_t(max_iter, tol, nb, int(no_fallback))
        info = solve_info_t()
        cs = cells.struct()
        r = self.lib.midas_compute_efields_mom(self.ctx, C.byref(cs), C.c_double(k0), C.c_int(quirks), C.c_int(ntr),
                                               _p(tth), _p(tph), C.c_int(nrx), _p(rth), _p(rph), C.c_int(mode),
                                               C.byref(opts), _p(S), _p(ce), _p(ca), _p(E), C.byref(info))
        d = info.as_dict()
        d["info"] = r
        if r == ERR_NOT_CONVERGED:
            raise NotConverged(self.lib.midas_last_error().decode(), (S, ce, ca, E, d))
        _check(r, "compute_efields_mom")
        return S, ce, ca, E, d

    # -- a14 / a16 ----------------------------------------------------------
    def calcul_matrixzij_aca(self, cells, k0, r0, nr, c0, nc, kmax=50, eps=1e-4, quirks=Q_ALL):
        cells = _cells(cells)
        U = np.zeros((3 * nr, kmax), np.complex128, order="F")
        V = np.zeros((kmax, 3 * nc), np.complex128, order="F")
        k = C.c_int(0)
        cs = cells.struct()
        _check(self.lib.midas_aca_block(self.ctx, C.byref(cs), C.c_double(k0), C.c_int(quirks), C.c_int(r0),
                                        C.c_int(nr), C.c_int(c0), C.c_int(nc), C.c_int(kmax), C.c_double(eps), _p(U),
                                        _p(V), C.byref(k)), "aca_block")
        kk = min(k.value, kmax)
        return k.value, U[:, :kk], V[:kk, :]

    def calcul_genmatrix_aca(self, A, kmax=50, eps=1e-4):
        A = np.asfortranarray(A, np.complex128)
        m, n = A.shape
        U = np.zeros((m, kmax), np.complex128, order="F")
        V = np.zeros((kmax, n), np.complex128, order="F")
        k = C.c_int(0)
        _check(self.lib.midas_aca_matrix(self.ctx, C.c_int(m), C.c_int(n), _p(A), C.c_long(m), C.c_int(kmax),
                                         C.c_double(eps), _p(U), _p(V), C.byref(k)), "aca_matrix")
        kk = min(k.value, kmax)
        return k.value, U[:, :kk], V[:kk, :]

    # -- Compute_EFields_CBFME(_ACA) -----------------------------------------
    def compute_efields_cbfme(self, cells, k0, blocks, ext, cbf_dirs, transmitters, use_aca=False, kmax=50,
                              eps_aca=1e-4, threshold=0.0, mode=SOLVE_FP64, quirks=Q_ALL,
                              return_cbfs=False, sr=False, res_sr=0.01, iws=1, e_dev=None):
        """e_dev: optional device address of a (3n x 2T) complex128 column-major
        buffer (ld 3n); the fields are written there and stay on the GPU
        (midas_dev_compute_efields_cbfm), and E is returned as None."""
        cells = _cells(cells)
        nb = len(blocks)
        starts = np.zeros(nb + 1, np.int32)
        for b, blk in enumerate(blocks):
            blk = np.asarray(blk)
            if len(blk) == 0 or np.any(np.diff(blk) != 1):
                raise MidasError("blocks must be non-empty contiguous cell ranges in block order")
            if b > 0 and blk[0] != starts[b]:
                raise MidasError("blocks must tile the cells in order")
            starts[b + 1] = blk[-1] + 1
        if starts[-1] != cells.n:
            raise MidasError("blocks must cover every cell")
        es = np.zeros(nb + 1, np.int32)
        for b, e in enumerate(ext):
            es[b + 1] = es[b] + len(e)
        ec = np.ascontiguousarray(np.concatenate([np.asarray(e, np.int32) for e in ext]) if es[-1] else
                                  np.zeros(1, np.int32), np.int32)
        bt = blocks_t(nb, _p(starts), _p(es), _p(ec))
        cth, cph = (np.ascontiguousarray(a, np.float64) for a in cbf_dirs)
        tth, tph = (np.ascontiguousarray(a, np.float64) for a in transmitters)
        E = None if e_dev is not None else np.empty((3 * cells.n, 2 * len(tth)), np.complex128, order="F")
        K = np.zeros(nb, np.int32)
        kcap = 2 * len(cth)
        Cj = np.zeros((3 * cells.n, kcap), np.complex128, order="F") if return_cbfs else None
        fsr = np.zeros(nb, np.float64)
        nnz = np.zeros(nb, np.int64)
        info = solve_info_t()
        cs = cells.struct()
        entry = self.lib.midas_compute_efields_cbfm if e_dev is None else self.lib.midas_dev_compute_efields_cbfm
        r = _check(entry(self.ctx, C.byref(cs), C.byref(bt), C.c_double(k0),
                                                       C.c_int(quirks), C.c_int(len(cth)), _p(cth), _p(cph),
                                                       C.c_int(iws), C.c_int(len(tth)), _p(tth), _p(tph),
                                                       C.c_int(int(use_aca)),
                                                       C.c_int(kmax), C.c_double(eps_aca), C.c_double(threshold),
                                                       C.c_int(mode), _p(E) if E is not None else C.c_void_p(e_dev),
                                                       C.c_long(3 * cells.n), _p(K),
                                                       None if Cj is None else _p(Cj), C.c_long(3 * cells.n),
                                                       C.c_int(kcap), C.c_int(int(sr)), C.c_double(res_sr),
                                                       _p(fsr), _p(nnz), C.byref(info)),
                   "compute_efields_cbfm")
        d = info.as_dict()
        d["info"] = r
        if sr:
            d["fsr"], d["nnz"] = fsr, nnz
        if return_cbfs:
            return E, K, d, Cj[:, :max(int(K.max()), 1)]
        return E, K, d
