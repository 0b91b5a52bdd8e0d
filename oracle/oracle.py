"""Python face of the oracle (TEST INFRASTRUCTURE ONLY).

Loads oracle/liboracle.so (the C restatement in oracle.c) and, when it has been
built from the reference tree, oracle/_ref/libmidas_ref.so (the reference's
own Fortran routines).  Dense solves use LAPACK through numpy (zgesv), the
CBFM pipeline is restated with numpy (block solves, SVD, sigma-ratio
truncation, Zc = C^T Z C, Vc, E = C alpha).

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may
import this module; the product never does.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = C.CDLL(path)
        d, i, v = C.c_double, C.c_int, C.c_void_p
        _LIB.or_pi.restype = d
        _LIB.or_pi.argtypes = [i]
        _LIB.or_aca.restype = i
        _LIB.or_aca_gen.restype = i
        _LIB.or_block_fn.restype = d
        _LIB.or_sr_block.restype = C.c_long
    return _LIB


def ref():
    """The reference's own routines (oracle/_ref), or None when not built."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libmidas_ref.so")
        if not os.path.exists(path):
            return None
        # lazy binding: one unused reference routine has an unresolved helper
        _REF = C.CDLL(path, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
    return _REF


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


class Cells:
    """SoA cell table (the fields of the reference's `type Cell` the hot path reads)."""

    def __init__(self, x, y, z, sc, eps, k0, quirk=1):
        self.x, self.y, self.z = _f64(x), _f64(y), _f64(z)
        self.sc = _f64(sc)
        self.eps = _c128(eps)
        self.n = len(self.x)
        self.k0 = float(k0)
        self.quirk = quirk
        self.id = np.arange(1, self.n + 1, dtype=np.int32)
        self.a = np.empty(self.n)
        self.che = np.empty(self.n, np.complex128)
        self.znnpp = np.empty(self.n, np.complex128)
        self.kappa = np.empty(self.n)
        lib().or_cell_params(C.c_int(self.n), _p(self.sc), _p(self.eps), C.c_double(self.k0), C.c_int(quirk),
                             _p(self.a), _p(self.che), _p(self.znnpp), _p(self.kappa))

    def subset(self, idx):
        """Cells idx (0-based), keeping their global identifiers."""
        c = object.__new__(Cells)
        for f in ("x", "y", "z", "sc", "eps", "a", "che", "znnpp", "kappa", "id"):
            setattr(c, f, np.ascontiguousarray(getattr(self, f)[idx]))
        c.n, c.k0, c.quirk = len(c.x), self.k0, self.quirk
        return c


def pi(quirk=1):
    return lib().or_pi(quirk)


def mom_fill(cells: Cells):
    n3 = 3 * cells.n
    Z = np.empty((n3, n3), np.complex128, order="F")
    lib().or_mom_fill(C.c_int(cells.n), _p(cells.x), _p(cells.y), _p(cells.z), _p(cells.kappa), _p(cells.che),
                      _p(cells.znnpp), C.c_double(cells.k0), C.c_int(cells.quirk), _p(Z))
    return Z


def block_fill(c1: Cells, c2: Cells, fp32=True):
    Z = np.empty((3 * c1.n, 3 * c2.n), np.complex128, order="F")
    fn = lib().or_block_fill_f32 if fp32 else lib().or_block_fill_f64
    fn(C.c_int(c1.n), _p(c1.id), _p(c1.x), _p(c1.y), _p(c1.z), C.c_int(c2.n), _p(c2.id), _p(c2.x), _p(c2.y),
       _p(c2.z), _p(c2.kappa), _p(c2.che), _p(c2.znnpp), C.c_double(c1.k0), C.c_int(c1.quirk), _p(Z))
    return Z


def block_row(cells: Cells, grow, c0, ncols):
    """computeBlockRow: row `grow` (0-based global) over cells c0 .. c0+ncols/3-1, FP32 intermediates."""
    row = np.empty(ncols, np.complex128)
    lib().or_block_row_f32(C.c_int(grow), C.c_int(c0), C.c_int(ncols), _p(cells.x), _p(cells.y), _p(cells.z),
                           _p(cells.kappa), _p(cells.che), C.c_double(cells.k0), C.c_int(cells.quirk), _p(row))
    return row


def block_col(cells: Cells, gcol, r0, nrows):
    """computeBlockCol: column `gcol` (0-based global) over cells r0 .. r0+nrows/3-1."""
    col = np.empty(nrows, np.complex128)
    lib().or_block_col_f32(C.c_int(gcol), C.c_int(r0), C.c_int(nrows), _p(cells.x), _p(cells.y), _p(cells.z),
                           _p(cells.kappa), _p(cells.che), C.c_double(cells.k0), C.c_int(cells.quirk), _p(col))
    return col


def mom_rhs(cells: Cells, th_deg, ph_deg):
    th, ph = _f64(th_deg), _f64(ph_deg)
    V = np.empty((3 * cells.n, 2 * len(th)), np.complex128, order="F")
    lib().or_mom_rhs(C.c_int(cells.n), _p(cells.x), _p(cells.y), _p(cells.z), C.c_int(len(th)), _p(th), _p(ph),
                     C.c_double(cells.k0), C.c_int(cells.quirk), _p(V))
    return V


def incident_field_f32(cells: Cells, th_deg, ph_deg, t0=0, t1=None):
    th, ph = _f64(th_deg), _f64(ph_deg)
    t1 = len(th) if t1 is None else t1
    E = np.empty((3 * cells.n, 2 * (t1 - t0)), np.complex128, order="F")
    lib().or_incident_field_f32(C.c_int(cells.n), _p(cells.x), _p(cells.y), _p(cells.z), C.c_int(t0), C.c_int(t1),
                                _p(th), _p(ph), C.c_double(cells.k0), C.c_int(cells.quirk), _p(E))
    return E


def incident_field_sph(cells: Cells, th_deg, ph_deg):
    """Incident_Field_Spherical (Incident_Field.f90:65-166) over the cells given
    (an extended block, block cells first).  Block centre = sequential sums / n
    (:88-90), R_block = largest distance to it (:91), sources at
    Rsource = 10 max(R_block, 2 Pi / k0) (:94) in the directions (:107-109);
    phase = e^{i k0 R} / R and the components are complex(4) (:76, :126-148),
    the products formed in double (mixed-kind promotion) before the rounding.
    Pi is the reference's constant (single-precision acos(-1.) under quirk)."""
    th, ph = _f64(th_deg), _f64(ph_deg)
    n = cells.n
    x, y, z = _f64(cells.x), _f64(cells.y), _f64(cells.z)
    cx = float(np.cumsum(x)[-1]) / n     # cumsum: strictly sequential
    cy = float(np.cumsum(y)[-1]) / n
    cz = float(np.cumsum(z)[-1]) / n
    rbl = np.sqrt(np.max((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2))
    PI = pi(cells.quirk)
    rs = 10.0 * max(rbl, 2.0 * PI / cells.k0)
    f32 = lambda a: a.astype(np.float32).astype(np.float64)
    E = np.zeros((3 * n, 2 * len(th)), np.complex128, order="F")
    for t in range(len(th)):
        tr, pr_ = th[t] * PI / 180.0, ph[t] * PI / 180.0
        xs = cx + rs * np.cos(tr)
        ys = cy + rs * np.sin(tr) * np.cos(pr_)
        zs = cz + rs * np.sin(tr) * np.sin(pr_)
        dx, dy, dz = x - xs, y - ys, z - zs
        R = np.sqrt(dx * dx + dy * dy + dz * dz)
        ux, uy, uz = dx / R, dy / R, dz / R
        kr = cells.k0 * R
        ph_r, ph_i = f32(np.cos(kr) / R), f32(np.sin(kr) / R)
        theta, phi = np.arccos(ux), np.arctan2(uz, uy)
        st, ct, sp, cp = np.sin(theta), np.cos(theta), np.sin(phi), np.cos(phi)
        cv = lambda re, im: f32(re) + 1j * f32(im)
        E[0::3, t] = cv(-(ph_r * st), -(ph_i * st))
        E[1::3, t] = cv(ph_r * ct * cp, ph_i * ct * cp)
        E[2::3, t] = cv(ph_r * ct * sp, ph_i * ct * sp)
        T = t + len(th)
        E[1::3, T] = cv(-(ph_r * sp), -(ph_i * sp))
        E[2::3, T] = cv(ph_r * cp, ph_i * cp)
    return E


def ff_coeff(cells: Cells, th, ph):
    ff = np.empty(cells.n, np.complex128)
    lib().or_ff_coeff(C.c_int(cells.n), _p(cells.x), _p(cells.y), _p(cells.z), _p(cells.kappa), _p(cells.che),
                      C.c_double(th), C.c_double(ph), C.c_double(cells.k0), C.c_int(cells.quirk), _p(ff))
    return ff


def green_s_dt(cells: Cells, rso, th, ph):
    """Green_s_dt (Green_functions.f90:1335-1397), FP64: the 3 x 3n dyad
    +F kappa chi from every cell to a receiver at rso (cos th, sin th cos ph,
    sin th sin ph), Gr = e^{i k r} / (4 Pi r^2), Tau = i k - 1/r."""
    PI = pi(cells.quirk)
    k = cells.k0
    t, p = th * PI / 180.0, ph * PI / 180.0
    xc, yc, zc = rso * np.cos(t), rso * np.sin(t) * np.cos(p), rso * np.sin(t) * np.sin(p)
    rx, ry, rz = xc - cells.x, yc - cells.y, zc - cells.z
    r = np.sqrt(rx ** 2 + ry ** 2 + rz ** 2)
    gr = np.exp(1j * k * r) / (4 * PI * r ** 2)
    tau = 1j * k - 1 / r
    f = cells.kappa * cells.che
    diag = lambda ra: gr * (tau + k ** 2 * (r - ra ** 2 / r) - 3.0 * ra ** 2 / r ** 2 * tau)
    off = lambda ra, rb: gr / r * (ra * rb) * (-k ** 2 - 3.0 * tau / r)
    F = [[diag(rx), off(rx, ry), off(rx, rz)], [off(rx, ry), diag(ry), off(ry, rz)],
         [off(rx, rz), off(ry, rz), diag(rz)]]
    G = np.zeros((3, 3 * cells.n), np.complex128)
    for a in range(3):
        for q in range(3):
            G[a, q::3] = F[a][q] * f
    return G


def near_field(cells: Cells, E, rso, rth, rph):
    """Compute_Scattered_Fields (Compute_Scattered_Fields.f90:75-136): scattered
    field at the receivers from E (3n x 2T), projected on (theta, phi):
    Es (nrx x 4T: Vv, Vh, Hv, Hh per transmitter)."""
    PI = pi(cells.quirk)
    T = E.shape[1] // 2
    S = np.zeros((len(rth), 4 * T), np.complex128)
    for r in range(len(rth)):
        G = green_s_dt(cells, rso, rth[r], rph[r])
        Ev, Eh = G @ E[:, :T], G @ E[:, T:]
        t, p = rth[r] * PI / 180.0, rph[r] * PI / 180.0
        st, ct, sp, cp = np.sin(t), np.cos(t), np.sin(p), np.cos(p)
        S[r, 0::4] = -Ev[0] * st + Ev[1] * ct * cp + Ev[2] * ct * sp
        S[r, 1::4] = -Ev[1] * sp + Ev[2] * cp
        S[r, 2::4] = -Eh[0] * st + Eh[1] * ct * cp + Eh[2] * ct * sp
        S[r, 3::4] = -Eh[1] * sp + Eh[2] * cp
    return S


def scattering_matrix(cells: Cells, E, rth, rph, beta_deg=(0.0,)):
    E = np.asfortranarray(E, dtype=np.complex128)
    ntr = E.shape[1] // 2
    rth, rph, beta = _f64(rth), _f64(rph), _f64(beta_deg)
    S = np.empty((len(rth), 4 * ntr * len(beta)), np.complex128, order="F")
    lib().or_scattering_matrix(C.c_int(cells.n), _p(cells.x), _p(cells.y), _p(cells.z), _p(cells.kappa),
                               _p(cells.che), _p(E), C.c_int(ntr), C.c_int(len(rth)), _p(rth), _p(rph),
                               C.c_int(len(beta)), _p(beta), C.c_double(cells.k0), C.c_int(cells.quirk), _p(S))
    return S


def cext_cabs(cells: Cells, E, th_deg, ph_deg, q6=1):
    E = np.asfortranarray(E, dtype=np.complex128)
    th, ph = _f64(th_deg), _f64(ph_deg)
    ce = np.empty(len(th), np.complex128)
    ca = np.empty(len(th), np.complex128)
    lib().or_cext_cabs(C.c_int(cells.n), _p(cells.x), _p(cells.y), _p(cells.z), _p(cells.sc), _p(cells.che), _p(E),
                       C.c_int(len(th)), _p(th), _p(ph), C.c_double(cells.k0), C.c_int(cells.quirk), C.c_int(q6),
                       _p(ce), _p(ca))
    return ce, ca


RX_SD, RX_RF, RX_LB = 0, 1, 2


def scattering_quantities(S, nrx, ntr, rx_th, rx_ph, tr_th, tr_ph, k0, dm, cext, cabs, rx_type=RX_SD,
                          rx_w=None, tr_type=RX_SD, tr_w=None, cext_intfields=False, quirk=1):
    """Compute_Scattering_Quantities_2 (Compute_Scattering_Quantities.f90:593-1013),
    receiver rules sd / rf (equal weights 4 pi / NRx, :812-831) and lb
    (Lebedev weights, :790-811).  S: NRx_tot x 4 NTr (Vv, Vh, Hv, Hh per
    transmitter).  Returns Q (ntr x 5: Q_ext, Q_abs, Q_sca, Q_bks, g) and the
    orientation averages (5).  Sums are sequential, as Fortran's SUM."""
    PI = pi(quirk)
    ap = dm / 2.0
    X = k0 * dm / 2.0
    a2 = lambda z: np.hypot(z.real, z.imag) ** 2
    rad = lambda d: d * PI / 180.0
    Q = np.zeros((ntr, 5))
    for t in range(ntr):
        kix = -(np.sin(rad(tr_th[t])) * np.cos(rad(tr_ph[t])))
        kiy = -(np.sin(rad(tr_th[t])) * np.sin(rad(tr_ph[t])))
        kiz = -np.cos(rad(tr_th[t]))
        fw = t if (ntr == nrx and rx_type in (RX_SD, RX_LB)) else nrx + t
        bk = nrx + t if (ntr == nrx and rx_type in (RX_SD, RX_LB, RX_RF)) else nrx + ntr + t
        q_ext = (2 / (X * X)) * (abs(S[fw, 4 * t + 3].imag) + abs(S[fw, 4 * t + 0].imag))
        q_ext_intf = (cext[t] / (PI * ap ** 2)).real
        Svv, Svh, Shv, Shh = (S[:nrx, 4 * t + j] for j in range(4))
        f = ((a2(Shh) + a2(Svv)) + a2(Svh)) + a2(Shv)
        if rx_type == RX_LB:
            f = rx_w * f
        ksx = -(np.sin(rad(rx_th)) * np.cos(rad(rx_ph)))
        ksy = -(np.sin(rad(rx_th)) * np.sin(rad(rx_ph)))
        ksz = -np.cos(rad(rx_th))
        cg = kix * ksx + kiy * ksy + kiz * ksz
        fg = f * cg
        s1, s2 = float(np.cumsum(f)[-1]), float(np.cumsum(fg)[-1])
        if rx_type == RX_LB:
            q_sd = s1
            c_sca = 1.0 / (2 * k0 ** 2) * q_sd
            q_sca = (4 * PI) / (2 * PI * X * X) * q_sd
            g = 1.0 / (2.0 * k0 ** 2.0 * c_sca) * s2
        else:
            q_sd = (4 * PI / nrx) * s1
            c_sca = 1.0 / (2 * k0 ** 2) * q_sd
            q_sca = 1 / (2 * PI * X * X) * q_sd
            g = 1.0 / (2.0 * k0 ** 2.0 * c_sca) * (4 * PI / nrx) * s2
        q_bks = (1.0 / (2.0 * PI * X * X)) * (((a2(S[bk, 4 * t + 0]) + a2(S[bk, 4 * t + 3])) + a2(S[bk, 4 * t + 1]))
                                              + a2(S[bk, 4 * t + 2]))
        q_abs_intf = (cabs[t] / (PI * ap ** 2)).real
        Q[t] = [q_ext_intf if cext_intfields else q_ext, q_abs_intf, q_sca, q_bks, g]
    if ntr == 1:
        av = Q[0].copy()
    elif tr_type in (RX_SD, RX_RF):
        av = np.array([float(np.cumsum(Q[:, j])[-1]) / ntr for j in range(5)])
    else:
        av = np.zeros(5)
        for i in range(ntr):
            av = av + tr_w[i] * Q[i]
    return Q, av


def aca(cells: Cells, r0, nr, c0, nc, kmax=50, eps=1e-4, fp32=True):
    """ACA of the block rows cells[r0:r0+nr] x columns cells[c0:c0+nc] (0-based)."""
    U = np.zeros((3 * nr, kmax), np.complex128, order="F")
    V = np.zeros((kmax, 3 * nc), np.complex128, order="F")
    rows = np.zeros(kmax, np.int32)
    cols = np.zeros(kmax, np.int32)
    k = lib().or_aca(C.c_int(r0), C.c_int(nr), C.c_int(c0), C.c_int(nc), _p(cells.x), _p(cells.y), _p(cells.z),
                     _p(cells.kappa), _p(cells.che), _p(cells.znnpp), C.c_double(cells.k0), C.c_int(cells.quirk),
                     C.c_int(0 if fp32 else 1), C.c_int(kmax), C.c_double(eps), _p(U), _p(V), _p(rows), _p(cols))
    kk = min(k, kmax)
    return k, U[:, :kk], V[:kk, :], rows[:kk], cols[:kk]


def aca_gen(A, kmax=50, eps=1e-4):
    A = np.asfortranarray(A, dtype=np.complex128)
    m, n = A.shape
    U = np.zeros((m, kmax), np.complex128, order="F")
    V = np.zeros((kmax, n), np.complex128, order="F")
    k = lib().or_aca_gen(C.c_int(m), C.c_int(n), _p(A), C.c_int(kmax), C.c_double(eps), _p(U), _p(V))
    kk = min(k, kmax)
    return k, U[:, :kk], V[:kk, :]


def block_fn(cells: Cells):
    """Green_s_tr_partial_FN of the block over all of `cells` (receivers = sources)."""
    return lib().or_block_fn(C.c_int(cells.n), _p(cells.id), _p(cells.x), _p(cells.y), _p(cells.z), _p(cells.kappa),
                             _p(cells.che), _p(cells.znnpp), C.c_double(cells.k0), C.c_int(cells.quirk))


def sr_block(cells: Cells, fsr, dense=True):
    """SR mask of the block (SR_Green_s_tr_partial / _FN): (nnz, Frobenius norm, dense masked Z or None)."""
    n3 = 3 * cells.n
    Z = np.empty((n3, n3), np.complex128, order="F") if dense else None
    fn = C.c_double(0)
    nnz = lib().or_sr_block(C.c_int(cells.n), _p(cells.x), _p(cells.y), _p(cells.z), _p(cells.kappa), _p(cells.che),
                            _p(cells.znnpp), C.c_double(cells.k0), C.c_int(cells.quirk), C.c_double(fsr), C.byref(fn),
                            _p(Z))
    return nnz, fn.value, Z


FSR_TESTS = (1e2, 1e3, 2e3, 3e3, 4e3, 5e3, 6e3, 7e3, 8e3, 9e3, 1e4)


def setfsr(cells: Cells, res_sr=0.01):
    """setfSR (getParameters_CBFM.f90:113-168): the first fSR of the list whose
    mask loses at most res_sr percent of the block's Frobenius norm (else the
    last).  Returns (fsr, nnz)."""
    fn = block_fn(cells)
    res, j, f, nnz = 1e3, 0, FSR_TESTS[0], 0
    while j < len(FSR_TESTS) and res > res_sr:
        f = FSR_TESTS[j]
        nnz, fns, _ = sr_block(cells, f, dense=False)
        res = abs(fn - fns) / fn * 100.0
        j += 1
    return f, nnz


def transmitters_un(niws):
    """getTransmitters_CBFM (getParameters_CBFM.f90:1-110) with distr_iws = 1:
    the grid uniform in cos(theta) and phi closest to niws directions
    (theta from +x, degrees).  The grid-size expression is single precision in
    the reference (integer and default-real operands)."""
    f32 = np.float32
    delta = float(f32(f32(540.0) ** f32(2.0)) + f32(4 * (niws - 1) * (360 * 180)))
    step = (540 + np.sqrt(delta)) / (2 * (niws - 1))
    nth = int(np.rint(180.0 / step)) + 1
    nph = int(np.rint(360.0 / step)) + 1
    pq = pi(1)
    step_ph = 360.0 / (nph - 1)
    c0 = np.cos(0.0 * pq / 180.0)
    step_th = (np.cos(180.0 * pq / 180.0) - c0) / (nth - 1)
    th, ph = [], []
    for i in range(nph):
        for k in range(nth):
            ph.append(i * step_ph)
            th.append(np.arccos(1.0 * k * step_th + c0) * 180.0 / pq)
    return np.array(th), np.array(ph)


# --------------------------------------------------------------------------
# Dense solve and CBFM pipeline (numpy restatement)
# --------------------------------------------------------------------------

def block_frames(x, y, z, blocks):
    """SetBlockCtrs (Division_blocks.f90:1631-1660): per block box min, max,
    extent, the box centre and the largest distance to it (13 values)."""
    out = []
    for b in blocks:
        p = np.stack([np.asarray(x)[b], np.asarray(y)[b], np.asarray(z)[b]], 1)
        lo, hi = p.min(0), p.max(0)
        c = (lo + hi) / 2.0
        r = float(np.sqrt(((p - c) ** 2).sum(1)).max())
        out.append(np.concatenate([lo, hi, hi - lo, c, [r]]))
    return np.array(out)


def extend_blocks(x, y, z, blocks, sc, nc_extended=1):
    """Extend_blocks (Extend_blocks.f90:1-162), restated with numpy: adjacency
    by bounding spheres (:51-79), the zone = box +- (nc_extended sc + sc/2)
    (:83-88), cells of adjacent blocks inside it in block then cell order
    (:97-133), adjacent blocks kept only if they gave cells (:130-134)."""
    nb = len(blocks)
    if nb == 1 or nc_extended == 0:
        return [np.zeros(0, np.int64) for _ in blocks], [np.zeros(0, np.int64) for _ in blocks]
    fr = block_frames(x, y, z, blocks)
    P = np.stack([x, y, z], 1)
    L = nc_extended * sc
    ext, adj = [], []
    for B in range(nb):
        d = np.sqrt(((fr[:, 9:12] - fr[B, 9:12]) ** 2).sum(1))
        near = [Bp for Bp in range(nb) if Bp != B and d[Bp] <= fr[B, 12] + fr[Bp, 12] + L]
        lo = fr[B, 0:3] - L - sc / 2.0
        hi = fr[B, 3:6] + L + sc / 2.0
        e, a = [], []
        for Bp in near:
            idx = np.asarray(blocks[Bp])
            inside = np.all((P[idx] >= lo) & (P[idx] <= hi), axis=1)
            if inside.any():
                e.append(idx[inside])
                a.append(Bp)
        ext.append(np.concatenate(e).astype(np.int64) if e else np.zeros(0, np.int64))
        adj.append(np.array(a, np.int64))
    return ext, adj


def _f_nint(v):
    """Fortran NINT: nearest integer, halves away from zero."""
    return int(np.floor(abs(v) + 0.5)) * (1 if v >= 0 else -1)


def _f_ceiling(v):
    """Fortran CEILING to a default integer; a value outside the integer range
    (an infinite quotient) gives the x86 conversion's -2**31."""
    if not np.isfinite(v) or abs(v) >= 2.0 ** 31:
        return -(2 ** 31)
    return int(np.ceil(v))


def _octree_split(P, cells, fr):
    """HierarchicalOctree_subdivision (Division_blocks.f90:1403-1537): split the
    cells (positions P[cells]) at fr of their centre extent per axis; octant
    number 1 + ix + 2 iy + 4 iz; empty octants dropped; cells stay in input
    order inside an octant.  Returns the list of octant cell lists."""
    q = P[cells]
    lo = q.min(0)
    h = (q.max(0) - lo) * np.asarray(fr)
    octs = [[] for _ in range(8)]
    for c, p in zip(cells, q):
        i = [0 if p[a] - lo[a] <= h[a] else 1 for a in range(3)]
        octs[i[0] + 2 * i[1] + 4 * i[2]].append(c)
    return [o for o in octs if o]


_THIRD, _TWO_THIRDS = float(np.float32(1.0 / 3.0)), float(np.float32(2.0 / 3.0))
_SPH_FRACTIONS = [(0.5, 0.5, 0.5)] + [(a, b, c) for a in (_THIRD, _TWO_THIRDS) for b in (_THIRD, _TWO_THIRDS)
                                      for c in (_THIRD, _TWO_THIRDS)]


def partition_sph(x, y, z, navg, best_combin=0):
    """Division_blocks with div_type SPH (Division_blocks.f90:95-132, SPH_v = 2
    -> best_combin 0, SPH_v = 3 -> 1) = Division_blocks_sph_H (:1252-1401).
    First octree split of all cells, then the block list is scanned: a block of
    N > navg cells with |N - navg| > |nint(N/8.) - navg| is split again in
    place and the scan resumes at its first piece.  With best_combin the split
    fractions are the one of 9 (1/2 everywhere, or 1/3. / 2/3. per axis) whose
    pieces deviate least from their mean.  Returns (order, sizes, hblock):
    order = input indices in the new cell order, sizes per block,
    hblock = 2 max block radius (:160)."""
    P = np.stack([np.asarray(x, float), np.asarray(y, float), np.asarray(z, float)], 1)
    blocks = _octree_split(P, list(range(len(P))), _SPH_FRACTIONS[0])
    ii = 0
    while ii < len(blocks):
        N = len(blocks[ii])
        if N > navg and abs(N - navg) > abs(_f_nint(float(np.float32(N) / np.float32(8.0))) - navg):
            fr = _SPH_FRACTIONS[0]
            if best_combin:
                best = None
                for f in _SPH_FRACTIONS:
                    sub = _octree_split(P, blocks[ii], f)
                    if len(sub) > 1:
                        mean = N // len(sub)
                        dev = max(abs(len(s) - mean) for s in sub)
                        if best is None or dev < best[0]:
                            best = (dev, f)
                if best is not None:
                    fr = best[1]
            sub = _octree_split(P, blocks[ii], fr)
            if len(sub) < 2:
                raise ValueError("partition_sph: a block of coincident cells cannot be split")
            blocks[ii:ii + 1] = sub
        else:
            ii += 1
    order = np.array([c for b in blocks for c in b], np.int64)
    sizes = np.array([len(b) for b in blocks], np.int64)
    starts = np.concatenate([[0], np.cumsum(sizes)])
    fr = block_frames(P[order, 0], P[order, 1], P[order, 2],
                      [np.arange(starts[i], starts[i + 1]) for i in range(len(sizes))])
    return order, sizes, 2.0 * float(fr[:, 12].max())


def _distribute(pos, cells, nbl, dmin, dstep):
    """distributeCells (Division_blocks.f90:1668-1709): slab ceiling((c - dmin)
    / dstep) clamped to [1, nbl], cells in input order per slab."""
    out = [[] for _ in range(nbl)]
    for c in cells:
        with np.errstate(divide="ignore", invalid="ignore"):
            q = np.float64(pos[c] - dmin) / np.float64(dstep)
        out[min(max(_f_ceiling(q), 1), nbl) - 1].append(c)
    return out


NB_BLOCK_S_MAX = 10000   # NbBlock_s_max (Initialization.f90:54)


def _csh_once(P, order, sc, navg, h_block, box_min, dims, type_s):
    """Division_blocks_csh (Division_blocks.f90:190-652) on the cells in the
    given order.  Returns the new order and block sizes, or None where the
    reference sets error_division."""
    dirs = []
    d = list(dims)
    for _ in range(3):
        m = max(d)
        k = d.index(m)
        dirs.append(k)
        d[k] = 0.0
    dir0, dir1, dir2 = dirs
    hbox = dims[dir0]
    div = int(hbox / h_block + 1)
    h0 = hbox / div
    if div > 1:
        dstep = _f_nint(h0 / sc) * sc
        blocks = _distribute(P[:, dir0], order, div, box_min[dir0], dstep)
    else:
        blocks = [list(order)]
    if type_s == 3:
        blocks = [b for b in blocks if b]
        return [c for b in blocks for c in b], [len(b) for b in blocks]
    nbmax = min(min(len(b) for b in blocks), navg)
    possible = True
    while possible:
        possible = False
        new = []
        for blk in blocks:
            n = len(blk)
            if n > nbmax and abs(nbmax - n) > abs(nbmax - n // 2):
                possible = True
                q = P[blk]
                ext = q.max(0) - q.min(0)
                dd = dir1 if ext[dir1] > ext[dir2] else dir2
                h = ext[dd] / 2
                dmin = q[:, dd].min() - sc / 2.0
                pos = P[:, dd]
                sub = _distribute(pos, blk, 2, dmin, _f_nint(h / sc) * sc)
                ref = n // 2
                f32 = np.float32
                err1 = float(f32(abs(len(sub[0]) - ref)) / f32(ref)) if ref else float("nan")
                prev_err, prev_h = 2 * err1, None
                step, tr = h / 2.0, 1
                while err1 <= prev_err and tr <= 20:
                    prev_err, prev_h = err1, h
                    h = h - step if len(sub[0]) > len(sub[1]) else h + step
                    sub = _distribute(pos, blk, 2, dmin, _f_nint(h / sc) * sc)
                    err1 = float(f32(abs(len(sub[0]) - ref)) / f32(ref))
                    tr += 1
                    step = step / 2.0
                if prev_h is None:
                    return None
                sub = _distribute(pos, blk, 2, dmin, _f_nint(prev_h / sc) * sc)
                new.extend(sub)
            else:
                new.append(blk)
            if len(new) > NB_BLOCK_S_MAX:
                return None
        blocks = new
    return [c for b in blocks for c in b], [len(b) for b in blocks]


def partition_csh(x, y, z, sc, navg, type_s=2):
    """Division_blocks with div_type CSH (Division_blocks.f90:133-180): up to 8
    trial divisions (hBlock halved after each, the cells reordered by every
    trial), then the division with the hBlock whose mean block size came
    closest to navg (or hBlock = the largest box side when the scatterer has
    at most 2 navg cells).  The box is the scatterer's (Discretization.f90:
    510-519): centre extremes -+ sc/2.  type_s 3 (cylinder): slabs only.
    Returns (order, sizes, hblock)."""
    P = np.stack([np.asarray(x, float), np.asarray(y, float), np.asarray(z, float)], 1)
    nbc = len(P)
    box_min = P.min(0) - sc / 2.0
    dims = list((P.max(0) + sc / 2.0) - box_min)
    dmax = max(dims)
    h_block = dmax
    order = list(range(nbc))
    diff = [nbc] * 9
    htest = [None] * 9
    i, N = 0, nbc
    while i < 8 and N > navg:
        r = _csh_once(P, order, sc, navg, h_block, box_min, dims, type_s)
        if r is None:
            if i == 0:
                raise ValueError("partition_csh: the division failed")
            break
        order, sizes = r
        htest[i] = h_block
        N = nbc // len(sizes)
        diff[i] = abs(N - navg)
        h_block = h_block * 0.5
        i += 1
    if nbc <= 2 * navg:
        h_block = dmax
    else:
        h_block = htest[int(np.argmin(diff))]
    r = _csh_once(P, order, sc, navg, h_block, box_min, dims, type_s)
    if r is None:
        raise ValueError("partition_csh: the division failed")
    order, sizes = r
    order = np.array(order, np.int64)
    sizes = np.array(sizes, np.int64)
    starts = np.concatenate([[0], np.cumsum(sizes)])
    fr = block_frames(P[order, 0], P[order, 1], P[order, 2],
                      [np.arange(starts[i], starts[i + 1]) for i in range(len(sizes))])
    return order, sizes, 2.0 * float(fr[:, 12].max())


def solve(Z, B):
    """PZGESV semantics (Compute_EFields_MoM.f90:173): LU with partial pivoting."""
    return np.linalg.solve(Z, B)


def cbf_threshold(eps_cells, aca_variant):
    """Compute_EFields_CBFME.f90:286-288 (m = sqrt(eps)) and Compute_EFields_CBFME_ACA.f90:124.
    The literals 1e-3 and 1e-10 are single precision widened to double."""
    if aca_variant:
        return float(np.float32(1e-3))
    mmax = np.max(np.abs(np.sqrt(np.asarray(eps_cells))))
    return max(10.0 ** (-3.0 - (np.ceil(mmax) - 2.0)), float(np.float32(1e-10)))


def cbfm(cells: Cells, blocks, ext, th_cbf, ph_cbf, th, ph, use_aca=False, kmax=50, eps_aca=1e-4,
         threshold=None, fp32_blocks=True, sr=False, res_sr=0.01, fsr_out=None, iws=1, transpose="T"):
    """Dense-block CBFM-E (Compute_EFields_CBFME.f90:348-449, 630-849, 896-1006)
    or its ACA variant (Compute_EFields_CBFME_ACA.f90:296-385, 569-636).
    iws: CBF excitation, 1 plane waves, 2 / 3 spherical waves (type_iws,
    Compute_EFields_CBFME.f90:363-369).
    transpose: "T" is the reference's plain transpose C_i^T in Zc and Vc
    (ZGEMM 'T', Compute_EFields_CBFME.f90:666-669, 844-847); "H" (conjugate
    transpose) exists only so that tests can show their checks tell the two apart.

    blocks: list of 0-based cell-index arrays (contiguous ranges in block order);
    ext: list of extension cell-index arrays.  Returns (E, K per block, C list).
    """
    if threshold is None:
        threshold = cbf_threshold(cells.eps, use_aca)
    Cs, Ks = [], []
    for b, e in zip(blocks, ext):
        idx = np.concatenate([b, e]).astype(np.int64)
        cb = cells.subset(idx)
        if sr:   # SR mask (Compute_EFields_CBFME.f90:398-409), solved densely
            f, _ = setfsr(cb, res_sr)
            _, _, Zb = sr_block(cb, f)
            if fsr_out is not None:
                fsr_out.append(f)
        else:
            Zb = block_fill(cb, cb, fp32=fp32_blocks)
        Eref = incident_field_f32(cb, th_cbf, ph_cbf) if iws == 1 else incident_field_sph(cb, th_cbf, ph_cbf)
        Eb = np.linalg.solve(Zb, Eref)
        U, s, _ = np.linalg.svd(Eb, full_matrices=False)
        K = int(np.sum(s / s[0] >= threshold))
        Cs.append(np.asfortranarray(U[: 3 * len(b), :K]))
        Ks.append(K)
    if transpose not in ("T", "H"):
        raise ValueError(transpose)
    tr = (lambda M: M.T) if transpose == "T" else (lambda M: M.conj().T)
    offs = np.concatenate([[0], np.cumsum(Ks)])
    Kt = offs[-1]
    Zc = np.zeros((Kt, Kt), np.complex128)
    for i, bi in enumerate(blocks):
        ci = cells.subset(bi)
        for j, bj in enumerate(blocks):
            cj = cells.subset(bj)
            if use_aca and i != j:
                k, U, V, _, _ = aca(cells, int(bi[0]), len(bi), int(bj[0]), len(bj), kmax, eps_aca)
                if k < kmax:
                    Zc[offs[i]:offs[i + 1], offs[j]:offs[j + 1]] = tr(Cs[i]) @ (U @ (V @ Cs[j]))
                    continue
            Zij = block_fill(ci, cj, fp32=fp32_blocks)
            Zc[offs[i]:offs[i + 1], offs[j]:offs[j + 1]] = tr(Cs[i]) @ Zij @ Cs[j]
    Vc = np.zeros((Kt, 2 * len(th)), np.complex128)
    for i, bi in enumerate(blocks):
        Vc[offs[i]:offs[i + 1]] = tr(Cs[i]) @ incident_field_f32(cells.subset(bi), th, ph)
    alpha = np.linalg.solve(Zc, Vc)
    E = np.zeros((3 * cells.n, 2 * len(th)), np.complex128)
    for i, bi in enumerate(blocks):
        rows = (3 * np.asarray(bi)[:, None] + np.arange(3)[None, :]).ravel()
        E[rows] = Cs[i] @ alpha[offs[i]:offs[i + 1]]
    return E, Ks, Cs


# --------------------------------------------------------------------------
# The reference's own routines (oracle/_ref), same calling shapes as above
# --------------------------------------------------------------------------

def _ref_globals(r, k0, nbc=1, ntr=1, homogs=0, nb_it=50, eps_aca=1e-4):
    r.ref_set_globals(C.c_double(k0), C.c_int(nbc), C.c_int(ntr), C.c_int(homogs), C.c_int(nb_it),
                      C.c_double(eps_aca))


def ref_cell_params(sc, eps, k0):
    r = ref()
    _ref_globals(r, k0)
    sc, eps = _f64(sc), _c128(eps)
    n = len(sc)
    a, kap = np.empty(n), np.empty(n)
    che, zn = np.empty(n, np.complex128), np.empty(n, np.complex128)
    r.ref_cell_params(C.c_int(n), _p(sc), _p(eps), _p(a), _p(che), _p(zn), _p(kap))
    return a, che, zn, kap


def ref_block_fill(c1: Cells, c2: Cells):
    r = ref()
    _ref_globals(r, c1.k0)
    Z = np.empty((3 * c1.n, 3 * c2.n), np.complex128, order="F")
    r.ref_block_fill(C.c_int(c1.n), _p(c1.id), _p(c1.x), _p(c1.y), _p(c1.z), C.c_int(c2.n), _p(c2.id), _p(c2.x),
                     _p(c2.y), _p(c2.z), _p(c2.kappa), _p(c2.che), _p(c2.znnpp), _p(Z))
    return Z


def ref_total_fill(c: Cells):
    r = ref()
    _ref_globals(r, c.k0, nbc=c.n)
    Z = np.empty((3 * c.n, 3 * c.n), np.complex128, order="F")
    r.ref_total_fill(C.c_int(c.n), _p(c.x), _p(c.y), _p(c.z), _p(c.kappa), _p(c.che), _p(c.znnpp), _p(Z))
    return Z


def ref_block_row(c: Cells, grow1, col_inf1, ncols):
    r = ref()
    _ref_globals(r, c.k0, nbc=c.n)
    row = np.empty(ncols, np.complex128)
    r.ref_block_row(C.c_int(c.n), _p(c.x), _p(c.y), _p(c.z), _p(c.kappa), _p(c.che), C.c_int(grow1),
                    C.c_int(col_inf1), C.c_int(ncols), _p(row))
    return row


def ref_block_col(c: Cells, row_inf1, nrows, gcol1):
    r = ref()
    _ref_globals(r, c.k0, nbc=c.n)
    col = np.empty(nrows, np.complex128)
    r.ref_block_col(C.c_int(c.n), _p(c.x), _p(c.y), _p(c.z), _p(c.kappa), _p(c.che), C.c_int(row_inf1),
                    C.c_int(nrows), C.c_int(gcol1), _p(col))
    return col


def ref_incident_field(c: Cells, th, ph, t0=1, t1=None):
    r = ref()
    th, ph = _f64(th), _f64(ph)
    t1 = len(th) if t1 is None else t1
    _ref_globals(r, c.k0, nbc=c.n, ntr=len(th))
    E = np.empty((3 * c.n, 2 * (t1 - t0 + 1)), np.complex128, order="F")
    r.ref_incident_field(C.c_int(c.n), _p(c.x), _p(c.y), _p(c.z), C.c_int(len(th)), _p(th), _p(ph), C.c_int(t0),
                         C.c_int(t1), _p(E))
    return E


def ref_incident_field_sph(c: Cells, th, ph, t0=1, t1=None):
    r = ref()
    th, ph = _f64(th), _f64(ph)
    t1 = len(th) if t1 is None else t1
    _ref_globals(r, c.k0, nbc=c.n, ntr=len(th))
    E = np.empty((3 * c.n, 2 * (t1 - t0 + 1)), np.complex128, order="F")
    r.ref_incident_field_sph(C.c_int(c.n), _p(c.x), _p(c.y), _p(c.z), C.c_int(len(th)), _p(th), _p(ph),
                             C.c_int(t0), C.c_int(t1), _p(E))
    return E


def ref_green_s_dt(c: Cells, rso, th, ph):
    r = ref()
    _ref_globals(r, c.k0, nbc=c.n)
    G = np.empty((3, 3 * c.n), np.complex128, order="F")
    r.ref_green_s_dt(C.c_int(c.n), _p(c.x), _p(c.y), _p(c.z), _p(c.kappa), _p(c.che), C.c_double(rso),
                     C.c_double(th), C.c_double(ph), _p(G))
    return G


def ref_ff_coeff(c: Cells, th, ph):
    r = ref()
    _ref_globals(r, c.k0, nbc=c.n)
    ff = np.empty(c.n, np.complex128)
    r.ref_ff_coeff(C.c_int(c.n), _p(c.x), _p(c.y), _p(c.z), _p(c.kappa), _p(c.che), C.c_double(th), C.c_double(ph),
                   _p(ff))
    return ff


def ref_aca(c: Cells, r0, nr, c0, nc, kmax=50, eps=1e-4):
    """Reference ACA of cells[r0:r0+nr] x cells[c0:c0+nc] (0-based cell ranges)."""
    r = ref()
    _ref_globals(r, c.k0, nbc=c.n, nb_it=kmax, eps_aca=eps)
    U = np.zeros((3 * nr, kmax), np.complex128, order="F")
    V = np.zeros((kmax, 3 * nc), np.complex128, order="F")
    k = C.c_int(0)
    r.ref_aca(C.c_int(c.n), _p(c.x), _p(c.y), _p(c.z), _p(c.kappa), _p(c.che), C.c_int(3 * r0 + 1),
              C.c_int(3 * c0 + 1), C.c_int(3 * nr), C.c_int(3 * nc), C.byref(k), _p(U), _p(V))
    kk = min(k.value, kmax)
    return k.value, U[:, :kk], V[:kk, :]


def ref_aca_gen(A, kmax=50, eps=1e-4):
    r = ref()
    _ref_globals(r, 1.0, nb_it=kmax, eps_aca=eps)
    A = np.asfortranarray(A, dtype=np.complex128)
    m, n = A.shape
    U = np.zeros((m, kmax), np.complex128, order="F")
    V = np.zeros((kmax, n), np.complex128, order="F")
    k = C.c_int(0)
    r.ref_aca_gen(C.c_int(m), C.c_int(n), _p(A), C.c_int(kmax), C.byref(k), _p(U), _p(V))
    kk = min(k.value, kmax)
    return k.value, U[:, :kk], V[:kk, :]


def _ref_block_args(c: Cells):
    return (C.c_int(c.n), _p(c.id), _p(c.x), _p(c.y), _p(c.z), _p(c.kappa), _p(c.che), _p(c.znnpp))


def ref_block_fn(c: Cells):
    r = ref()
    _ref_globals(r, c.k0, nbc=c.n)
    fn = C.c_double(0)
    r.ref_block_fn(*_ref_block_args(c), C.byref(fn))
    return fn.value


def ref_sr_fn(c: Cells, fsr):
    r = ref()
    _ref_globals(r, c.k0, nbc=c.n)
    fn, nnz = C.c_double(0), C.c_int(0)
    r.ref_sr_fn(*_ref_block_args(c), C.c_double(fsr), C.byref(nnz), C.byref(fn))
    return nnz.value, fn.value


def ref_sr_csr(c: Cells, fsr):
    """SR_Green_s_tr_partial as (val, irow, icol), 1-based CSR as the reference builds it."""
    nnz, _ = ref_sr_fn(c, fsr)
    r = ref()
    val = np.zeros(nnz, np.complex128)
    irow = np.zeros(3 * c.n + 1, np.int32)
    icol = np.zeros(nnz, np.int32)
    r.ref_sr_csr(*_ref_block_args(c), C.c_double(fsr), C.c_int(nnz), _p(val), _p(irow), _p(icol))
    return val, irow, icol


def ref_setfsr(c: Cells, res_sr=0.01):
    r = ref()
    _ref_globals(r, c.k0, nbc=c.n)
    f, nnz = C.c_double(0), C.c_int(0)
    r.ref_setfsr(*_ref_block_args(c), C.c_double(res_sr), C.byref(f), C.byref(nnz))
    return f.value, nnz.value


def ref_transmitters_un(niws, nmax=100000):
    r = ref()
    th, ph = np.zeros(nmax), np.zeros(nmax)
    n = C.c_int(0)
    r.ref_transmitters_un(C.c_int(niws), C.c_int(nmax), _p(th), _p(ph), C.byref(n))
    return th[:n.value].copy(), ph[:n.value].copy()


def ref_adaptive_quad(f, nth, nph, ph, jj, a0, b0):
    """The reference's getAdaptiveQuadPts2 (compiled): f (d1 x d2) grid values."""
    r = ref()
    f = np.asfortranarray(f, dtype=np.float64)
    out = C.c_double(0)
    r.ref_adaptive_quad(C.c_int(nth), C.c_int(nph), C.c_int(f.shape[0]), C.c_int(f.shape[1]), _p(f),
                        C.c_int(1 if ph else 0), C.c_int(jj), C.c_double(a0), C.c_double(b0), C.byref(out))
    return out.value


def ref_trap2d(f, a, b, c, d):
    r = ref()
    f = np.asfortranarray(f, dtype=np.float64)
    out = C.c_double(0)
    r.ref_trap2d(C.c_int(f.shape[0]), C.c_int(f.shape[1]), _p(f), C.c_double(a), C.c_double(b), C.c_double(c),
                 C.c_double(d), C.byref(out))
    return out.value


def ref_gauss_legendre(n, a, b):
    r = ref()
    x, w = np.zeros(n), np.zeros(n)
    r.ref_gauss_legendre(C.c_int(n), C.c_double(a), C.c_double(b), _p(x), _p(w))
    return x, w


# ---------------------------------------------------------------------------
# Compute_Scattering_Quantities_1 (Compute_Scattering_Quantities.f90:1-589) and
# its quadrature rules, restated.  QUADPACK qag / qk61 (quadpack.f90; the
# published algorithm of Piessens et al. 1983, 61-point Gauss-Kronrod rule) in
# single precision, reading the integrand at the nearest grid node
# (getAdaptiveQuadPts2, Numerical_Integration.f90:213-245).
# ---------------------------------------------------------------------------
_F = np.float32
_XGK = np.array([0.999484410050490637571325895705811, 0.996893484074649540271630050918695,
                 0.991630996870404594858628366109486, 0.983668123279747209970032581605663,
                 0.973116322501126268374693868423707, 0.960021864968307512216871025581798,
                 0.944374444748559979415831324037439, 0.926200047429274325879324277080474,
                 0.905573307699907798546522558925958, 0.882560535792052681543116462530226,
                 0.857205233546061098958658510658944, 0.829565762382768397442898119732502,
                 0.799727835821839083013668942322683, 0.767777432104826194917977340974503,
                 0.733790062453226804726171131369528, 0.697850494793315796932292388026640,
                 0.660061064126626961370053668149271, 0.620526182989242861140477556431189,
                 0.579345235826361691756024932172540, 0.536624148142019899264169793311073,
                 0.492480467861778574993693061207709, 0.447033769538089176780609900322854,
                 0.400401254830394392535476211542661, 0.352704725530878113471037207089374,
                 0.304073202273625077372677107199257, 0.254636926167889846439805129817805,
                 0.204525116682309891438957671002025, 0.153869913608583546963794672743256,
                 0.102806937966737030147096751318001, 0.051471842555317695833025213166723, 0.0], _F)
_WGK = np.array([0.001389013698677007624551591226760, 0.003890461127099884051267201844516,
                 0.006630703915931292173319826369750, 0.009273279659517763428441146892024,
                 0.011823015253496341742232898853251, 0.014369729507045804812451432443580,
                 0.016920889189053272627572289420322, 0.019414141193942381173408951050128,
                 0.021828035821609192297167485738339, 0.024191162078080601365686370725232,
                 0.026509954882333101610601709335075, 0.028754048765041292843978785354334,
                 0.030907257562387762472884252943092, 0.032981447057483726031814191016854,
                 0.034979338028060024137499670731468, 0.036882364651821229223911065617136,
                 0.038678945624727592950348651532281, 0.040374538951535959111995279752468,
                 0.041969810215164246147147541285970, 0.043452539701356069316831728117073,
                 0.044814800133162663192355551616723, 0.046059238271006988116271735559374,
                 0.047185546569299153945261478181099, 0.048185861757087129140779492298305,
                 0.049055434555029778887528165367238, 0.049795683427074206357811569379942,
                 0.050405921402782346840893085653585, 0.050881795898749606492297473049805,
                 0.051221547849258772170656282604944, 0.051426128537459025933862879215781,
                 0.051494729429451567558340433647099], _F)
_WG = np.array([0.007968192496166605615465883474674, 0.018466468311090959142302131912047,
                0.028784707883323369349719179611292, 0.038799192569627049596801936446348,
                0.048402672830594052902938140422808, 0.057493156217619066481721689402056,
                0.065974229882180495128128515115962, 0.073755974737705206268243850022191,
                0.080755895229420215354694938460530, 0.086899787201082979802387530715126,
                0.092122522237786128717632707087619, 0.096368737174644259639468626351810,
                0.099593420586795267062780282103569, 0.101762389748405504596428952168554,
                0.102852652893558840341285636705415], _F)
_FEPS, _FTINY = _F(np.finfo(np.float32).eps), _F(np.finfo(np.float32).tiny)


def _nint32(v):
    v = float(v)
    return int(np.floor(abs(v) + 0.5)) * (1 if v >= 0 else -1)


def _qk61(fcol, n, a0, b0, a, b):
    """qk61 (quadpack.f90:486-725) on the grid column fcol (float64 values,
    read as float32 at node nint(x (n-1) / (b0 - a0)) + 1)."""
    def at(x):
        return fcol[_nint32(_F(_F(x) * _F(n - 1)) / _F(b0 - a0))]
    centr, hlgth = _F(_F(0.5) * _F(b + a)), _F(_F(0.5) * _F(b - a))
    dhlgth = abs(hlgth)
    resg = _F(0)
    fc = at(centr)
    resk = _F(_WGK[30] * fc)
    resabs = abs(resk)
    fv1, fv2 = [_F(0)] * 30, [_F(0)] * 30
    for j in range(1, 16):
        jtw = 2 * j
        absc = _F(hlgth * _XGK[jtw - 1])
        f1, f2 = at(_F(centr - absc)), at(_F(centr + absc))
        fv1[jtw - 1], fv2[jtw - 1] = f1, f2
        fsum = _F(f1 + f2)
        resg = _F(resg + _F(_WG[j - 1] * fsum))
        resk = _F(resk + _F(_WGK[jtw - 1] * fsum))
        resabs = _F(resabs + _F(_WGK[jtw - 1] * _F(abs(f1) + abs(f2))))
    for j in range(1, 16):
        jm = 2 * j - 1
        absc = _F(hlgth * _XGK[jm - 1])
        f1, f2 = at(_F(centr - absc)), at(_F(centr + absc))
        fv1[jm - 1], fv2[jm - 1] = f1, f2
        fsum = _F(f1 + f2)
        resk = _F(resk + _F(_WGK[jm - 1] * fsum))
        resabs = _F(resabs + _F(_WGK[jm - 1] * _F(abs(f1) + abs(f2))))
    reskh = _F(resk * _F(0.5))
    resasc = _F(_WGK[30] * abs(_F(fc - reskh)))
    for j in range(30):
        resasc = _F(resasc + _F(_WGK[j] * _F(abs(_F(fv1[j] - reskh)) + abs(_F(fv2[j] - reskh)))))
    result = _F(resk * hlgth)
    resabs = _F(resabs * dhlgth)
    resasc = _F(resasc * dhlgth)
    abserr = abs(_F(_F(resk - resg) * hlgth))
    if resasc != 0 and abserr != 0:
        abserr = _F(resasc * min(_F(1), _F(np.power(_F(_F(_F(200) * abserr) / resasc), _F(1.5)))))
    if resabs > _F(_FTINY / _F(_F(50) * _FEPS)):
        abserr = max(_F(_F(_FEPS * _F(50)) * resabs), abserr)
    return result, abserr, resabs, resasc


def _qpsrt(limit, last, maxerr, elist, iord, nrmax):
    """qsort of quadpack.f90 (QUADPACK qpsrt), 1-based indices."""
    E = lambda i: elist[i - 1]   # noqa: E731
    if last <= 2:
        iord[0], iord[1] = 1, 2
    else:
        errmax = E(maxerr)
        for _ in range(1, nrmax):
            isucc = iord[nrmax - 2]
            if errmax <= E(isucc):
                break
            iord[nrmax - 1] = isucc
            nrmax -= 1
        jupbn = last
        if limit // 2 + 2 < last:
            jupbn = limit + 3 - last
        errmin = E(last)
        jbnd, ibeg = jupbn - 1, nrmax + 1
        i, found = ibeg, False
        while i <= jbnd:
            isucc = iord[i - 1]
            if E(isucc) <= errmax:
                found = True
                break
            iord[i - 2] = isucc
            i += 1
        if not found:
            iord[jbnd - 1] = maxerr
            iord[jupbn - 1] = last
        else:
            iord[i - 2] = maxerr
            k, placed = jbnd, False
            for _ in range(i, jbnd + 1):
                isucc = iord[k - 1]
                if errmin < E(isucc):
                    iord[k] = last
                    placed = True
                    break
                iord[k] = isucc
                k -= 1
            if not placed:
                iord[i - 1] = last
    maxerr = iord[nrmax - 1]
    return maxerr, E(maxerr), nrmax


def adaptive_quad(f, nth, nph, ph, jj, a0, b0):
    """getAdaptiveQuadPts2: qag (key 6, epsabs 0, epsrel 1e-3, limit 500) of
    column jj (1-based) of f over [a0, b0], single precision."""
    fcol = np.asarray(f, np.float64)[:, jj - 1].astype(np.float32)   # read as single precision
    n = nph if ph else nth
    a, b = _F(a0), _F(b0)
    A0, B0 = a, b
    limit, epsabs, epsrel = 500, _F(0), _F(0.001)
    alist, blist, rlist, elist = [_F(0)] * limit, [_F(0)] * limit, [_F(0)] * limit, [_F(0)] * limit
    iord = [0] * limit
    ier = 0
    alist[0], blist[0] = a, b
    result, abserr, defabs, resabs = _qk61(fcol, n, A0, B0, a, b)
    rlist[0], elist[0], iord[0] = result, abserr, 1
    errbnd = max(epsabs, _F(epsrel * abs(result)))
    if abserr <= _F(_F(_F(50) * _FEPS) * defabs) and errbnd < abserr:
        ier = 2
    if ier != 0 or (abserr <= errbnd and abserr != resabs) or abserr == 0:
        return float(result)
    errmax, maxerr, area, errsum, nrmax, iroff1, iroff2 = abserr, 1, result, abserr, 1, 0, 0
    last = 1
    for last in range(2, limit + 1):
        a1 = alist[maxerr - 1]
        b1 = _F(_F(0.5) * _F(alist[maxerr - 1] + blist[maxerr - 1]))
        a2, b2 = b1, blist[maxerr - 1]
        area1, error1, _, defab1 = _qk61(fcol, n, A0, B0, a1, b1)
        area2, error2, _, defab2 = _qk61(fcol, n, A0, B0, a2, b2)
        area12, erro12 = _F(area1 + area2), _F(error1 + error2)
        errsum = _F(_F(errsum + erro12) - errmax)
        area = _F(_F(area + area12) - rlist[maxerr - 1])
        if defab1 != error1 and defab2 != error2:
            if abs(_F(rlist[maxerr - 1] - area12)) <= _F(_F(1e-5) * abs(area12)) and _F(_F(0.99) * errmax) <= erro12:
                iroff1 += 1
            if 10 < last and errmax < erro12:
                iroff2 += 1
        rlist[maxerr - 1], rlist[last - 1] = area1, area2
        errbnd = max(epsabs, _F(epsrel * abs(area)))
        if errbnd < errsum:
            if iroff1 >= 6 or iroff2 >= 20:
                ier = 2
            if last == limit:
                ier = 1
            if max(abs(a1), abs(b2)) <= _F(_F(_F(1) + _F(_F(_F(6) * _F(1e3)) * _FEPS)) *
                                           _F(abs(a2) + _F(_F(1e4) * _FTINY))):
                ier = 3
        if error2 <= error1:
            alist[last - 1], blist[maxerr - 1], blist[last - 1] = a2, b1, b2
            elist[maxerr - 1], elist[last - 1] = error1, error2
        else:
            alist[maxerr - 1], alist[last - 1], blist[last - 1] = a2, a1, b1
            rlist[maxerr - 1], rlist[last - 1] = area2, area1
            elist[maxerr - 1], elist[last - 1] = error2, error1
        maxerr, errmax, nrmax = _qpsrt(limit, last, maxerr, elist, iord, nrmax)
        if ier != 0 or errsum <= errbnd:
            break
    s = _F(0)
    for v in rlist[:last]:
        s = _F(s + v)
    return float(s)


def trap2d(f, a, b, c, d):
    """trap_2Dc (Numerical_Integration.f90:251-305)."""
    f = np.asarray(f, np.float64)
    nx, ny = f.shape
    with np.errstate(divide="ignore", invalid="ignore"):
        hx, hy = (b - a) / np.float64(nx - 1), (d - c) / np.float64(ny - 1)
    s = f[0, 0] + f[0, ny - 1] + f[nx - 1, 0] + f[nx - 1, ny - 1]
    for i in range(1, nx - 1):
        s = s + 2.0 * (f[i, 0] + f[i, ny - 1])
    for j in range(1, ny - 1):
        s = s + 2.0 * (f[0, j] + f[nx - 1, j])
    for i in range(1, nx - 1):
        for j in range(1, ny - 1):
            s = s + 4.0 * f[i, j]
    with np.errstate(invalid="ignore"):
        return 0.25 * hx * hy * s


def gauss_legendre(n, a, b):
    """Gauss-Legendre rule on [a, b], nodes descending (getGaussLegendreQuadPts3)."""
    x, w = np.polynomial.legendre.leggauss(n)
    return ((a * (1 - x) + b * (1 + x)) / 2)[::-1], (w * (b - a) / 2)[::-1]


INT_SD, INT_LB, INT_UN, INT_GL, INT_TR, INT_SM = range(6)


def _unique_first(v):
    out = []
    for x in v:
        if x not in out:
            out.append(x)
    return out


def _txrx(th_all, ph_all, th, ph):
    """getTxRxIndex (get_trans_Receiv.f90:690-713), 0-based, -1 if none."""
    eps = float(np.float32(1e-2))
    for i, (t, p) in enumerate(zip(th_all, ph_all)):
        if abs(t - th) <= eps and abs(p - ph) <= eps:
            return i
    return -1


def scattering_quantities_grid(S, nrx, rx_th, rx_ph, rx_type, tr_th, tr_ph, tr_type, k0, dm, cext, cabs,
                               rx_w=None, cext_intfields=False, quirks=15):
    """Compute_Scattering_Quantities_1 (Compute_Scattering_Quantities.f90:1-589):
    per transmitter Q_ext (forward S, or internal field), Q_abs (internal
    field), Q_sca and g (receiver integral by the rule rx_type), Q_bks
    (backward S); averages over the transmitter grid by tr_type.  rx_th,
    rx_ph hold all NRx_tot receivers (the forward / backward ones are looked
    up there).  Returns (Q (ntr x 5), Q_av (5))."""
    pi = float(np.float32(np.arccos(np.float32(-1)))) if quirks & 1 else np.pi
    ap, X = dm / 2.0, k0 * dm / 2.0
    S = np.asarray(S)
    ntr = len(tr_th)
    sq = lambda z: abs(z) ** 2   # noqa: E731
    tth, tph = _unique_first(list(tr_th)), _unique_first(list(tr_ph))
    ntt, ntp = len(tth), len(tph)
    trth_min, trth_max = tth[0] * pi / 180., tth[-1] * pi / 180.
    trph_min, trph_max = tph[0] * pi / 180., tph[-1] * pi / 180.
    one_d = rx_type in (INT_SD, INT_LB)
    if not one_d:
        rth, rph = _unique_first(list(rx_th[:nrx])), _unique_first(list(rx_ph[:nrx]))
        nrt, nrp = len(rth), len(rph)
        rth_min, rth_max = rth[0] * pi / 180., rth[-1] * pi / 180.
        rph_min, rph_max = rph[0] * pi / 180., rph[-1] * pi / 180.
        sn_r = max(rph_max - rph_min, 1.0) * (math.cos(rth_min) - math.cos(rth_max))
    Q = np.zeros((ntr, 5))
    for t in range(ntr):
        thi, phi = tr_th[t], tr_ph[t]
        ki = np.array([-math.sin(thi * pi / 180.) * math.cos(phi * pi / 180.),
                       -math.sin(thi * pi / 180.) * math.sin(phi * pi / 180.), -math.cos(thi * pi / 180.)])

        def cg(ts, ps):
            ks = (-math.sin(ts * pi / 180.) * math.cos(ps * pi / 180.), -math.sin(ts * pi / 180.) * math.sin(ps * pi / 180.),
                  -math.cos(ts * pi / 180.))
            return ki[0] * ks[0] + ki[1] * ks[1] + ki[2] * ks[2]

        def s_it(r):
            return ((sq(S[r, 4 * t + 3]) + sq(S[r, 4 * t])) + sq(S[r, 4 * t + 1])) + sq(S[r, 4 * t + 2])
        ifw = _txrx(rx_th, rx_ph, thi, phi)
        q_ext = (2 / (X * X)) * (abs(S[ifw, 4 * t + 3].imag) + abs(S[ifw, 4 * t].imag))
        q_sca = g = 0.0
        if one_d:
            f = [(rx_w[r] if rx_type == INT_LB else 1.0) * s_it(r) for r in range(nrx)]
            sf = sum(f)
            sg = sum(fr * cg(rx_th[r], rx_ph[r]) for r, fr in enumerate(f))
            if rx_type == INT_LB:
                c_sca = 1. / (2 * k0 ** 2) * sf
                q_sca = (4 * pi) / (2 * pi * X * X) * sf
                g = 1. / (2. * k0 ** 2. * c_sca) * sg
            else:
                q_sd = (4 * pi / nrx) * sf
                c_sca = 1. / (2 * k0 ** 2) * q_sd
                q_sca = 1 / (2 * pi * X * X) * q_sd
                g = 1. / (2. * k0 ** 2. * c_sca) * (4 * pi / nrx) * sg
        else:
            fxy, fxyg = np.zeros((nrt, nrp)), np.zeros((nrt, nrp))
            for ii in range(nrt):
                s = math.sin(rth[ii] * pi / 180.)
                for jj in range(nrp):
                    v = s_it(jj * nrt + ii)
                    fxy[ii, jj] = v * s
                    fxyg[ii, jj] = v * s * cg(rth[ii], rph[jj])
            qi = qg = 0.0
            if rx_type == INT_GL:
                _, wt = gauss_legendre(nrt, rth_min, rth_max)
                wp = gauss_legendre(nrp, rph_min, rph_max)[1] if nrp > 1 else np.ones(1)
                wt, wp = wt[::-1], wp[::-1]
                for ii in range(nrt):
                    for jj in range(nrp):
                        qi = qi + wp[jj] * wt[ii] * fxy[ii, jj]
                for ii in range(nrt):
                    for jj in range(nrp):
                        qg = qg + wp[jj] * wt[ii] * fxyg[ii, jj]
            elif rx_type == INT_UN:
                fy = np.array([[adaptive_quad(fxy, nrt, nrp, False, jj, rth_min, rth_max)] for jj in range(1, nrp + 1)])
                qi = adaptive_quad(fy, nrt, nrp, True, 1, rph_min, rph_max) if nrp > 1 else fy[0, 0]
                fy = np.array([[adaptive_quad(fxyg, nrt, nrp, False, jj, rth_min, rth_max)] for jj in range(1, nrp + 1)])
                qg = adaptive_quad(fy, nrt, nrp, True, 1, rph_min, rph_max) if nrp > 1 else fy[0, 0]
            elif rx_type == INT_TR:
                qi = trap2d(fxy, rth_min, rth_max, rph_min, rph_max)
                qg = trap2d(fxyg, rth_min, rth_max, rph_min, rph_max)
            if rx_type != INT_SM:
                c_sca = 1. / (2 * k0 ** 2) * qi
                q_sca = 2. / (sn_r * X * X) * qi
                g = 1. / (2. * k0 ** 2. * c_sca) * qg
        ibk = _txrx(rx_th, rx_ph, 180. - thi, np.fmod(phi + 180., 360.))
        q_bks = (1. / (2. * pi * X * X)) * (((sq(S[ibk, 4 * t]) + sq(S[ibk, 4 * t + 3])) + sq(S[ibk, 4 * t + 1])) +
                                          sq(S[ibk, 4 * t + 2]))
        Q[t] = [(cext[t].real / (pi * ap * ap)) if cext_intfields else q_ext, cabs[t].real / (pi * ap * ap), q_sca,
                q_bks, g]
    if ntr == 1:
        return Q, Q[0].copy()
    sn_t = max(trph_max - trph_min, 1.0) * (math.cos(trth_min) - math.cos(trth_max))
    G = np.zeros((5, ntt, ntp))
    for ii in range(ntt):
        s = math.sin(tr_th[ii] * pi / 180.)
        for jj in range(ntp):
            G[:, ii, jj] = Q[jj * ntt + ii] * s
    integ = np.zeros(5)
    if tr_type == INT_GL:
        _, w1 = gauss_legendre(ntt, trth_min, trth_max)
        w2 = gauss_legendre(ntp, trph_min, trph_max)[1] if ntp > 1 else np.ones(1)
        w1, w2 = w1[::-1], w2[::-1]
        for ii in range(ntt):
            for jj in range(ntp):
                integ = integ + w2[jj] * w1[ii] * G[:, ii, jj]
    elif tr_type == INT_UN:
        for j in range(5):
            fy = np.array([[adaptive_quad(G[j], ntt, ntp, False, jj, trth_min, trth_max)] for jj in range(1, ntp + 1)])
            integ[j] = adaptive_quad(fy, ntt, ntp, True, 1, trph_min, trph_max) if ntp > 1 else fy[0, 0]
    elif tr_type == INT_TR:
        for j in range(5):
            integ[j] = trap2d(G[j], trth_min, trth_max, trph_min, trph_min)
    return Q, 1. / sn_t * integ
