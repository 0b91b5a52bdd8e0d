"""Writes tests/golden/ref_vectors.npz: inputs and outputs of the reference's
own routines, compiled from the reference sources into oracle/_ref by
oracle/Makefile.ref (flang).  Run in a container that has the reference tree:

    make -C oracle ref && python tests/golden/make_golden.py

Routines captured (reference file:line):
  SetCellsParams            SetCellsParams.f90:1-50
  Green_s_tr_total          Green_functions.f90:5-93   (full matrix, FP32 intermediates)
  Green_s_tr_partial        Green_functions.f90:96-173
  computeBlockRow / Col     Green_functions.f90:905-1012, 1119-1225
  Incident_Field            Incident_Field.f90:1-60
  GetFFieldCoeff            Green_functions.f90:1399-1438
  Calcul_MatrixZij_ACA      Calcul_ACA_matrices.f90:6-219 (converged and k = kmax)
  Calcul_GenMatrix_ACA      Calcul_ACA_matrices.f90:222-424
  Green_s_tr_partial_FN     Green_functions.f90:176-287
  SR_Green_s_tr_partial(_FN) Green_functions.f90:290-806 (CSR mask, norm)
  setfSR                    getParameters_CBFM.f90:113-168
  getTransmitters_CBFM      getParameters_CBFM.f90:1-110 (uniform grid)
  Incident_Field_Spherical  Incident_Field.f90:65-166
  Green_s_dt                Green_functions.f90:1335-1397

The fixture holds data only (inputs and expected outputs).
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as o  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_vectors.npz")


def cluster(rng, n, centre, sc):
    """n cells on a jittered lattice around `centre` (metres)."""
    side = int(np.ceil(n ** (1 / 3)))
    g = np.stack(np.meshgrid(*[np.arange(side)] * 3, indexing="ij"), -1).reshape(-1, 3)[:n]
    p = (g - (side - 1) / 2) * sc + np.asarray(centre) + rng.normal(scale=0.02 * sc, size=(n, 3))
    return p


def main():
    if o.ref() is None:
        raise SystemExit("oracle/_ref/libmidas_ref.so is not built (make -C oracle ref)")
    rng = np.random.default_rng(20240611)
    d = {}

    # --- cell parameters: spread of sizes and permittivities -----------------
    k0 = 2 * np.pi / 1.0e-3
    sc = 1e-5 * (1.0 + 3.0 * rng.random(16))
    eps = (1.2 + 2.5 * rng.random(16)) + 1j * (0.3 * rng.random(16))
    eps[0], eps[1] = 1.0 + 0j, 3.0 + 1.0j
    a, che, zn, kap = o.ref_cell_params(sc, eps, k0)
    d.update(cp_sc=sc, cp_eps=eps, cp_k0=np.float64(k0), cp_a=a, cp_che=che, cp_zn=zn, cp_kap=kap)

    # --- a small two-material scatterer (30 cells) ----------------------------
    sc1 = 2.0e-5
    p = cluster(rng, 30, (0, 0, 0), sc1)
    eps1 = np.where(rng.random(30) < 0.5, 1.6 + 0.02j, 2.5 + 0.4j)
    k1 = 2 * np.pi / 2.0e-4
    d.update(g1_x=p[:, 0].copy(), g1_y=p[:, 1].copy(), g1_z=p[:, 2].copy(), g1_sc=np.full(30, sc1), g1_eps=eps1,
             g1_k0=np.float64(k1))
    c1 = o.Cells(p[:, 0], p[:, 1], p[:, 2], np.full(30, sc1), eps1, k1)
    rc = o.Cells(p[:, 0], p[:, 1], p[:, 2], np.full(30, sc1), eps1, k1)
    a, che, zn, kap = o.ref_cell_params(rc.sc, rc.eps, k1)
    rc.a, rc.che, rc.znnpp, rc.kappa = a, che, zn, kap
    d["g1_total"] = o.ref_total_fill(rc)
    r_idx, c_idx = np.arange(0, 15), np.arange(10, 30)
    d["g1_part_rows"], d["g1_part_cols"] = r_idx, c_idx
    d["g1_partial"] = o.ref_block_fill(rc.subset(r_idx), rc.subset(c_idx))
    d["g1_partial_diag"] = o.ref_block_fill(rc.subset(c_idx), rc.subset(c_idx))
    # block rows / columns: 1-based global row (col) and 1-based first column (row)
    d["g1_row_args"] = np.array([[7, 31, 45], [1, 1, 90], [58, 4, 30]], np.int32)
    d["g1_rows"] = np.concatenate([o.ref_block_row(rc, *map(int, a_)) for a_ in d["g1_row_args"]])
    d["g1_col_args"] = np.array([[31, 45, 7], [1, 90, 90], [4, 30, 58]], np.int32)
    d["g1_cols"] = np.concatenate([o.ref_block_col(rc, *map(int, a_)) for a_ in d["g1_col_args"]])
    th = np.array([0.0, 37.5, 90.0, 151.0])
    ph = np.array([0.0, 12.0, 200.0, 333.3])
    d["g1_th"], d["g1_ph"] = th, ph
    d["g1_einc"] = o.ref_incident_field(rc, th, ph)
    d["g1_ff"] = np.stack([o.ref_ff_coeff(rc, float(t), float(f)) for t, f in zip(th, ph)])

    # --- ACA: two separated clusters (converges) and two touching (k = kmax) --
    sc2 = 2.0e-5
    pa = cluster(rng, 48, (0, 0, 0), sc2)
    pb = cluster(rng, 40, (9 * sc2, 3 * sc2, 0), sc2)
    pc = cluster(rng, 27, (0, 5 * sc2, 0), sc2)
    pp = np.concatenate([pa, pb, pc])
    n2 = len(pp)
    eps2 = np.where(np.arange(n2) < 48, 1.5 + 0.01j, 1.8 + 0.6j)
    k2 = 2 * np.pi / 3.0e-4
    d.update(g2_x=pp[:, 0].copy(), g2_y=pp[:, 1].copy(), g2_z=pp[:, 2].copy(), g2_sc=np.full(n2, sc2), g2_eps=eps2,
             g2_k0=np.float64(k2))
    rc2 = o.Cells(pp[:, 0], pp[:, 1], pp[:, 2], np.full(n2, sc2), eps2, k2)
    a, che, zn, kap = o.ref_cell_params(rc2.sc, rc2.eps, k2)
    rc2.a, rc2.che, rc2.znnpp, rc2.kappa = a, che, zn, kap
    cases = [(0, 48, 48, 40, 50, 1e-4), (48, 40, 0, 48, 50, 1e-6), (0, 48, 88, 27, 6, 1e-4)]
    d["g2_aca_args"] = np.array([c[:5] for c in cases], np.int64)
    d["g2_aca_eps"] = np.array([c[5] for c in cases])
    for i, (r0, nr, c0, nc, km, ep) in enumerate(cases):
        k, U, V = o.ref_aca(rc2, r0, nr, c0, nc, kmax=km, eps=ep)
        d[f"g2_aca{i}_k"], d[f"g2_aca{i}_U"], d[f"g2_aca{i}_V"] = np.int64(k), U, V

    # --- ACA of an explicit matrix -------------------------------------------
    m, n, r = 60, 45, 6
    A = (rng.normal(size=(m, r)) + 1j * rng.normal(size=(m, r))) @ (
        rng.normal(size=(r, n)) + 1j * rng.normal(size=(r, n)))
    A += 1e-7 * (rng.normal(size=(m, n)) + 1j * rng.normal(size=(m, n)))
    k, U, V = o.ref_aca_gen(A, kmax=20, eps=1e-4)
    d.update(gen_A=A, gen_k=np.int64(k), gen_U=U, gen_V=V)

    # --- SR mask of an extended block (two materials) and the fSR search -----
    sr_idx = np.concatenate([np.arange(0, 48), np.arange(88, 100)])
    cb = rc2.subset(sr_idx)
    d["sr_idx"] = sr_idx
    d["sr_fn"] = np.float64(o.ref_block_fn(cb))
    d["sr_fsr_list"] = np.array([1e2, 1e3, 5e3, 1e4])
    d["sr_nnz"] = np.array([o.ref_sr_fn(cb, f)[0] for f in d["sr_fsr_list"]], np.int64)
    d["sr_fns"] = np.array([o.ref_sr_fn(cb, f)[1] for f in d["sr_fsr_list"]])
    val, irow, icol = o.ref_sr_csr(cb, 1e3)
    d["sr_csr_val"], d["sr_csr_irow"], d["sr_csr_icol"] = val, irow, icol
    for res in (0.01, 0.001):
        f, nnz = o.ref_setfsr(cb, res)
        d[f"sr_setfsr_{res}"] = np.array([f, nnz])

    # --- CBF excitation grid (uniform in cos theta and phi) -------------------
    for ni in (94, 422):
        th, ph = o.ref_transmitters_un(ni)
        d[f"un_{ni}_th"], d[f"un_{ni}_ph"] = th, ph

    # --- spherical-wave CBF excitation (Incident_Field_Spherical) -----------
    # the 30-cell scatterer (Rsource = 10 wavelengths) and a block wider than a
    # wavelength (Rsource = 10 block radii) on the 94-direction grid
    d["sph1_E"] = o.ref_incident_field_sph(rc, d["g1_th"], d["g1_ph"])
    pw = cluster(rng, 120, (1e-4, -2e-4, 3e-4), 1.5e-4)
    kw = 2 * np.pi / 2.5e-4
    cw = o.Cells(pw[:, 0], pw[:, 1], pw[:, 2], np.full(120, 1.5e-4), np.full(120, 1.5 + 0.1j), kw)
    d.update(sph2_x=pw[:, 0].copy(), sph2_y=pw[:, 1].copy(), sph2_z=pw[:, 2].copy(), sph2_k0=np.float64(kw))
    d["sph2_E"] = o.ref_incident_field_sph(cw, d["un_94_th"], d["un_94_ph"])

    # --- near-field dyad Green_s_dt (receivers at Rso) ------------------------
    d["gdt_args"] = np.array([[0.05, 0.0, 0.0], [0.05, 37.5, 12.0], [3e-4, 151.0, 333.3]])
    d["gdt_G"] = np.stack([o.ref_green_s_dt(rc, *map(float, a_)) for a_ in d["gdt_args"]])

    np.savez_compressed(OUT, **d)
    print("wrote", OUT, os.path.getsize(OUT), "bytes;", {k_: v.shape for k_, v in d.items() if np.ndim(v) > 1})


if __name__ == "__main__":
    main()
