/*
 * MIDAS hot-path ORACLE -- CPU restatement of the reference arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline).  The product (midas_amd, libmidas_gpu.so) never
 * links, loads or calls it.
 *
 * Every routine follows one reference routine; the citation is on each.
 * Arithmetic is FP64 unless a routine states that it reproduces the
 * reference's single-precision (default `Complex`) intermediates.
 *
 * Conventions (same as the reference):
 *   - unknown index i = 3*cell + p (0-based here), p in {x,y,z};
 *   - matrices column-major, complex interleaved (re, im);
 *   - angles in degrees, theta measured from +x (Incident_Field.f90:31-34).
 *
 * Pinning (how each routine is checked, see tests/test_oracle.py):
 *   or_cell_params        : reference SetCellsParams compiled in oracle/_ref;
 *                           static limit (Znnpp -> -chi/3, kappa -> Sc^3).
 *   or_mom_fill           : reference Green_s_tr_total (FP32 intermediates) in
 *                           oracle/_ref to 1e-6; complex symmetry for homogeneous
 *                           chi*kappa; hand-computed 2-cell case; static dipole.
 *   or_block_fill_f32     : reference Green_s_tr_partial in oracle/_ref.
 *   or_block_row/col_f32  : reference computeBlockRow/computeBlockCol.
 *   or_mom_rhs            : closed form |E_V| = |E_H| = 1, transversality, and
 *                           reference Incident_Field (FP32) to 1e-6.
 *   or_incident_field_f32 : reference Incident_Field in oracle/_ref.
 *   or_ff_coeff           : reference GetFFieldCoeff in oracle/_ref.
 *   or_scattering_matrix  : beta=0 reduces to the MoM projection; Mie series
 *                           for the sphere (physics, ~1e-2); optical theorem.
 *   or_cext_cabs          : dot-product variant (q6 = 0): Mie series, energy
 *                           balance and optical theorem; component-sum variant
 *                           (q6 = 1, the reference's formula): identities it
 *                           fixes against the dot product (E = (a, a, a) gives
 *                           3x Cext and 3x Cabs, E = (a, -a, 0) gives 0), and a
 *                           one-cell case at the origin worked by hand.
 *   or_aca                : reference Calcul_MatrixZij_ACA in oracle/_ref
 *                           (pivots, rank, U, V).
 *   or_aca_gen            : reference Calcul_GenMatrix_ACA in oracle/_ref.
 *   or_block_fn           : reference Green_s_tr_partial_FN in oracle/_ref.
 *   or_sr_block           : reference SR_Green_s_tr_partial (CSR) and
 *                           SR_Green_s_tr_partial_FN in oracle/_ref; the fSR
 *                           search built on them against setfSR.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double _Complex zc;
typedef float _Complex fc;

/* Initialization.f90:39  Pi = Acos(-1.) evaluates acos of a single literal
 * (quirk Q1).  quirk=0 gives the exact double pi. */
double or_pi(int quirk) { return quirk ? (double)acosf(-1.0f) : acos(-1.0); }

static inline zc ld(const double *p, long i) { return p[2 * i] + I * p[2 * i + 1]; }
static inline void st(double *p, long i, zc v) { p[2 * i] = creal(v); p[2 * i + 1] = cimag(v); }

/* SetCellsParams.f90:37-48.  With quirk!=0 the single-precision literals
 * (0.75/Pi)**(1./3.) and (2./3.) of the reference are reproduced (Q2). */
void or_cell_params(int n, const double *sc, const double *eps, double k0, int quirk,
                    double *a_out, double *che_out, double *znnpp_out, double *kappa_out)
{
    const double pi = or_pi(quirk);
    const double third = quirk ? (double)(1.0f / 3.0f) : 1.0 / 3.0;
    const double twothird = quirk ? (double)(2.0f / 3.0f) : 2.0 / 3.0;
    for (int c = 0; c < n; ++c) {
        double a = sc[c] * pow(0.75 / pi, third);
        zc che = ld(eps, c) - 1.0;
        zc e = cexp(I * k0 * a);
        zc znnpp = (twothird * e * (1.0 - I * k0 * a) - 1.0) * che;
        double kap = 4.0 * pi * (sin(k0 * a) - k0 * a * cos(k0 * a)) / pow(k0, 3.0);
        a_out[c] = a;
        st(che_out, c, che);
        st(znnpp_out, c, znnpp);
        kappa_out[c] = kap;
    }
}

/* compute_Green_s_tr_single_element (Compute_EFields_MoM.f90:433-523), FP64,
 * for receiver cell o, source cell s, components po (row), ps (column). */
static zc mom_elem(int o, int s, int po, int ps, const double *x, const double *y, const double *z,
                   const double *kappa, const double *che, const double *znnpp, double k0, double pi)
{
    if (o == s) return (po == ps) ? 1.0 - ld(znnpp, s) : 0.0;
    double rx = x[o] - x[s], ry = y[o] - y[s], rz = z[o] - z[s];
    double r = sqrt(rx * rx + ry * ry + rz * rz);
    zc gr = cexp(I * k0 * r) / (4.0 * pi * r * r);
    zc tau = I * k0 - 1.0 / r;
    zc f = kappa[s] * ld(che, s);
    double rv[3] = {rx, ry, rz};
    zc F;
    if (po == ps) {
        double rp = rv[po];
        F = gr * (tau + k0 * k0 * (r - rp * rp / r) - 3.0 * rp * rp / (r * r) * tau);
    } else {
        F = gr / r * (rv[po] * rv[ps]) * (-k0 * k0 - 3.0 * tau / r);
    }
    return -F * f;
}

/* Full FP64 MoM matrix, Z(3n x 3n) column-major (Compute_EFields_MoM.f90:138-145). */
void or_mom_fill(int n, const double *x, const double *y, const double *z, const double *kappa,
                 const double *che, const double *znnpp, double k0, int quirk, double *Z)
{
    const double pi = or_pi(quirk);
    const long N3 = 3L * n;
#pragma omp parallel for schedule(dynamic, 4)
    for (int s = 0; s < n; ++s)
        for (int ps = 0; ps < 3; ++ps)
            for (int o = 0; o < n; ++o)
                for (int po = 0; po < 3; ++po)
                    st(Z, (3L * s + ps) * N3 + 3L * o + po,
                       mom_elem(o, s, po, ps, x, y, z, kappa, che, znnpp, k0, pi));
}

/* FP64 block fill between receiver list B1 and source list B2 (same element
 * as above, diagonal decided by cell identifiers as Green_s_tr_partial does,
 * Green_functions.f90:125-127).  Output (3*n1 x 3*n2) column-major. */
void or_block_fill_f64(int n1, const int *id1, const double *x1, const double *y1, const double *z1,
                       int n2, const int *id2, const double *x2, const double *y2, const double *z2,
                       const double *kappa2, const double *che2, const double *znnpp2, double k0,
                       int quirk, double *Z)
{
    const double pi = or_pi(quirk);
    const long ld3 = 3L * n1;
#pragma omp parallel for schedule(dynamic, 4)
    for (int s = 0; s < n2; ++s)
        for (int o = 0; o < n1; ++o) {
            if (id1[o] == id2[s]) {
                for (int ps = 0; ps < 3; ++ps)
                    for (int po = 0; po < 3; ++po)
                        st(Z, (3L * s + ps) * ld3 + 3L * o + po, po == ps ? 1.0 - ld(znnpp2, s) : 0.0);
                continue;
            }
            double rx = x1[o] - x2[s], ry = y1[o] - y2[s], rz = z1[o] - z2[s];
            double r = sqrt(rx * rx + ry * ry + rz * rz);
            zc gr = cexp(I * k0 * r) / (4.0 * pi * r * r);
            zc tau = I * k0 - 1.0 / r;
            zc f = kappa2[s] * ld(che2, s);
            double rv[3] = {rx, ry, rz};
            for (int ps = 0; ps < 3; ++ps)
                for (int po = 0; po < 3; ++po) {
                    zc F;
                    if (po == ps)
                        F = gr * (tau + k0 * k0 * (r - rv[po] * rv[po] / r) - 3.0 * rv[po] * rv[po] / (r * r) * tau);
                    else
                        F = gr / r * (rv[po] * rv[ps]) * (-k0 * k0 - 3.0 * tau / r);
                    st(Z, (3L * s + ps) * ld3 + 3L * o + po, -F * f);
                }
        }
}

/* ---------------------------------------------------------------------------
 * Single-precision intermediates of Green_functions.f90 (quirk Q3).
 * Gr_mn, Tau_mn, f_kapChe and F** are default `Complex`; the mixed
 * expressions promote to double, the stores round to float, and the final
 * -F*f_kapChe is a complex(4) product.
 * ------------------------------------------------------------------------- */
static inline fc fmulc(fc a, fc b)
{
    float ar = crealf(a), ai = cimagf(a), br = crealf(b), bi = cimagf(b);
    volatile float t1 = ar * br, t2 = ai * bi, t3 = ar * bi, t4 = ai * br;
    float re = t1 - t2, im = t3 + t4;
    return re + I * im;
}
static inline fc tofc(zc v) { return (float)creal(v) + I * (float)cimag(v); }
static inline zc tozc(fc v) { return (double)crealf(v) + I * (double)cimagf(v); }

/* The single-precision dyad F (Fxx .. Fzz) and f_kapChe of one cell pair,
 * Green_functions.f90:139-160 (the same expressions appear in the _FN and SR
 * routines). */
static void dyad_F32(double rx, double ry, double rz, double k0, double pi, double kappa, zc che, fc F[3][3],
                     fc *fout)
{
    double r = sqrt(rx * rx + ry * ry + rz * rz);
    fc gr = tofc(cexp(I * k0 * r) / (4.0 * pi * r * r));
    fc tau = tofc(I * k0 - 1.0 / r);
    *fout = tofc(kappa * che);
    double rv[3] = {rx, ry, rz};
    for (int p = 0; p < 3; ++p) {
        double rp = rv[p];
        F[p][p] = tofc(tozc(gr) * (tozc(tau) + k0 * k0 * (r - rp * rp / r) - 3.0 * rp * rp / (r * r) * tozc(tau)));
    }
    /* 3.*Tau_mn is a complex(4) x real(4) product before the division by r. */
    fc tau3 = (3.0f * crealf(tau)) + I * (3.0f * cimagf(tau));
    for (int p = 0; p < 3; ++p)
        for (int q = p + 1; q < 3; ++q) {
            zc v = tozc(gr) / r * (rv[p] * rv[q]) * (-k0 * k0 - tozc(tau3) / r);
            F[p][q] = F[q][p] = tofc(v);
        }
}

/* The nine dyad entries for one cell pair, Green_functions.f90:139-167.
 * out[3*ps+po] is the entry (row component po, column component ps). */
static void dyad_f32(double rx, double ry, double rz, double k0, double pi, double kappa, zc che, zc out[9])
{
    fc F[3][3], f;
    dyad_F32(rx, ry, rz, k0, pi, kappa, che, F, &f);
    for (int ps = 0; ps < 3; ++ps)
        for (int po = 0; po < 3; ++po) out[3 * ps + po] = tozc(-fmulc(F[po][ps], f));
}

/* Green_s_tr_partial (Green_functions.f90:96-173): block fill with FP32
 * intermediates, stored to complex128.  Output (3*n1 x 3*n2). */
void or_block_fill_f32(int n1, const int *id1, const double *x1, const double *y1, const double *z1,
                       int n2, const int *id2, const double *x2, const double *y2, const double *z2,
                       const double *kappa2, const double *che2, const double *znnpp2, double k0,
                       int quirk, double *Z)
{
    const double pi = or_pi(quirk);
    const long ld3 = 3L * n1;
#pragma omp parallel for schedule(dynamic, 4)
    for (int s = 0; s < n2; ++s)
        for (int o = 0; o < n1; ++o) {
            zc e[9];
            if (id1[o] == id2[s]) {
                for (int q = 0; q < 9; ++q) e[q] = 0.0;
                e[0] = e[4] = e[8] = 1.0 - ld(znnpp2, s);
            } else {
                dyad_f32(x1[o] - x2[s], y1[o] - y2[s], z1[o] - z2[s], k0, pi, kappa2[s], ld(che2, s), e);
            }
            for (int ps = 0; ps < 3; ++ps)
                for (int po = 0; po < 3; ++po) st(Z, (3L * s + ps) * ld3 + 3L * o + po, e[3 * ps + po]);
        }
}

/* computeBlockRow (Green_functions.f90:1119-1225): global row grow (0-based),
 * columns of cells c0 .. c0+ncols/3-1.  FP32 intermediates; no diagonal case
 * (ACA is only used between distinct blocks). */
void or_block_row_f32(int grow, int c0, int ncols, const double *x, const double *y, const double *z,
                      const double *kappa, const double *che, double k0, int quirk, double *row)
{
    const double pi = or_pi(quirk);
    int o = grow / 3, po = grow % 3;
    for (int j = 0; j < ncols / 3; ++j) {
        int s = c0 + j;
        zc e[9];
        dyad_f32(x[o] - x[s], y[o] - y[s], z[o] - z[s], k0, pi, kappa[s], ld(che, s), e);
        for (int ps = 0; ps < 3; ++ps) st(row, 3L * j + ps, e[3 * ps + po]);
    }
}

/* computeBlockCol (Green_functions.f90:905-1012). */
void or_block_col_f32(int gcol, int r0, int nrows, const double *x, const double *y, const double *z,
                      const double *kappa, const double *che, double k0, int quirk, double *col)
{
    const double pi = or_pi(quirk);
    int s = gcol / 3, ps = gcol % 3;
    for (int i = 0; i < nrows / 3; ++i) {
        int o = r0 + i;
        zc e[9];
        dyad_f32(x[o] - x[s], y[o] - y[s], z[o] - z[s], k0, pi, kappa[s], ld(che, s), e);
        for (int po = 0; po < 3; ++po) st(col, 3L * i + po, e[3 * ps + po]);
    }
}

/* Same generators in FP64 (the MoM element formula). */
void or_block_row_f64(int grow, int c0, int ncols, const double *x, const double *y, const double *z,
                      const double *kappa, const double *che, const double *znnpp, double k0, int quirk,
                      double *row)
{
    const double pi = or_pi(quirk);
    int o = grow / 3, po = grow % 3;
    for (int j = 0; j < ncols; ++j)
        st(row, j, mom_elem(o, c0 + j / 3, po, j % 3, x, y, z, kappa, che, znnpp, k0, pi));
}
void or_block_col_f64(int gcol, int r0, int nrows, const double *x, const double *y, const double *z,
                      const double *kappa, const double *che, const double *znnpp, double k0, int quirk,
                      double *col)
{
    const double pi = or_pi(quirk);
    int s = gcol / 3, ps = gcol % 3;
    for (int i = 0; i < nrows; ++i)
        st(col, i, mom_elem(r0 + i / 3, s, i % 3, ps, x, y, z, kappa, che, znnpp, k0, pi));
}

/* Incident_Field_single_element (Compute_EFields_MoM.f90:529-598), FP64:
 * V(3n x 2*ntr), columns 0..ntr-1 V-pol, ntr..2ntr-1 H-pol. */
void or_mom_rhs(int n, const double *x, const double *y, const double *z, int ntr, const double *th_deg,
                const double *ph_deg, double k0, int quirk, double *V)
{
    const double pi = or_pi(quirk);
    const long ld3 = 3L * n;
    for (int t = 0; t < ntr; ++t) {
        double th = th_deg[t] * pi / 180.0, ph = ph_deg[t] * pi / 180.0;
        double kx = k0 * cos(th), ky = k0 * sin(th) * cos(ph), kz = k0 * sin(th) * sin(ph);
        for (int c = 0; c < n; ++c) {
            zc e = cexp(I * kx * x[c]) * cexp(I * ky * y[c]) * cexp(I * kz * z[c]);
            st(V, (long)t * ld3 + 3L * c + 0, -e * sin(th));
            st(V, (long)t * ld3 + 3L * c + 1, e * cos(th) * cos(ph));
            st(V, (long)t * ld3 + 3L * c + 2, e * cos(th) * sin(ph));
            st(V, (long)(t + ntr) * ld3 + 3L * c + 0, 0.0);
            st(V, (long)(t + ntr) * ld3 + 3L * c + 1, -e * sin(ph));
            st(V, (long)(t + ntr) * ld3 + 3L * c + 2, e * cos(ph));
        }
    }
}

/* Incident_Field (Incident_Field.f90:1-63): K11x..K11z and Ex..Ez are default
 * Complex (quirk Q4).  Output columns: t-t0 V-pol, (t-t0)+nt H-pol. */
void or_incident_field_f32(int n, const double *x, const double *y, const double *z, int t0, int t1,
                           const double *th_deg, const double *ph_deg, double k0, int quirk, double *E)
{
    const double pi = or_pi(quirk);
    const long ld3 = 3L * n;
    const int nt = t1 - t0;
    for (int t = t0; t < t1; ++t) {
        double th = th_deg[t] * pi / 180.0, ph = ph_deg[t] * pi / 180.0;
        float kx = (float)(k0 * cos(th)), ky = (float)(k0 * sin(th) * cos(ph)), kz = (float)(k0 * sin(th) * sin(ph));
        long cv = (long)(t - t0) * ld3, ch = (long)(t - t0 + nt) * ld3;
        for (int c = 0; c < n; ++c) {
            zc e = cexp(I * (double)kx * x[c]) * cexp(I * (double)ky * y[c]) * cexp(I * (double)kz * z[c]);
            st(E, cv + 3L * c + 0, tozc(tofc(-e * sin(th))));
            st(E, cv + 3L * c + 1, tozc(tofc(e * cos(th) * cos(ph))));
            st(E, cv + 3L * c + 2, tozc(tofc(e * cos(th) * sin(ph))));
            st(E, ch + 3L * c + 0, 0.0);
            st(E, ch + 3L * c + 1, tozc(tofc(-e * sin(ph))));
            st(E, ch + 3L * c + 2, tozc(tofc(e * cos(ph))));
        }
    }
}

/* GetFFieldCoeff (Green_functions.f90:1399-1438). */
void or_ff_coeff(int n, const double *x, const double *y, const double *z, const double *kappa,
                 const double *che, double th_deg, double ph_deg, double k0, int quirk, double *ff)
{
    const double pi = or_pi(quirk);
    double ux = cos(th_deg * pi / 180.0);
    double uy = sin(th_deg * pi / 180.0) * cos(ph_deg * pi / 180.0);
    double uz = sin(th_deg * pi / 180.0) * sin(ph_deg * pi / 180.0);
    for (int c = 0; c < n; ++c) {
        zc ffc = cexp(-I * k0 * (ux * x[c] + uy * y[c] + uz * z[c]));
        zc f = kappa[c] * ld(che, c);
        st(ff, c, ffc * k0 * k0 * f / (4.0 * pi) * k0);
    }
}

/* Far-field contraction and projection (Compute_Scattering_Matrices.f90:122-207;
 * with beta = 0 it is Compute_EFields_MoM.f90:306-369).  E is (3n x 2*ntr).
 * S is (nrx x 4*ntr*npol) column-major: column 4*npol*t + 4*ipol + {Vv,Vh,Hv,Hh}. */
void or_scattering_matrix(int n, const double *x, const double *y, const double *z, const double *kappa,
                          const double *che, const double *E, int ntr, int nrx, const double *rth,
                          const double *rph, int npol, const double *beta_deg, double k0, int quirk,
                          double *S)
{
    const double pi = or_pi(quirk);
    const long ld3 = 3L * n;
#pragma omp parallel for schedule(dynamic, 1)
    for (int r = 0; r < nrx; ++r) {
        double *ff = (double *)malloc(sizeof(double) * 2 * (size_t)n);
        or_ff_coeff(n, x, y, z, kappa, che, rth[r], rph[r], k0, quirk, ff);
        double th = rth[r] * pi / 180.0, ph = rph[r] * pi / 180.0;
        double st_ = sin(th), ct = cos(th), sp = sin(ph), cp = cos(ph);
        for (int t = 0; t < ntr; ++t) {
            zc ev[3] = {0, 0, 0}, eh[3] = {0, 0, 0};
            for (int c = 0; c < n; ++c) {
                zc f = ld(ff, c);
                for (int p = 0; p < 3; ++p) {
                    ev[p] += f * ld(E, (long)t * ld3 + 3L * c + p);
                    eh[p] += f * ld(E, (long)(t + ntr) * ld3 + 3L * c + p);
                }
            }
            for (int ip = 0; ip < npol; ++ip) {
                double b = beta_deg[ip] * pi / 180.0, cb = cos(b), sb = sin(b);
                zc evp[3], ehp[3];
                for (int p = 0; p < 3; ++p) {
                    evp[p] = cb * ev[p] + sb * eh[p];
                    ehp[p] = -sb * ev[p] + cb * eh[p];
                }
                zc vv = -evp[0] * cb * st_ + evp[1] * (cb * ct * cp - sb * sp) + evp[2] * (cb * ct * sp + sb * cp);
                zc vh = evp[0] * sb * st_ - evp[1] * (sb * ct * cp + cb * sp) + evp[2] * (cb * cp - sb * ct * sp);
                zc hv = -ehp[0] * cb * st_ + ehp[1] * (cb * ct * cp - sb * sp) + ehp[2] * (cb * ct * sp + sb * cp);
                zc hh = ehp[0] * sb * st_ - ehp[1] * (sb * ct * cp + cb * sp) + ehp[2] * (cb * cp - sb * ct * sp);
                long col = 4L * npol * t + 4L * ip;
                st(S, (col + 0) * nrx + r, vv);
                st(S, (col + 1) * nrx + r, vh);
                st(S, (col + 2) * nrx + r, hv);
                st(S, (col + 3) * nrx + r, hh);
            }
        }
        free(ff);
    }
}

/* Compute_ExtAbsCsec_fromIntField (Compute_ExtAbsCsec_fromIntField.f90:45-64).
 * quirk_q6 != 0 reproduces the reference's sum-of-components formula (Q6);
 * quirk_q6 == 0 uses the dot products E.E* and chi E.Einc*.  Einc comes from
 * or_incident_field_f32 as in the reference. */
void or_cext_cabs(int n, const double *x, const double *y, const double *z, const double *sc,
                  const double *che, const double *E, int ntr, const double *th_deg, const double *ph_deg,
                  double k0, int quirk, int quirk_q6, double *cext, double *cabs_out)
{
    const long ld3 = 3L * n;
    double *einc = (double *)malloc(sizeof(double) * 2 * (size_t)ld3 * 2);
    for (int t = 0; t < ntr; ++t) {
        or_incident_field_f32(n, x, y, z, t, t + 1, th_deg, ph_deg, k0, quirk, einc);
        double ext[2] = {0, 0}, abs_[2] = {0, 0};
        for (int pol = 0; pol < 2; ++pol) {
            long cE = (long)(t + pol * ntr) * ld3, cI = (long)pol * ld3;
            for (int c = 0; c < n; ++c) {
                zc chi = ld(che, c);
                double v = sc[c] * sc[c] * sc[c];
                if (quirk_q6) {
                    zc se = ld(E, cE + 3L * c) + ld(E, cE + 3L * c + 1) + ld(E, cE + 3L * c + 2);
                    zc si = ld(einc, cI + 3L * c) + ld(einc, cI + 3L * c + 1) + ld(einc, cI + 3L * c + 2);
                    double a = cabs(se);
                    abs_[pol] += cimag(chi) * a * a * v;
                    ext[pol] += cimag(chi * se * conj(si)) * v;
                } else {
                    double ee = 0;
                    zc ei = 0;
                    for (int p = 0; p < 3; ++p) {
                        zc e = ld(E, cE + 3L * c + p);
                        ee += creal(e) * creal(e) + cimag(e) * cimag(e);
                        ei += e * conj(ld(einc, cI + 3L * c + p));
                    }
                    abs_[pol] += cimag(chi) * ee * v;
                    ext[pol] += cimag(chi * ei) * v;
                }
            }
        }
        cext[2 * t] = k0 * (ext[0] + ext[1]) / 2.0;
        cext[2 * t + 1] = 0.0;
        cabs_out[2 * t] = k0 * (abs_[0] + abs_[1]) / 2.0;
        cabs_out[2 * t + 1] = 0.0;
    }
    free(einc);
}

/* ---------------------------------------------------------------------------
 * Partially pivoted ACA, Calcul_MatrixZij_ACA (Calcul_ACA_matrices.f90:6-219).
 * The block is rows of cells r0.. (nrow = 3*cells), columns of cells c0..
 * (ncol = 3*cells), generated on the fly.  prec=0: FP32-intermediate rows and
 * columns as the reference computes them (Q3); prec=1: FP64 element.
 * U is (nrow x kmax) and V is (kmax x ncol), column-major.  Returns k as the
 * reference does (kmax+1 if the tolerance is never met).  Pivots follow
 * maxloc (first index on ties); the first row is always row 1 (Q8).
 * ------------------------------------------------------------------------- */
typedef struct {
    const double *x, *y, *z, *kappa, *che, *znnpp;
    double k0;
    int quirk, prec;
} aca_cells;

static void aca_row(const aca_cells *C, int grow, int c0, int ncol, zc *out)
{
    double *tmp = (double *)malloc(sizeof(double) * 2 * (size_t)ncol);
    if (C->prec == 0) or_block_row_f32(grow, c0, ncol, C->x, C->y, C->z, C->kappa, C->che, C->k0, C->quirk, tmp);
    else or_block_row_f64(grow, c0, ncol, C->x, C->y, C->z, C->kappa, C->che, C->znnpp, C->k0, C->quirk, tmp);
    for (int j = 0; j < ncol; ++j) out[j] = ld(tmp, j);
    free(tmp);
}
static void aca_col(const aca_cells *C, int gcol, int r0, int nrow, zc *out)
{
    double *tmp = (double *)malloc(sizeof(double) * 2 * (size_t)nrow);
    if (C->prec == 0) or_block_col_f32(gcol, r0, nrow, C->x, C->y, C->z, C->kappa, C->che, C->k0, C->quirk, tmp);
    else or_block_col_f64(gcol, r0, nrow, C->x, C->y, C->z, C->kappa, C->che, C->znnpp, C->k0, C->quirk, tmp);
    for (int i = 0; i < nrow; ++i) out[i] = ld(tmp, i);
    free(tmp);
}

/* Generic ACA core: getrow/getcol return the exact matrix row / column. */
typedef void (*rowfn)(void *, int i, zc *out);
typedef void (*colfn)(void *, int j, zc *out);

static int aca_core(int m, int n, int kmax, double eps, void *ctx, rowfn getrow, colfn getcol, zc *U, zc *V,
                    int *rows_out, int *cols_out)
{
    int *Ir = (int *)calloc(kmax + 2, sizeof(int)), *Jc = (int *)calloc(kmax + 2, sizeof(int));
    double *nf = (double *)calloc(kmax + 2, sizeof(double));
    zc *row = (zc *)malloc(sizeof(zc) * n), *col = (zc *)malloc(sizeof(zc) * m);
    for (long q = 0; q < (long)m * kmax; ++q) U[q] = 0;
    for (long q = 0; q < (long)n * kmax; ++q) V[q] = 0;
#define Uel(i, k) U[(long)(k) * m + (i)]
#define Vel(k, j) V[(long)(j) * kmax + (k)]
    /* k = 1 */
    Ir[0] = 0;
    getrow(ctx, 0, row);
    int jb = 0;
    double best = -1.0;
    for (int j = 0; j < n; ++j) if (cabs(row[j]) > best) { best = cabs(row[j]); jb = j; }
    Jc[0] = jb;
    zc piv = row[jb];
    for (int j = 0; j < n; ++j) Vel(0, j) = row[j] / piv;
    getcol(ctx, jb, col);
    for (int i = 0; i < m; ++i) Uel(i, 0) = col[i];
    double v2 = 0;
    for (int i = 0; i < m; ++i) v2 += cabs(Uel(i, 0)) * cabs(Uel(i, 0));
    double nu = sqrt(v2);
    v2 = 0;
    for (int j = 0; j < n; ++j) v2 += cabs(Vel(0, j)) * cabs(Vel(0, j));
    double nv = sqrt(v2);
    nf[0] = sqrt(0.0 + nu * nu * nv * nv);
    {
        int ib = 0;
        best = -1.0;
        for (int i = 0; i < m; ++i) {
            double a = (i == Ir[0]) ? 0.0 : cabs(Uel(i, 0));
            if (a > best) { best = a; ib = i; }
        }
        Ir[1] = ib;
    }
    int k = 2; /* 1-based iteration counter as in the reference */
    while (k <= kmax) {
        int ik = Ir[k - 1];
        getrow(ctx, ik, row);
        for (int j = 0; j < n; ++j) {
            zc acc = 0;
            for (int l = 0; l < k - 1; ++l) acc += Uel(ik, l) * Vel(l, j);
            row[j] -= acc;
        }
        int jk = 0;
        best = -1.0;
        for (int j = 0; j < n; ++j) {
            int used = 0;
            for (int l = 0; l < k - 1; ++l) if (Jc[l] == j) used = 1;
            double a = used ? 0.0 : cabs(row[j]);
            if (a > best) { best = a; jk = j; }
        }
        Jc[k - 1] = jk;
        piv = row[jk];
        for (int j = 0; j < n; ++j) Vel(k - 1, j) = row[j] / piv;
        getcol(ctx, jk, col);
        for (int i = 0; i < m; ++i) {
            zc acc = 0;
            for (int l = 0; l < k - 1; ++l) acc += Vel(l, jk) * Uel(i, l);
            Uel(i, k - 1) = col[i] - acc;
        }
        double prev = nf[k - 2];
        v2 = 0;
        for (int i = 0; i < m; ++i) v2 += cabs(Uel(i, k - 1)) * cabs(Uel(i, k - 1));
        nu = sqrt(v2);
        v2 = 0;
        for (int j = 0; j < n; ++j) v2 += cabs(Vel(k - 1, j)) * cabs(Vel(k - 1, j));
        nv = sqrt(v2);
        double sum = 0;
        for (int l = 0; l < k - 1; ++l) {
            zc du = 0, dv = 0; /* non-conjugated inner products (Q8) */
            for (int i = 0; i < m; ++i) du += Uel(i, l) * Uel(i, k - 1);
            for (int j = 0; j < n; ++j) dv += Vel(k - 1, j) * Vel(l, j);
            sum += cabs(du) * cabs(dv);
        }
        nf[k - 1] = sqrt(prev * prev + 2.0 * sum + nu * nu * nv * nv);
        double res = (nu * nv) / nf[k - 1];
        if (res <= eps) break;
        int ib = 0;
        best = -1.0;
        for (int i = 0; i < m; ++i) {
            int used = 0;
            for (int l = 0; l < k; ++l) if (Ir[l] == i) used = 1;
            double a = used ? 0.0 : cabs(Uel(i, k - 1));
            if (a > best) { best = a; ib = i; }
        }
        Ir[k] = ib;
        k = k + 1;
    }
#undef Uel
#undef Vel
    int kk = k < kmax ? k : kmax;
    if (rows_out) for (int l = 0; l < kk; ++l) rows_out[l] = Ir[l];
    if (cols_out) for (int l = 0; l < kk; ++l) cols_out[l] = Jc[l];
    free(Ir); free(Jc); free(nf); free(row); free(col);
    return k;
}

typedef struct { const aca_cells *C; int r0, c0, m, n; } aca_blk;
static void blk_row(void *p, int i, zc *out) { aca_blk *b = (aca_blk *)p; aca_row(b->C, 3 * b->r0 + i, b->c0, b->n, out); }
static void blk_col(void *p, int j, zc *out) { aca_blk *b = (aca_blk *)p; aca_col(b->C, 3 * b->c0 + j, b->r0, b->m, out); }

int or_aca(int r0, int ncell_r, int c0, int ncell_c, const double *x, const double *y, const double *z,
           const double *kappa, const double *che, const double *znnpp, double k0, int quirk, int prec,
           int kmax, double eps, double *U, double *V, int *rows, int *cols)
{
    aca_cells C = {x, y, z, kappa, che, znnpp, k0, quirk, prec};
    aca_blk b = {&C, r0, c0, 3 * ncell_r, 3 * ncell_c};
    return aca_core(b.m, b.n, kmax, eps, &b, blk_row, blk_col, (zc *)U, (zc *)V, rows, cols);
}

/* Calcul_GenMatrix_ACA (Calcul_ACA_matrices.f90:222-424): ACA of an explicit
 * m x n matrix (column-major). */
typedef struct { const zc *A; int m, n; } gen_blk;
static void gen_row(void *p, int i, zc *out) { gen_blk *g = (gen_blk *)p; for (int j = 0; j < g->n; ++j) out[j] = g->A[(long)j * g->m + i]; }
static void gen_col(void *p, int j, zc *out) { gen_blk *g = (gen_blk *)p; memcpy(out, g->A + (long)j * g->m, sizeof(zc) * g->m); }

int or_aca_gen(int m, int n, const double *A, int kmax, double eps, double *U, double *V)
{
    gen_blk g = {(const zc *)A, m, n};
    return aca_core(m, n, kmax, eps, &g, gen_row, gen_col, (zc *)U, (zc *)V, NULL, NULL);
}


/* ---------------------------------------------------------------------------
 * SR (sparsified) block of the CBF solves.
 * ------------------------------------------------------------------------- */

/* Green_s_tr_partial_FN (Green_functions.f90:176-239, homogs = 0): Frobenius
 * norm of the block over cells 0..n-1 (receivers = sources).  Each pair's term
 * (|Fxx|^2 + 2|Fxy|^2 + 2|Fxz|^2 + |Fyy|^2 + 2|Fyz|^2 + |Fzz|^2) |f|^2 is
 * single precision, accumulated in double; a self pair adds 3 |1 - Znnpp|^2. */
double or_block_fn(int n, const int *id, const double *x, const double *y, const double *z, const double *kappa,
                   const double *che, const double *znnpp, double k0, int quirk)
{
    const double pi = or_pi(quirk);
    double val = 0.0;
    for (int s = 0; s < n; ++s)
        for (int o = 0; o < n; ++o) {
            if (id[o] == id[s]) {
                double a = cabs(1.0 - ld(znnpp, s));
                val += 3.0 * (a * a);
            } else {
                fc F[3][3], f;
                dyad_F32(x[o] - x[s], y[o] - y[s], z[o] - z[s], k0, pi, kappa[s], ld(che, s), F, &f);
                volatile float axx = cabsf(F[0][0]), axy = cabsf(F[0][1]), axz = cabsf(F[0][2]);
                volatile float ayy = cabsf(F[1][1]), ayz = cabsf(F[1][2]), azz = cabsf(F[2][2]), af = cabsf(f);
                volatile float t = axx * axx;
                t = t + 2.0f * (axy * axy);
                t = t + 2.0f * (axz * axz);
                t = t + ayy * ayy;
                t = t + 2.0f * (ayz * ayz);
                t = t + azz * azz;
                t = t * (af * af);
                val += (double)t;
            }
        }
    return sqrt(val);
}

/* The SR mask of SR_Green_s_tr_partial(_FN) (Green_functions.f90:290-561,
 * 564-806, homogs = 0): threshold = |1 - Znnpp(first cell)| / fsr; an
 * off-cell entry g (single precision) is kept when |g| > threshold; the
 * diagonal of a self pair is always kept, its cross-component zeros never.
 * Returns nnz; *fn = sqrt(sum of |g|^2 over kept entries) (single-precision
 * squares accumulated in double, rows in order); when Zout is given the kept
 * entries are written into the dense (3n x 3n) block, zeros elsewhere. */
long or_sr_block(int n, const double *x, const double *y, const double *z, const double *kappa, const double *che,
                 const double *znnpp, double k0, int quirk, double fsr, double *fn, double *Zout)
{
    const double pi = or_pi(quirk);
    const double thr = cabs(1.0 - ld(znnpp, 0)) / fsr;
    const long ld3 = 3L * n;
    if (Zout) memset(Zout, 0, sizeof(double) * 2 * ld3 * ld3);
    long nnz = 0;
    double g2 = 0.0;
    for (int o = 0; o < n; ++o)
        for (int po = 0; po < 3; ++po) {
            const long row = 3L * o + po;
            for (int s = 0; s < n; ++s) {
                if (s == o) {
                    zc d = 1.0 - ld(znnpp, o);
                    double a = cabs(d);
                    ++nnz;
                    g2 += a * a;
                    if (Zout) st(Zout, row + row * ld3, d);
                    continue;
                }
                zc e[9];
                dyad_f32(x[o] - x[s], y[o] - y[s], z[o] - z[s], k0, pi, kappa[s], ld(che, s), e);
                for (int ps = 0; ps < 3; ++ps) {
                    zc g = e[3 * ps + po];
                    volatile float a = cabsf((float)creal(g) + I * (float)cimag(g));
                    if ((double)a > thr) {
                        ++nnz;
                        volatile float a2 = a * a;
                        g2 += (double)a2;
                        if (Zout) st(Zout, row + (3L * s + ps) * ld3, g);
                    }
                }
            }
        }
    if (fn) *fn = sqrt(g2);
    return nnz;
}
