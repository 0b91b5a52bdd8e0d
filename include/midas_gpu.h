/*
 * midas_gpu.h -- C-ABI of the MI355X (gfx950) implementation of the MIDAS
 * volume-integral-equation MoM hot path: dyadic Green's-function fill, dense
 * complex solve (FP64 LU, or mixed-precision LU + FP64 iterative refinement),
 * far-field scattering matrix, Cext/Cabs, ACA and the CBFM compressed solver.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, and
 * returns 0 on success, > 0 for a LAPACK-style singular pivot (1-based index,
 * as PZGESV's INFO), < 0 for a bad argument or a HIP error; midas_last_error()
 * gives the message.  Host-pointer entry points are synchronous: results are in
 * host memory on return.  midas_dev_* entry points take device pointers and are
 * ordered on the context's stream (midas_stream()).
 *
 * Conventions follow the reference (Fortran, column-major):
 *   - unknown i = 3*cell + p (0-based here), p in {x, y, z};
 *   - complex numbers are interleaved double (re, im) = Fortran COMPLEX(real64);
 *   - matrices are column-major with an explicit leading dimension;
 *   - angles are degrees, theta measured from +x (Incident_Field.f90:31-34);
 *   - excitation columns 0..ntr-1 are V-pol, ntr..2ntr-1 H-pol
 *     (Compute_EFields_MoM.f90:557-563).
 */
#ifndef MIDAS_GPU_H
#define MIDAS_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct midas_ctx midas_ctx;

/* Quirk bits: reproduce the reference's precision quirks (default: all on). */
#define MIDAS_Q_PI 1          /* Pi = Acos(-1.) of a single literal (Initialization.f90:39)          */
#define MIDAS_Q_LITERALS 2    /* single (1./3.), (2./3.) literals (SetCellsParams.f90:37-44)         */
#define MIDAS_Q_CEXT_SUM 4    /* component-sum Cext/Cabs (Compute_ExtAbsCsec_fromIntField.f90:53-60)  */
#define MIDAS_Q_FP32_BLOCKS 8 /* FP32 intermediates of Green_s_tr_partial / Incident_Field (CBFM)    */
#define MIDAS_Q_ALL 15

/* Solver modes for the dense complex systems. */
#define MIDAS_SOLVE_FP64 0        /* FP64 LU (PZGESV semantics): the plain baseline                 */
#define MIDAS_SOLVE_C32_IR 1      /* complex-FP32 LU on FP32 MFMA + FP64 iterative refinement       */
#define MIDAS_SOLVE_F16_IR 2      /* FP32 LU, trailing GEMM on FP16 MFMA (scaled, 1 term) + FP64 IR */
#define MIDAS_SOLVE_F16X3_IR 3    /* ... FP16 split hi/lo, 3 product terms + FP64 IR                */
#define MIDAS_SOLVE_BF16_IR 4     /* ... BF16, 1 term + FP64 IR                                     */
#define MIDAS_SOLVE_BF16X3_IR 5   /* ... BF16 split hi/lo, 3 product terms + FP64 IR                */

/* Cell table in structure-of-arrays form (the fields of the reference's
 * `type Cell`, Initialization.f90:22-26, that the hot path reads).  If kappa,
 * che and znnpp are NULL they are computed from sc and eps as SetCellsParams
 * does; if given, they are used as they are.  id (n_cell) may be NULL (1..n). */
typedef struct {
    int n;
    const double *x, *y, *z;  /* centres [m]                         */
    const double *sc;         /* cell edge Sc [m]                    */
    const double *eps;        /* relative permittivity, 2n doubles   */
    const double *kappa;      /* Kappa_n (optional)                  */
    const double *che;        /* Che_n = eps - 1, 2n (optional)      */
    const double *znnpp;      /* Znnpp, 2n (optional)                */
    const int *id;            /* n_cell identifiers (optional)       */
} midas_cells_t;

/* Mixed-precision solves refine X in FP64 until the normwise backward error
 * max_j ||b_j - Z x_j||_inf / (||Z||_inf ||x_j||_inf + ||b_j||_inf) is <= tol.
 * When it is not reached (the sweeps stop gaining, or max_iter is spent), the
 * solve falls back to the FP64 LU of Z, as LAPACK's ZCGESV does
 * (status MIDAS_STATUS_FP64_FALLBACK), unless no_fallback is set: then X holds
 * the iterate with the smallest backward error, status is
 * MIDAS_STATUS_NOT_CONVERGED and the call returns MIDAS_ERR_NOT_CONVERGED. */
typedef struct {
    int max_iter;      /* refinement sweeps (default 30)                                      */
    double tol;        /* backward-error target (default max(1e-14, sqrt(n) * DBL_EPSILON))  */
    int nb;            /* LU panel width (0: default)                                        */
    int no_fallback;   /* 1: report a refinement that misses tol instead of the FP64 fallback */
} midas_solve_opts_t;

#define MIDAS_STATUS_CONVERGED 0       /* FP64 LU, or refinement reached tol                   */
#define MIDAS_STATUS_NOT_CONVERGED 1   /* refinement missed tol; X = best iterate (no_fallback) */
#define MIDAS_STATUS_FP64_FALLBACK 2   /* refinement missed tol; X from the FP64 LU of Z        */
#define MIDAS_ERR_NOT_CONVERGED (-4)

typedef struct {
    int info;          /* 0, or first exactly-zero pivot (1-based)                            */
    int iterations;    /* refinement sweeps done (0 for MIDAS_SOLVE_FP64)                     */
    double berr;       /* backward error of the returned X (definition above)                 */
    double anorm;      /* ||Z||_inf (PZLANGE 'I', Compute_EFields_MoM.f90:168)                */
    double t_factor_ms, t_solve_ms, t_refine_ms;
    int status;        /* MIDAS_STATUS_*                                                      */
    int pivot_search;  /* mixed LU: 0 partial pivoting over each whole column; 1 pivot search
                        * restricted to each panel's diagonal block (schedule parameter
                        * "lu_pivot_search": 1 for the MoM solves, 2 for every solve)         */
    double max_l21;    /* pivot_search 1: largest |l| below the diagonal blocks; <= 1/sqrt(2)
                        * certifies that partial pivoting over the whole columns would have
                        * chosen the same pivots                                              */
} midas_solve_info_t;

/* ---------------- context ---------------- */
int midas_init(int device, midas_ctx **ctx);
int midas_finalize(midas_ctx *ctx);
/* The solvers keep their device workspace (the FP32 copy of the matrix, LU
 * work buffers, refinement vectors; the CBFM call's block matrices, CBFs,
 * Gram and reduced matrices, ACA factors) in the context and grow it on
 * demand, so that repeated calls of one size allocate nothing; this frees it.
 * No reference counterpart (ScaLAPACK workspaces are the caller's). */
int midas_release_workspace(midas_ctx *ctx);
/* Memory budgets (bytes) of the CBFM batching: the padded block matrices of one
 * batched block-LU chunk, the column strips of the dense reduced-matrix
 * assembly, the dense block pairs filled side by side, and the ACA factors of
 * one batch of block columns.  0 keeps the default (24e9, 8 GiB, 4 GiB, 4e9).
 * Smaller budgets give more, smaller batches with the same results.  No
 * reference counterpart (the reference distributes blocks over MPI ranks). */
int midas_set_cbfm_budgets(midas_ctx *ctx, double lu_bytes, double strip_bytes, double dense_bytes,
                           double aca_bytes);
const char *midas_last_error(void);
void *midas_stream(midas_ctx *ctx);           /* hipStream_t of the context */
/* Schedule parameters (process-wide; defaults chosen by the project's
 * autotuner, midas_amd/autotune.py): "lu_nb_split", "lu_two_panel_min_n",
 * "lu_nb_mixed", "lu_nb_fp64" (LU panel widths), "panel_leaf" (recursive panel
 * leaves), "trail_group_kib" (L2 share of a row-tile group in the split GEMM),
 * "refine_graph" (1: refinement sweeps replayed from a captured hipGraph),
 * "cbfm_graph" (1: each batch of CBF block systems -- grouped fills, SR masks,
 * excitations, batched FP64 LU and solve -- recorded into a hipGraph and
 * launched as one, the executable graph kept and updated in place by the next
 * call), "zgemm_tile", "cgemm_tile", "fp64_gemm_library", "cbfm_block_mode",
 * "lu_pivot_search" (mixed LU: 1 the MoM solves search each panel's pivots in
 * its diagonal block, certified by max |L21| in the report; 2 every solve;
 * 0 whole columns), "mom_matrix_free" (1: MoM mixed solves with at most two
 * directions recompute the FP64 matrix for their residuals instead of storing
 * it), "u12_split" (1: LUs with a single-term split trailing update form the
 * U12 block rows on the same split-precision matrix cores, in place as
 * A12 -= (I - L11^-1) A12; 0 on the complex-FP32 GEMM), "l21_split" (1: in
 * their two-panel steps the diagonal-block pivot search's L21 = A21 U11^-1 the
 * same way, A21 -= A21 (I - U11^-1)), "lu_lookahead" (1: LU panels factored
 * one step ahead on a stream of their own; 0: on the main stream, one stream
 * in all, for profilers that serialise dispatches).
 * They change speed, not results beyond rounding.  No reference counterpart
 * (the reference fixes its 32 x 32 ScaLAPACK blocks, Initialization.f90:66-67). */
int midas_tune_set(const char *name, int value);
int midas_tune_get(const char *name, int *value);

/* Live kernel accounting: midas_profile(ctx, 1) starts recording HIP events
 * around the trailing-update GEMM launches of the LU (kind 0) and the matrix
 * fill launches (kind 1) on the context's stream; midas_profile_query sums
 * their count, elapsed time, algorithmic flops and algorithmic bytes.  No
 * reference counterpart (the reference prints wall-clock phase timers,
 * Calcul_time_spent.f90:1-60). */
int midas_profile(midas_ctx *ctx, int enable);
int midas_profile_query(midas_ctx *ctx, int kind, int *launches, double *total_ms, double *flops, double *bytes);

/* ---------------- a1: cell constants ----------------
 * SetCellsParams (SetCellsParams.f90:37-48): a_n, Che_n, Znnpp, Kappa_n. */
int midas_cell_params(midas_ctx *ctx, int n, const double *sc, const double *eps, double k0, int quirks,
                      double *a_n, double *che, double *znnpp, double *kappa);

/* ---------------- a2: MoM matrix ----------------
 * Z (3n x 3n, ldz) = compute_Green_s_tr_single_element for every (row, column)
 * (Compute_EFields_MoM.f90:138-145, 433-523), FP64. */
int midas_mom_fill(midas_ctx *ctx, const midas_cells_t *cells, double k0, int quirks, double *Z, long ldz);

/* ---------------- a7: block matrix ----------------
 * Green_s_tr_partial (Green_functions.f90:96-173): Z (3*b1.n x 3*b2.n) between
 * receiver cells b1 and source cells b2; diagonal blocks where identifiers
 * coincide.  quirks & MIDAS_Q_FP32_BLOCKS reproduces the FP32 intermediates. */
int midas_block_fill(midas_ctx *ctx, const midas_cells_t *b1, const midas_cells_t *b2, double k0, int quirks,
                     double *Z, long ldz);

/* ---------------- a3 / a10: plane-wave excitations ----------------
 * V (3n x 2ntr, ldv).  ref32 = 0: Incident_Field_single_element
 * (Compute_EFields_MoM.f90:529-598, FP64); ref32 = 1: Incident_Field
 * (Incident_Field.f90:1-63, FP32 locals). */
int midas_plane_wave(midas_ctx *ctx, const midas_cells_t *cells, double k0, int quirks, int ntr, const double *th,
                     const double *ph, int ref32, double *V, long ldv);

/* ---------------- a10: spherical-wave CBF excitations ----------------
 * Incident_Field_Spherical (Incident_Field.f90:65-166) over the n cells given
 * (an extended block, block cells first): sources at 10 max(R_block, lambda)
 * from the cells' centroid in the directions (th, ph); V (3n x 2ntr, ldv).
 * ref32 = 1 rounds the phase and the components to complex(4) as the
 * reference does. */
int midas_spherical_wave(midas_ctx *ctx, const midas_cells_t *cells, double k0, int quirks, int ntr, const double *th,
                         const double *ph, int ref32, double *V, long ldv);

/* ---------------- a4: dense solve ----------------
 * PZGESV (Compute_EFields_MoM.f90:173): X = Z^-1 B, overwriting B (n x nrhs).
 * Z is not modified. */
int midas_solve(midas_ctx *ctx, int n, int nrhs, const double *Z, long ldz, double *B, long ldb, int mode,
                const midas_solve_opts_t *opts, midas_solve_info_t *info);

/* ---------------- a5: far-field scattering matrix ----------------
 * GetFFieldCoeff + contraction + projection (Green_functions.f90:1399-1438,
 * Compute_EFields_MoM.f90:306-369, Compute_Scattering_Matrices.f90:122-207).
 * E (3n x 2ntr, lde); S (nrx x 4*ntr*npol, lds), column 4*npol*t + 4*ip +
 * {Vv, Vh, Hv, Hh}; npol = 1 with beta = 0 is the MoM S_total layout. */
int midas_scattering_matrix(midas_ctx *ctx, const midas_cells_t *cells, double k0, int quirks, const double *E,
                            long lde, int ntr, int nrx, const double *rth, const double *rph, int npol,
                            const double *beta, double *S, long lds);

/* ---------------- a6: extinction / absorption ----------------
 * Compute_ExtAbsCsec_fromIntField (Compute_ExtAbsCsec_fromIntField.f90:45-64);
 * C_ext, C_abs are complex (2*ntr doubles) as in the reference. */
int midas_cext_cabs(midas_ctx *ctx, const midas_cells_t *cells, double k0, int quirks, const double *E, long lde,
                    int ntr, const double *th, const double *ph, double *C_ext, double *C_abs);

/* ---------------- (f) near field ----------------
 * Compute_Scattered_Fields (Compute_Scattered_Fields.f90:75-136) with
 * Green_s_dt (Green_functions.f90:1335-1397), the path taken when the
 * far-field approximation is off (FFA = 0): the scattered field of E
 * (3n x 2ntr, lde) at receivers rso (cos th, sin th cos ph, sin th sin ph)
 * (Rso), projected on the receivers' theta / phi directions.  Es (nrx x 4ntr,
 * lds): Vv, Vh, Hv, Hh per transmitter.  FP64. */
int midas_near_field(midas_ctx *ctx, const midas_cells_t *cells, double k0, int quirks, const double *E, long lde,
                     int ntr, int nrx, const double *rth, const double *rph, double rso, double *Es, long lds);

/* ---------------- (f) orientation-averaged quantities ----------------
 * Compute_Scattering_Quantities_2 (Compute_Scattering_Quantities.f90:593-1013)
 * from S (nrx_tot x 4ntr, lds; Vv, Vh, Hv, Hh per transmitter, as
 * midas_scattering_matrix writes it): rows 0..nrx-1 are the scattering
 * receivers (rx_th, rx_ph), followed by the forward and backward directions
 * as get_trans_Receiv lays them out (forward = receiver t when ntr == nrx and
 * the receivers are sd / lb, else row nrx + t; backward = nrx + t when
 * ntr == nrx, else nrx + ntr + t).  rx_type: MIDAS_RX_SD / MIDAS_RX_RF (equal
 * weights 4 pi / nrx) or MIDAS_RX_LB (weights rx_w, Lebedev); tr_type the same
 * for the average over transmitters (tr_w for lb).  dm is the scatterer's
 * diameter (SimScatterer%dm): ap = dm/2, X = k0 dm/2.  C_ext, C_abs are the
 * complex internal-field cross sections (midas_cext_cabs).  Q (5 x ntr):
 * Q_ext (far field, or the internal-field value when cext_intfields),
 * Q_abs (internal field), Q_sca, Q_bks, g; Q_av (5): their averages. */
enum { MIDAS_RX_SD = 0, MIDAS_RX_RF = 1, MIDAS_RX_LB = 2 };
int midas_scattering_quantities(midas_ctx *ctx, const double *S, long lds, int nrx_tot, int nrx, int ntr,
                                const double *rx_th, const double *rx_ph, int rx_type, const double *rx_w,
                                const double *tr_th, const double *tr_ph, int tr_type, const double *tr_w, double k0,
                                double dm, int quirks, const double *C_ext, const double *C_abs, int cext_intfields,
                                double *Q, double *Q_av);

/* Compute_Scattering_Quantities_1 (Compute_Scattering_Quantities.f90:1-589),
 * the rule when the transmitters are integrated by tr_type = MIDAS_INT_UN
 * (the reference's single-precision QUADPACK qag, 61-point Gauss-Kronrod,
 * reading the integrand at the nearest grid node), _GL (Gauss-Legendre),
 * _TR (trapezoid cubature; the reference passes the lower phi bound twice,
 * so its average is 0) or _SM (no rule: 0).  Host routine, no context.
 * Transmitters (tr_th, tr_ph, degrees) form a theta x phi grid, theta
 * fastest; receivers rows 0..nrx-1 of S a grid of the same kind for rx_type
 * _UN / _GL / _TR / _SM, or a list for _SD (equal weights) / _LB (weights
 * rx_w).  The forward and backward directions are looked up among all nrx_tot
 * receivers (rx_th, rx_ph) within 1e-2 degree, as getTxRxIndex.  Other
 * arguments and outputs as midas_scattering_quantities. */
enum { MIDAS_INT_SD = 0, MIDAS_INT_LB = 1, MIDAS_INT_UN = 2, MIDAS_INT_GL = 3, MIDAS_INT_TR = 4, MIDAS_INT_SM = 5 };
int midas_scattering_quantities_grid(const double *S, long lds, int nrx_tot, const double *rx_th, const double *rx_ph,
                                     int nrx, int rx_type, const double *rx_w, int ntr, const double *tr_th,
                                     const double *tr_ph, int tr_type, double k0, double dm, int quirks,
                                     const double *C_ext, const double *C_abs, int cext_intfields, double *Q,
                                     double *Q_av);
/* The quadrature rules of the above, exposed for callers that integrate their
 * own grids: getAdaptiveQuadPts2 (Numerical_Integration.f90:213-245; ph = 0 for
 * 'th', 1 for 'ph'; f is d1 x d2, column jj 1-based), trap_2Dc (:251-305) and
 * the Gauss-Legendre rule of getGaussLegendreQuadPts3 (:88-145). */
int midas_quad_adaptive(int nth, int nph, int d1, int d2, const double *f, int ph, int jj, double a0, double b0,
                        double *I);
int midas_quad_trap2d(int nx, int ny, const double *f, double a, double b, double c, double d, double *I);
int midas_quad_gauss_legendre(int n, double a, double b, double *x, double *w);

/* ---------------- Compute_EFields_MoM drop-in ----------------
 * Compute_EFields_MoM(Cells, Transmitters, Receivers, S_total, C_ext, C_abs)
 * (Compute_EFields_MoM.f90:1-19): fill, solve, S_total (nrx x 4ntr),
 * C_ext/C_abs (ntr).  E (3n x 2ntr) is returned when non-NULL. */
int midas_compute_efields_mom(midas_ctx *ctx, const midas_cells_t *cells, double k0, int quirks, int ntr,
                              const double *tr_th, const double *tr_ph, int nrx, const double *rx_th,
                              const double *rx_ph, int mode, const midas_solve_opts_t *opts, double *S_total,
                              double *C_ext, double *C_abs, double *E, midas_solve_info_t *info);

/* ---------------- a14: ACA of one off-diagonal block ----------------
 * Calcul_MatrixZij_ACA (Calcul_ACA_matrices.f90:6-219): rows of cells
 * [r0, r0+nr), columns of cells [c0, c0+nc) of the full matrix over `cells`,
 * generated on the fly.  U (3nr x kmax), V (kmax x 3nc); *k_out as the
 * reference's nb_iter_out (kmax+1 when not converged). */
int midas_aca_block(midas_ctx *ctx, const midas_cells_t *cells, double k0, int quirks, int r0, int nr, int c0,
                    int nc, int kmax, double eps, double *U, double *V, int *k_out);

/* ---------------- a16: ACA of an explicit matrix ----------------
 * Calcul_GenMatrix_ACA (Calcul_ACA_matrices.f90:222-424). */
int midas_aca_matrix(midas_ctx *ctx, int m, int n, const double *A, long lda, int kmax, double eps, double *U,
                     double *V, int *k_out);

/* ---------------- a7-a15: CBFM-E (dense blocks) and CBFM-E + ACA ----------------
 * Compute_EFields_CBFME (Compute_EFields_CBFME.f90:1-1143) with the dense
 * block solve (SR = 0), and Compute_EFields_CBFME_ACA (use_aca = 1,
 * Compute_EFields_CBFME_ACA.f90:1-1017).  Blocks are contiguous cell ranges
 * block_start[b] .. block_start[b+1]-1 (cells already in block order, as
 * Division_blocks leaves them); extension cells of block b are
 * ext_cells[ext_start[b] .. ext_start[b+1]-1] (CBFM_Blocks_Ext).  CBF
 * excitations come from the ntr_cbf directions th_cbf, ph_cbf
 * (Transmitters_CBFM of getTransmitters_CBFM): iws = 1 plane waves
 * (Incident_Field), iws = 2 or 3 spherical waves from sources 10 block radii
 * around each extended block (Incident_Field_Spherical; type_iws,
 * Compute_EFields_CBFME.f90:363-369).  The problem excitations are ntr plane
 * waves.
 * threshold <= 0 selects the reference's rule (1e-3 with ACA, else
 * max(10^(-3-(ceil(max|m|)-2)), 1e-10) with |m| = sqrt(|eps|)).
 * Outputs: E (3n x 2ntr, lde); K_out[b] = number of CBFs of block b
 * (K_patchs); C_job (optional, may be NULL): the CBFs, block b's in rows
 * 3*block_start[b] .. and columns 0..K_b-1 (C_job_patchs), ldc >= 3n and
 * kcap columns; the call fails if some K_b exceeds kcap (K_b <= 2*ntr_cbf).
 * sr = 1 applies the reference's SR sparsification to every extended block
 * before its solve (setfSR with res_sr percent, SR_Green_s_tr_partial;
 * getParameters_CBFM.f90:113-168, Green_functions.f90:290-806); the masked
 * block is solved by the dense LU instead of PARDISO.  fsr_out[b] and
 * nnz_out[b] (optional) receive the chosen fSR and the kept entries. */
typedef struct {
    int nblocks;
    const int *block_start;   /* nblocks + 1 */
    const int *ext_start;     /* nblocks + 1 */
    const int *ext_cells;     /* 0-based cell indices */
} midas_blocks_t;

int midas_compute_efields_cbfm(midas_ctx *ctx, const midas_cells_t *cells, const midas_blocks_t *blocks, double k0,
                               int quirks, int ntr_cbf, const double *th_cbf, const double *ph_cbf, int iws, int ntr,
                               const double *tr_th, const double *tr_ph, int use_aca, int kmax, double eps_aca,
                               double threshold, int mode, double *E, long lde, int *K_out, double *C_job, long ldc,
                               int kcap, int sr, double res_sr, double *fsr_out, long *nnz_out,
                               midas_solve_info_t *info);

/* ---------------- a17: block tables (host routines, no device work) ----------------
 * Blocks are contiguous cell ranges block_start[b] .. block_start[b+1]-1.
 * midas_block_frames: SetBlockCtrs (Division_blocks.f90:1616-1660); 13 doubles
 * per block: BCubCont min x,y,z, max x,y,z, extent x,y,z, then BSphCont
 * centre x,y,z (the box centre) and radius (largest distance to it).
 * midas_extend_blocks: Extend_blocks (Extend_blocks.f90:1-162).  With sc the
 * cell size (SimScatterer%Sc) and nc_extended layers: blocks whose bounding
 * spheres lie within R_B + R_B' + nc_extended*sc are adjacent; the extension
 * of block B is every cell of an adjacent block inside B's box grown by
 * nc_extended*sc + sc/2, in cell order (0-based indices).  ext_start
 * (nblocks + 1) always receives the offsets; ext_cells (capacity ext_cap) may
 * be NULL to size it first.  adj_start / adj (optional) receive the adjacent
 * blocks that gave cells (CBFM_Block%Badj).  No extension when nblocks == 1 or
 * nc_extended == 0.  Fails (< 0) when an extension exceeds 10 times the
 * largest block (Fact_Nbext_max, where the reference stops). */
int midas_block_frames(const midas_cells_t *cells, int nblocks, const int *block_start, double *frames);

/* Division of the cells into CBFM blocks (host routines, no context).  Both
 * reorder the cells so that every block is a contiguous range, as the
 * reference leaves Cells: order[i] (n entries) is the input index of the cell
 * at position i; block_size (capacity cap; n always suffices) receives the
 * sizes in block order and *nblocks their count; *hblock (optional) = 2 x the
 * largest block radius, the reference's final hBlock.  Cells keep their input
 * order inside a block.
 * midas_partition_sph: div_type 'SPH' (Division_blocks.f90:95-132 with
 * SPH_v = 2, Division_blocks_sph_H :1252-1401, HierarchicalOctree_subdivision
 * :1403-1537): octree splits at 1/2 of the centre extent; a block of N >
 * navg_cells cells with |N - navg_cells| > |nint(N/8.) - navg_cells| is split
 * again.  best_combin = 1 is SPH_v = 3 (9 split fractions tried per block).
 * midas_partition_csh: div_type 'CSH' (Division_blocks.f90:133-180,
 * Division_blocks_csh :190-652): slabs along the longest side of the box
 * (centre extremes -+ sc/2, sc = SimScatterer%Sc), then halving cuts with the
 * reference's balancing search, over up to 8 trial hBlock values;
 * type_s = 3 (cylinder) stops after the slabs.  Fails (< 0) where the
 * reference reports a failed division. */
int midas_partition_sph(const midas_cells_t *cells, int navg_cells, int best_combin, int *order, int *block_size,
                        int cap, int *nblocks, double *hblock);
int midas_partition_csh(const midas_cells_t *cells, double sc, int navg_cells, int type_s, int *order,
                        int *block_size, int cap, int *nblocks, double *hblock);
int midas_extend_blocks(const midas_cells_t *cells, int nblocks, const int *block_start, double sc, int nc_extended,
                        int *ext_start, int *ext_cells, long ext_cap, int *adj_start, int *adj, long adj_cap);

/* ---------------- device-pointer variants (inputs resident in HBM) ----------------
 * Same semantics as above; the pointers named below are device memory; ordered
 * on midas_stream(ctx).  midas_dev_mom_solve fills Z (FP64) and, in FP64 mode
 * and in the FP64 fallback, overwrites it with its LU factors.  A mixed solve
 * with at most two directions (schedule parameter "mom_matrix_free", on by
 * default) computes its FP64 residuals and ||Z||_inf from the cells instead
 * and leaves Z untouched unless the fallback runs. */
int midas_dev_mom_fill(midas_ctx *ctx, const midas_cells_t *dcells, double k0, int quirks, double *Z, long ldz,
                       float *W32, long ldw);
int midas_dev_mom_solve(midas_ctx *ctx, const midas_cells_t *dcells, double k0, int quirks, int ntr,
                        const double *dth, const double *dph, int mode, const midas_solve_opts_t *opts, double *Z,
                        long ldz, double *X, long ldx, midas_solve_info_t *info, double *t_fill_ms);
/* midas_dev_mom_solve_noalloc: midas_dev_mom_solve that allocates no device
 * memory, for callers that cannot (code under stream capture, tight serving
 * loops, harnesses that forbid it).  No counterpart in the reference.  Every
 * work buffer comes from a static device array of 76 MiB that each GPU the
 * library runs on holds until the process ends; midas_init finds its address.
 * Limits: 3n <= 2048 unknowns, 2 ntr <= 32 right-hand sides (larger problems
 * fail with an error and no work done).  It runs the same kernels with
 * 256-column panels (128 in FP64 and in the FP64 fallback), without the
 * look-ahead stream, and only this library's GEMM kernels (no rocBLAS);
 * t_fill_ms is not measured (no timing events).  All calls on one GPU share
 * the array: calls that overlap on one GPU race. */
int midas_dev_mom_solve_noalloc(midas_ctx *ctx, const midas_cells_t *dcells, double k0, int quirks, int ntr,
                                const double *dth, const double *dph, int mode, const midas_solve_opts_t *opts,
                                double *Z, long ldz, double *X, long ldx, midas_solve_info_t *info);
/* midas_dev_compute_efields_cbfm: midas_compute_efields_cbfm with the fields
 * E_dev (3n x 2ntr, ld lde) in device memory, so they stay in HBM
 * (Compute_EFields_CBFME.f90 / Compute_EFields_CBFME_ACA.f90, the same
 * interface as midas_compute_efields_cbfm); the cell and block tables, the
 * directions, K_out, C_job and the other outputs stay host memory, as there. */
int midas_dev_compute_efields_cbfm(midas_ctx *ctx, const midas_cells_t *cells, const midas_blocks_t *blocks,
                                   double k0, int quirks, int ntr_cbf, const double *th_cbf, const double *ph_cbf,
                                   int iws, int ntr, const double *tr_th, const double *tr_ph, int use_aca, int kmax,
                                   double eps_aca, double threshold, int mode, double *E_dev, long lde, int *K_out,
                                   double *C_job, long ldc, int kcap, int sr, double res_sr, double *fsr_out,
                                   long *nnz_out, midas_solve_info_t *info);

#ifdef __cplusplus
}
#endif
#endif /* MIDAS_GPU_H */
