// Batched adaptive cross approximation (see midas_aca.hip).
#pragma once
#include "midas_common.h"

namespace midas {

constexpr int ACA_KMAX_SUPPORTED = 64;

// One ACA problem: an (m x n) block, rows of cells r0.., columns of cells c0..
// (cell generators), U at U + u_off (m x kmax, ld m), Vt at Vt + v_off
// (n x kmax, ld n; the transpose of the reference's Matrix_V).
struct AcaJob {
    int m, n;
    int r0, c0;
    long u_off, v_off;
};

// gram: aca_gram_bytes(kmax, njobs) of device scratch (the factors' Gram
// matrices, two kmax x kmax per job)
inline size_t aca_gram_bytes(int kmax, int njobs) { return (size_t)njobs * 2 * kmax * kmax * sizeof(double2); }
int launch_aca_cells(const DevCells &cells, bool ref32, double k0, double pi, const AcaJob *jobs, int njobs,
                     double2 *U, double2 *Vt, int kmax, double eps, int *kout, double2 *gram, hipStream_t st);
int launch_aca_matrix(const double2 *A, long lda, const AcaJob *jobs, int njobs, double2 *U, double2 *Vt, int kmax,
                      double eps, int *kout, double2 *gram, hipStream_t st);
int launch_vt_to_v(const double2 *Vt, int n, int kmax, double2 *V, long ldv, hipStream_t st);

}  // namespace midas
