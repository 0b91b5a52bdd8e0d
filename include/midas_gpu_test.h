/*
 * midas_gpu_test.h -- kernel-level entry points used by the test suite to
 * check the building blocks of the solver in isolation (no reference
 * counterpart; not part of the drop-in interface).  Host pointers,
 * synchronous, same return convention as midas_gpu.h.
 */
#ifndef MIDAS_GPU_TEST_H
#define MIDAS_GPU_TEST_H
#include "midas_gpu.h"
#ifdef __cplusplus
extern "C" {
#endif

/* C = alpha op(A) op(B) + beta C, complex, column-major; prec 64 (ZGEMM on
 * FP64 MFMA, 64 x 64 tiles; without transposes the kernel of 65), 65 (ZGEMM,
 * the kernel of the FP64 LU and the residuals, its tile shape the zgemm_tile
 * parameter; no transposes), 66 (that kernel as a strided batch of
 * three: A, B and C each hold three matrices side by side, m x k, k x n and
 * m x n) or 32 (CGEMM on FP32 MFMA, operands given as complex128 and
 * rounded); transA/transB: 0 = N, 1 = T. */
int midas_test_gemm(midas_ctx *ctx, int prec, int m, int n, int k, const double *alpha, const double *A, long lda,
                    int transA, const double *B, long ldb, int transB, const double *beta, double *C, long ldc);

/* LU with partial pivoting of one n x n complex matrix by the library's
 * factorisation: A (complex128, column-major, lda) receives the unit-lower L
 * and U; ipiv[j] (0-based) the row interchanged with row j.  prec 64 (FP64)
 * or 32 (FP32: A rounded to complex64, the factors returned widened); nb the
 * panel width (<= 512); pivot_search 1 restricts each panel's search to its
 * diagonal block and returns the largest |L21| in *max_l21 (else 0). */
int midas_test_getrf(midas_ctx *ctx, int prec, int n, double *A, long lda, int nb, int pivot_search, int *ipiv,
                     double *max_l21);

/* Eigen-decomposition of a batch of Hermitian n x n complex matrices G
 * (batch matrices one after another, column-major) by the CBF subspaces'
 * batched eigensolver: D (n per matrix) ascending, V (n x n per matrix) the
 * eigenvectors, column c for D[c].  path: 0 by order, 1 the one-stage kernel
 * (all pairs), 2 the large-order kernel, which with cut > 0 computes only the
 * pairs with lambda >= cut lambda_max (the others: 0 and zero columns). */
int midas_test_heev(midas_ctx *ctx, int n, int batch, const double *G, double *D, double *V, double cut, int path);

/* Pivoted Cholesky G ~= L L^H of a batch of Hermitian positive semi-definite
 * n x n matrices (the CBF Gram matrices' rank reduction), stopped where the
 * largest remaining diagonal is <= tolrel max diag(G): L n x n per matrix
 * (columns past the rank zero), rank[b] its number of columns. */
int midas_test_pchol(midas_ctx *ctx, int n, int batch, const double *G, double *L, int *rank, double tolrel);

/* C -= A B with the split FP16/BF16 trailing-update kernel (mode as
 * TrailingMode: 1 f16, 3 f16x3, 11 bf16, 13 bf16x3); operands complex128,
 * rounded to complex FP32 before the call.  A (lda x k) and C (ldc x n) keep
 * their leading dimensions on the device: the kernel works on the m-row
 * blocks inside them, and rows m .. ld-1 come back rounded to FP32. */
int midas_test_trailing(midas_ctx *ctx, int mode, int m, int n, int k, const double *A, long lda, const double *B,
                        long ldb, double *C, long ldc);

/* Matrix-core lane layout probe: which = 0 (f64 16x16x4), 1 (f32 16x16x4),
 * 2 (f16 16x16x32), 3 (bf16 16x16x32), 4 (f16 32x32x16), 5 (bf16 32x32x16).
 * out receives the product of A(i,k) = (i + 2k) % 7 - 3 and
 * B(k,j) = (3k + j) % 5 - 2 read back with the lane layout the kernels assume
 * (column-major 16x16, or 32x32 for which >= 4). */
int midas_test_mfma_probe(midas_ctx *ctx, int which, double *out);

#ifdef __cplusplus
}
#endif
#endif
