// Characteristic-basis-function method (CBFM-E) on gfx950.
//
//  Compute_EFields_CBFME      (Compute_EFields_CBFME.f90:348-449, 630-849, 896-1006)
//      with the dense block solve (SR = 0), and
//  Compute_EFields_CBFME_ACA  (Compute_EFields_CBFME_ACA.f90:296-385, 569-636).
//
// Pipeline (all on the device, one context stream):
//   1. per block b: cells = block cells then its extension cells (s_b of them);
//      Z_b = Green_s_tr_partial(ext block, ext block) -- with sr, sparsified by
//      the reference's SR mask and fSR search -- and the CBF excitations
//      E_b = Incident_Field(ext block, CBF directions) (FP32 intermediates when
//      the FP32-blocks quirk is set); X_b = Z_b^-1 E_b by the FP64 LU of this
//      library.  The reference takes the SVD X_b = U S W^H (GESVD, JOB='U');
//      only sigma and the span of the leading left vectors enter the result,
//      so here they come from the Gram matrix X^H X (or X X^H when X is wide):
//      one GEMM per block, pivoted Cholesky G = L L^H to the numerical rank r,
//      and ONE launch of this library's batched Hermitian eigensolver on the
//      r x r matrices L^H L (midas_heev.hip) for all blocks, U = X W S^-1.  For
//      thresholds below 1e-5 the squared spectrum is not resolved well enough
//      and each block takes rocSOLVER gesvd instead.
//      K_b = #{sigma_d / sigma_1 >= threshold}; C_b = U(1:3 n_b, 1:K_b).
//   2. reduced matrix Zc(K x K), block (i, j) = C_i^T Z_ij C_j (plain transpose,
//      as the reference's ZGEMM('T', ...)):
//        dense path: one fill of the column strip Z(:, block j) (3N x 3n_j),
//                    T = Z(:, j) C_j as one GEMM, then every C_i^T T(rows of i)
//                    in one grouped launch of small GEMM tiles;
//        ACA path:   diagonal blocks dense; every off-diagonal pair (i, j) by ACA
//                    (one workgroup per pair, a block column per launch batch),
//                    then (C_i^T U)(V C_j) for all pairs in two grouped launches
//                    when the ACA rank is below kmax, and the dense product
//                    otherwise (the reference's fallback).
//   3. Vc = C_i^T E_inc(block i) (FP32 Incident_Field quirk), alpha = Zc^-1 Vc
//      (FP64 LU or mixed-precision LU + FP64 refinement, the `mode`), and
//      E(block i) = C_i alpha_i.
#include "midas_aca.h"
#include "midas_elem.h"
#include "midas_host.h"
#include "midas_lu.h"
#include "midas_tune.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <vector>

namespace midas {

#define MIDAS_ROC(call)                                                                             \
    do {                                                                                            \
        rocblas_status s_ = (call);                                                                 \
        if (s_ != rocblas_status_success) {                                                         \
            ::midas::set_error("%s:%d %s: rocblas status %d", __FILE__, __LINE__, #call, (int)s_); \
            return -90;                                                                             \
        }                                                                                           \
    } while (0)

struct RocHandle {
    rocblas_handle h = nullptr;
    ~RocHandle()
    {
        if (h) rocblas_destroy_handle(h);
    }
};

// Compute_EFields_CBFME.f90:286-288 (m = sqrt(eps)) and Compute_EFields_CBFME_ACA.f90:124
// (a single-precision literal 1e-3 widened to double).
static double cbf_threshold_rule(const midas_cells_t *c, int use_aca)
{
    if (use_aca) return (double)1e-3f;
    double mx = 0.0;
    for (int i = 0; i < c->n; ++i) mx = std::max(mx, std::sqrt(std::hypot(c->eps[2 * i], c->eps[2 * i + 1])));
    return std::max(std::pow(10.0, -3.0 - (std::ceil(mx) - 2.0)), (double)1e-10f);
}

// (m x n) column-major sub-block copy, device to device
static int copy2d(const double2 *S, long lds, double2 *D, long ldd, int m, int n, hipStream_t st)
{
    if (m <= 0 || n <= 0) return 0;
    MIDAS_HIP(hipMemcpy2DAsync(D, sizeof(double2) * ldd, S, sizeof(double2) * lds, sizeof(double2) * m, n,
                               hipMemcpyDeviceToDevice, st));
    return 0;
}

typedef cplx<double> zc;
static const zc ONE = {1.0, 0.0}, ZERO = {0.0, 0.0};

// The Gram route squares the singular values: FP64 resolves sigma_d / sigma_1
// to a few digits down to ~1e-6 (eigenvalue ratio 1e-12 against an absolute
// eigenvalue error of ~n eps lambda_1); smaller thresholds take the SVD.
constexpr double CBF_GRAM_MIN_THRESHOLD = 1e-5;
// The Gram matrices are reduced to their numerical rank by pivoted Cholesky,
// stopped where the largest remaining diagonal is <= this times the largest
// diagonal: the dropped semi-definite remainder is at the level of the
// eigensolver's own backward error (a few hundred eps |G|), far below the
// smallest kept eigenvalue (threshold^2 |G| >= 1e-10 |G|).
constexpr double CBF_GRAM_RANK_TOL = 100.0 * 2.220446049250313e-16;

// out(:, c) = W(0:rows, rpad-1-c) * lambda[rpad-1-c]^(-power/2), c < K: the
// eigenvectors of the K largest eigenvalues (ascending storage), largest
// first, scaled by lambda^-1/2 (power 1) or lambda^-1 (power 2).
__global__ void k_pick_eigvecs(const double2 *W, long ldw, const double *lam, int rpad, int rows, int K, int power,
                               double2 *out, long ldo)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (r >= rows || c >= K) return;
    const int src = rpad - 1 - c;
    const double s = power == 2 ? 1.0 / lam[src] : power == 1 ? 1.0 / sqrt(lam[src]) : 1.0;
    double2 v = W[(long)src * ldw + r];
    out[(long)c * ldo + r] = make_double2(v.x * s, v.y * s);
}
static int launch_pick_eigvecs(const double2 *W, long ldw, const double *lam, int rpad, int rows, int K, int power,
                               double2 *out, long ldo, hipStream_t st)
{
    hipLaunchKernelGGL(k_pick_eigvecs, dim3((rows + 255) / 256, K), dim3(256), 0, st, W, ldw, lam, rpad, rows, K,
                       power, out, ldo);
    MIDAS_HIP(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// Mixed-precision block solves: the batch of extended-block systems
// Z_b X_b = E_b factored in complex FP32 (trailing updates on the split FP16 /
// BF16 or FP32 matrix cores, one batched LU) and refined in FP64 against the
// FP64 Z_b until every block's normwise backward error
// max_j |e_j - Z x_j|_inf / (|Z|_inf |x_j|_inf + |e_j|_inf) is <= tol; a
// block that does not get there is solved again by the FP64 LU.  The
// reference's sparse block solver refines its direct solution the same way
// (pardiso_solver.f90:123-142, iparm(8) refinement steps).
// ---------------------------------------------------------------------------
__global__ void k_flat_z2c(const double2 *S, float2 *D, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        D[i] = make_float2((float)S[i].x, (float)S[i].y);
}
__global__ void k_flat_c2z_add(const float2 *S, double2 *D, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        D[i] = make_double2(D[i].x + (double)S[i].x, D[i].y + (double)S[i].y);
}
// row sums of |Z_b| (row r of block blockIdx.y), then their maximum per block
__global__ void k_blk_rowsum(const double2 *Z, long ld, long zs, int m, double *rs)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m) return;
    const double2 *z = Z + blockIdx.y * zs + r;
    double s = 0.0;
    for (int c = 0; c < m; ++c) {
        const double2 v = z[(long)c * ld];
        s += sqrt(v.x * v.x + v.y * v.y);
    }
    rs[(long)blockIdx.y * m + r] = s;
}
__global__ void k_blk_max(const double *v, int m, double *out)
{
    __shared__ double red[4];
    double x = 0.0;
    for (int i = threadIdx.x; i < m; i += blockDim.x) x = fmax(x, v[(long)blockIdx.x * m + i]);
    for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}
// column maxima of |R|, |X|, |E| for column blockIdx.x of block blockIdx.y
__global__ void k_blk_colmax3(const double2 *R, const double2 *X, const double2 *E, int m, long ld, long xs, int ncol,
                              double *out)
{
    __shared__ double red[3][4];
    const long off = blockIdx.y * xs + (long)blockIdx.x * ld;
    double a = 0.0, b = 0.0, c = 0.0;
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
        const double2 r = R[off + i], x = X[off + i], e = E[off + i];
        a = fmax(a, fmax(fabs(r.x), fabs(r.y)));
        b = fmax(b, fmax(fabs(x.x), fabs(x.y)));
        c = fmax(c, fmax(fabs(e.x), fabs(e.y)));
    }
    for (int o = 32; o > 0; o >>= 1) {
        a = fmax(a, __shfl_xor(a, o, 64));
        b = fmax(b, __shfl_xor(b, o, 64));
        c = fmax(c, __shfl_xor(c, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = a;
        red[1][threadIdx.x >> 6] = b;
        red[2][threadIdx.x >> 6] = c;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int q = threadIdx.x;
        out[((long)blockIdx.y * ncol + blockIdx.x) * 3 + q] =
            fmax(fmax(red[q][0], red[q][1]), fmax(red[q][2], red[q][3]));
    }
}

// X (nc blocks of mpc x T2c at stride mpc T2c, ld mpc) holds E on entry and
// X on return; Z (nc blocks of mpc x mpc, ld mpc) stays unmodified.
static int block_solve_mixed(midas_ctx *ctx, int mode, int nc, int mpc, int T2c, const double2 *Z, double2 *X,
                             int *ipiv, int *perm, std::vector<int> &hinfo, BlockMixedWs &bw, int nb, int nb64,
                             hipStream_t st)
{
    const long zs = (long)mpc * mpc, xs = (long)mpc * T2c, nx = xs * nc;
    const dim3 fg(2048), fb(256);
    TRY(bw.W.reserve(sizeof(float2) * (size_t)zs * nc));
    hipLaunchKernelGGL(k_flat_z2c, fg, fb, 0, st, Z, bw.W.as<float2>(), zs * nc);
    TRY(bw.E.reserve(sizeof(double2) * (size_t)nx));
    TRY(bw.R.reserve(sizeof(double2) * (size_t)nx));
    TRY(bw.D.reserve(sizeof(float2) * (size_t)nx));
    MIDAS_HIP(hipMemcpyAsync(bw.E.p, X, sizeof(double2) * nx, hipMemcpyDeviceToDevice, st));
    // ||Z_b||_inf
    TRY(bw.rs.reserve(sizeof(double) * (size_t)mpc * nc));
    TRY(bw.anorm.reserve(sizeof(double) * nc));
    hipLaunchKernelGGL(k_blk_rowsum, dim3((mpc + 255) / 256, nc), dim3(256), 0, st, Z, (long)mpc, zs, mpc,
                       bw.rs.as<double>());
    hipLaunchKernelGGL(k_blk_max, dim3(nc), dim3(256), 0, st, bw.rs.as<double>(), mpc, bw.anorm.as<double>());
    MIDAS_HIP(hipGetLastError());
    // batched complex-FP32 LU (workspace kept with the batch buffers)
    LuWork<float> &w = bw.lu;
    w.share_stream(ctx->ps, ctx->pev[0], ctx->pev[1]);
    TRY(w.alloc(mpc, nb, nc, zs));
    TrailingGemm<float> tg;
    tg.mode = mode == MIDAS_SOLVE_C32_IR ? TRAIL_NATIVE : mode == MIDAS_SOLVE_BF16_IR ? TRAIL_BF16_1 : TRAIL_F16_1;
    tg.batch = nc;
    tg.stride = zs;
    if (tg.mode != TRAIL_NATIVE) {
        TRY(bw.scr.reserve(trail16_batch_bytes(mpc, mpc, nb, nc)));
        tg.scratch = bw.scr.p;
    }
    // the block systems are MoM matrices of the extended blocks: the pivot
    // search stays in the panels' diagonal blocks (refinement and the FP64
    // fallback below keep each block's answer whatever the pivots)
    w.diag_pivot = tune(TUNE_LU_PIVOT_SEARCH) >= 1;
    if (w.diag_pivot) {
        const size_t sl = sizeof(float2) * (size_t)mpc * nb * nc;
        TRY(bw.l21.reserve(sl + 256));
        w.l21 = bw.l21.as<c32>();
        w.lmax = (unsigned *)((char *)bw.l21.p + sl);
    }
    TRY(getrf<float>(bw.W.as<float2>(), mpc, mpc, ipiv, w, st, tg));
    TRY(ipiv_to_perm_dev(ipiv, perm, mpc, nc, st));
    TRY(copy_sync(hinfo.data(), w.info, sizeof(int) * nc, hipMemcpyDeviceToHost, st));
    // refinement: X = 0, R = E; D = LU^-1 R (FP32), X += D, R = E - Z X (FP64)
    TRY(bw.tmp.reserve(sizeof(float2) * (size_t)getrs_tmp_elems(mpc, nb, T2c) * nc));
    TRY(bw.cm.reserve(sizeof(double) * 3 * (size_t)T2c * nc));
    MIDAS_HIP(hipMemsetAsync(X, 0, sizeof(double2) * nx, st));
    MIDAS_HIP(hipMemcpyAsync(bw.R.p, bw.E.p, sizeof(double2) * nx, hipMemcpyDeviceToDevice, st));
    const double tol = std::max(1e-14, std::sqrt((double)mpc) * DBL_EPSILON);
    std::vector<double> cm(3 * (size_t)T2c * nc), an(nc), berr(nc, INFINITY), prev(nc, INFINITY);
    TRY(copy_sync(an.data(), bw.anorm.p, sizeof(double) * nc, hipMemcpyDeviceToHost, st));
    std::vector<char> open_(nc, 1);   // still refining
    for (int it = 0; it < 10; ++it) {
        hipLaunchKernelGGL(k_flat_z2c, fg, fb, 0, st, bw.R.as<double2>(), bw.D.as<float2>(), nx);
        TRY(getrs<float>(bw.W.as<float2>(), mpc, mpc, perm, w, bw.D.as<float2>(), mpc, T2c, bw.tmp.as<float2>(), st,
                         xs));
        hipLaunchKernelGGL(k_flat_c2z_add, fg, fb, 0, st, bw.D.as<float2>(), X, nx);
        MIDAS_HIP(hipMemcpyAsync(bw.R.p, bw.E.p, sizeof(double2) * nx, hipMemcpyDeviceToDevice, st));
        TRY(launch_zgemm_nn(mpc, T2c, mpc, z64{-1.0, 0.0}, Z, mpc, zs, X, mpc, xs, z64{1.0, 0.0}, bw.R.p, mpc, xs, nc,
                            st, tune(TUNE_ZGEMM_TILE)));
        hipLaunchKernelGGL(k_blk_colmax3, dim3(T2c, nc), dim3(256), 0, st, bw.R.as<double2>(), X, bw.E.as<double2>(),
                           mpc, (long)mpc, xs, T2c, bw.cm.as<double>());
        MIDAS_HIP(hipGetLastError());
        TRY(copy_sync(cm.data(), bw.cm.p, sizeof(double) * cm.size(), hipMemcpyDeviceToHost, st));
        bool any = false;
        for (int q = 0; q < nc; ++q) {
            double e = 0.0;
            for (int j = 0; j < T2c; ++j) {
                const double *c = &cm[((size_t)q * T2c + j) * 3];
                const double den = an[q] * c[1] + c[2];
                e = std::max(e, den > 0.0 ? c[0] / den : (c[0] > 0.0 ? INFINITY : 0.0));
            }
            prev[q] = berr[q];
            berr[q] = e;
            // a block keeps refining while above tol and still gaining
            open_[q] = std::isfinite(e) && e > tol && e < 0.5 * prev[q];
            any = any || open_[q];
        }
        if (!any) break;
    }
    // blocks above tol: FP64 LU of their Z (unmodified) and E
    for (int q = 0; q < nc; ++q) {
        if (berr[q] <= tol) continue;
        LuWork<double> w64;
        w64.share_stream(ctx->ps, ctx->pev[0], ctx->pev[1]);
        TRY(w64.alloc(mpc, nb64));
        TrailingGemm<double> tg64;
        DBuf Ab;
        TRY(Ab.alloc(sizeof(double2) * (size_t)zs));
        double2 *A = Ab.as<double2>();
        MIDAS_HIP(hipMemcpyAsync(A, Z + q * zs, sizeof(double2) * zs, hipMemcpyDeviceToDevice, st));
        TRY(getrf<double>(A, mpc, mpc, ipiv, w64, st, tg64));
        TRY(ipiv_to_perm_dev(ipiv, perm, mpc, 1, st));
        MIDAS_HIP(hipMemcpyAsync(X + q * xs, bw.E.as<double2>() + q * xs, sizeof(double2) * xs, hipMemcpyDeviceToDevice,
                                 st));
        DBuf t64;
        TRY(t64.alloc(sizeof(double2) * (size_t)getrs_tmp_elems(mpc, nb64, T2c)));
        TRY(getrs<double>(A, mpc, mpc, perm, w64, X + q * xs, mpc, T2c, t64.as<double2>(), st));
        int inf1 = 0;
        TRY(copy_sync(&inf1, w64.info, sizeof(int), hipMemcpyDeviceToHost, st));
        hinfo[q] = inf1;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Grouped small GEMMs: every job is one 32 x 32 tile of C = op(A) B, complex
// FP64, op = none or plain transpose.  The reduced-matrix assembly has
// thousands of independent products of a few tens of columns (C_i^T T, C_i^T U,
// V C_j, L R); one launch carries all of them instead of one launch each.
// ---------------------------------------------------------------------------
struct GJob {
    const double2 *A, *B;
    double2 *C;
    int lda, ldb, ldc;
    int m, n, k;    // op(A) m x k, B k x n
    int tA;         // 1: op(A) = A^T (A stored k x m)
    int i0, j0;     // tile origin in C
};

constexpr int GT = 32, GK = 16;
typedef double gd4 __attribute__((ext_vector_type(4)));
// One 32 x 32 tile per workgroup on the FP64 matrix cores: wave w owns the
// 16 x 16 quadrant (16 (w & 1), 16 (w >> 1)) and runs v_mfma_f64_16x16x4_f64
// (lane l supplies A(l % 16, l / 16) and B(l / 16, l % 16); it receives rows
// l / 16 + 4 q of column l % 16).  The four real products of the complex
// multiply-add accumulate in four independent chains, summed at the end.
__global__ __launch_bounds__(256) void k_ggemm(const GJob *__restrict__ jobs)
{
    const GJob jb = jobs[blockIdx.x];
    __shared__ double2 sA[GK][GT + 1], sB[GK][GT + 1];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = (w & 1) * 16, wn = (w >> 1) * 16;
    gd4 rr = {0, 0, 0, 0}, ii = {0, 0, 0, 0}, ri = {0, 0, 0, 0}, ir = {0, 0, 0, 0};
    // this thread's two elements of each stage's op(A) tile (GT rows x GK k)
    // and B tile (GK k x GT cols); the next stage's are loaded into registers
    // while the current one computes
    int ak[2], aa[2], bk[2], bb[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const int idx = t + 256 * e;
        if (jb.tA) { ak[e] = idx & 15; aa[e] = idx >> 4; }   // A stored k x m: unit stride along k
        else { aa[e] = idx & 31; ak[e] = idx >> 5; }          // A stored m x k: unit stride along m
        bk[e] = idx & 15;
        bb[e] = idx >> 4;
    }
    double2 ra[2], rb[2];
    auto gload = [&](int k0) {
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const int a = jb.i0 + aa[e], kg = k0 + ak[e];
            ra[e] = make_double2(0.0, 0.0);
            if (a < jb.m && kg < jb.k) ra[e] = jb.tA ? jb.A[(long)a * jb.lda + kg] : jb.A[(long)kg * jb.lda + a];
            const int b = jb.j0 + bb[e], kb = k0 + bk[e];
            rb[e] = (b < jb.n && kb < jb.k) ? jb.B[(long)b * jb.ldb + kb] : make_double2(0.0, 0.0);
        }
    };
    gload(0);
    for (int k0 = 0; k0 < jb.k; k0 += GK) {
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            sA[ak[e]][aa[e]] = ra[e];
            sB[bk[e]][bb[e]] = rb[e];
        }
        __syncthreads();
        if (k0 + GK < jb.k) gload(k0 + GK);
#pragma unroll
        for (int kk = 0; kk < GK; kk += 4) {
            const double2 x = sA[kk + (lane >> 4)][wm + (lane & 15)];
            const double2 y = sB[kk + (lane >> 4)][wn + (lane & 15)];
            rr = __builtin_amdgcn_mfma_f64_16x16x4f64(x.x, y.x, rr, 0, 0, 0);
            ii = __builtin_amdgcn_mfma_f64_16x16x4f64(x.y, y.y, ii, 0, 0, 0);
            ri = __builtin_amdgcn_mfma_f64_16x16x4f64(x.x, y.y, ri, 0, 0, 0);
            ir = __builtin_amdgcn_mfma_f64_16x16x4f64(x.y, y.x, ir, 0, 0, 0);
        }
        __syncthreads();
    }
    const int b = jb.j0 + wn + (lane & 15);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int a = jb.i0 + wm + (lane >> 4) + 4 * q;
        if (a < jb.m && b < jb.n) jb.C[(long)b * jb.ldc + a] = make_double2(rr[q] - ii[q], ri[q] + ir[q]);
    }
}

// All tiles of C = op(A) B (m x n) appended to a job list.
static void add_gemm_tiles(std::vector<GJob> &jobs, const double2 *A, int lda, int tA, const double2 *B, int ldb,
                           double2 *C, int ldc, int m, int n, int k)
{
    for (int j0 = 0; j0 < n; j0 += GT)
        for (int i0 = 0; i0 < m; i0 += GT) jobs.push_back(GJob{A, B, C, lda, ldb, ldc, m, n, k, tA, i0, j0});
}

static int run_ggemm(const std::vector<GJob> &jobs, DBuf &djobs, hipStream_t st)
{
    if (jobs.empty()) return 0;
    TRY(djobs.reserve(sizeof(GJob) * jobs.size()));
    MIDAS_HIP(hipMemcpyAsync(djobs.p, jobs.data(), sizeof(GJob) * jobs.size(), hipMemcpyHostToDevice, st));
    for (size_t q0 = 0; q0 < jobs.size(); q0 += 1u << 30) {
        const unsigned cnt = (unsigned)std::min<size_t>(jobs.size() - q0, 1u << 30);
        hipLaunchKernelGGL(k_ggemm, dim3(cnt), dim3(256), 0, st, djobs.as<GJob>() + q0);
    }
    MIDAS_HIP(hipGetLastError());
    // the host job list may be reused right after: wait for the copy
    MIDAS_HIP(hipStreamSynchronize(st));
    return 0;
}

// ---------------------------------------------------------------------------
// SR (sparsified) block solves: setfSR (getParameters_CBFM.f90:113-168) with
// Green_s_tr_partial_FN and SR_Green_s_tr_partial_FN (Green_functions.f90:
// 176-287, 564-806), and the mask of SR_Green_s_tr_partial (290-561).  The
// reference hands the masked block to PARDISO; here the same masked matrix is
// solved by the dense LU (a direct solve of the same linear system).
// ---------------------------------------------------------------------------
constexpr int NFSR = 11;
struct FsrList { double f[NFSR]; };
static const FsrList FSR_TESTS = {{1e2, 1e3, 2e3, 3e3, 4e3, 5e3, 6e3, 7e3, 8e3, 9e3, 1e4}};
struct SrStats {
    double fn2;                     // squared Frobenius norm of the block
    double g2[NFSR];                // squared norm of the SR mask, per fSR
    unsigned long long nnz[NFSR];   // entries kept, per fSR
};

// One thread per (receiver, source) pair of block b's extended cell list.
__global__ __launch_bounds__(64) void k_sr_stats(DevCells cells, CellSet S, double k0, double pi, FsrList fl,
                                                 SrStats *out)
{
    const int o = blockIdx.x * 64 + threadIdx.x;
    const int sl = blockIdx.y;
    double fn2 = 0.0, g2[NFSR];
    unsigned nz[NFSR];
#pragma unroll
    for (int j = 0; j < NFSR; ++j) {
        g2[j] = 0.0;
        nz[j] = 0;
    }
    if (o < S.n) {
        const int go = S.cell(o), gs = S.cell(sl), g0 = S.cell(0);
        const double v0 = hypot(1.0 - cells.znnpp[2 * g0], -cells.znnpp[2 * g0 + 1]);
        if (o == sl) {
            const double a = hypot(1.0 - cells.znnpp[2 * gs], -cells.znnpp[2 * gs + 1]);
            fn2 = 3.0 * (a * a);
#pragma unroll
            for (int j = 0; j < NFSR; ++j) {
                g2[j] = 3.0 * (a * a);
                nz[j] = 3;
            }
        } else {
#pragma clang fp contract(off)
            float2 F[3][3], f;
            dyad_ref32_F(cells.x[go] - cells.x[gs], cells.y[go] - cells.y[gs], cells.z[go] - cells.z[gs], k0, pi,
                         cells.kappa[gs], z64{cells.che[2 * gs], cells.che[2 * gs + 1]}, F, f);
            const float axx = hypotf(F[0][0].x, F[0][0].y), axy = hypotf(F[0][1].x, F[0][1].y);
            const float axz = hypotf(F[0][2].x, F[0][2].y), ayy = hypotf(F[1][1].x, F[1][1].y);
            const float ayz = hypotf(F[1][2].x, F[1][2].y), azz = hypotf(F[2][2].x, F[2][2].y);
            const float af = hypotf(f.x, f.y);
            float t = axx * axx;
            t = t + 2.0f * (axy * axy);
            t = t + 2.0f * (axz * axz);
            t = t + ayy * ayy;
            t = t + 2.0f * (ayz * ayz);
            t = t + azz * azz;
            t = t * (af * af);
            fn2 = (double)t;
            for (int po = 0; po < 3; ++po)
                for (int ps = 0; ps < 3; ++ps) {
                    const z64 g = neg_fprod(F[po][ps], f);
                    const float a = hypotf((float)g.r, (float)g.i);
                    const float a2 = a * a;
#pragma unroll
                    for (int j = 0; j < NFSR; ++j)
                        if ((double)a > v0 / fl.f[j]) {
                            g2[j] += (double)a2;
                            ++nz[j];
                        }
                }
        }
    }
    // wave sums, then one atomic per value per wave
    auto wsum = [](double v) {
        for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
        return v;
    };
    fn2 = wsum(fn2);
    if (threadIdx.x == 0) atomicAdd(&out->fn2, fn2);
#pragma unroll
    for (int j = 0; j < NFSR; ++j) {
        double gs = wsum(g2[j]);
        double ns = wsum((double)nz[j]);
        if (threadIdx.x == 0) {
            atomicAdd(&out->g2[j], gs);
            atomicAdd(&out->nnz[j], (unsigned long long)ns);
        }
    }
}

// Zero the entries of a filled (m x m) block that the SR mask drops: off the
// self pairs, |g| (single precision) <= |1 - Znnpp(first cell)| / fsr.
__device__ __forceinline__ void sr_mask_entry(double2 *Z, long ld, int i, int j, const DevCells &cells, int g0,
                                              double fsr)
{
    const double thr = hypot(1.0 - cells.znnpp[2 * g0], -cells.znnpp[2 * g0 + 1]) / fsr;
    double2 *p = Z + (long)j * ld + i;
    const double2 v = *p;
    if (!((double)hypotf((float)v.x, (float)v.y) > thr)) *p = make_double2(0.0, 0.0);
}
// Every matrix of a grouped batch (blockIdx.z), each with its own fSR.
__global__ void k_sr_mask_group(const BlockJob *__restrict__ jobs, const int *__restrict__ idx, DevCells cells,
                                double2 *Z, long ld, long zstride)
{
    const BlockJob jb = jobs[blockIdx.z];
    const int m = 3 * jb.n;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= m || j >= m || i / 3 == j / 3) return;
    sr_mask_entry(Z + blockIdx.z * zstride, ld, i, j, cells, idx[jb.off], jb.fsr);
}

// Padding of a grouped batch of block systems in mp x mp slots (matrix q holds
// m = 3 jobs[q].n rows): zeros and a unit diagonal outside the block in Z, and
// zero rows past m in the ncx columns of X, so the padded unknowns solve to
// zero.  Columns [0, mp) of grid y are Z's, the next ncx are X's.
__global__ void k_pad_group(const BlockJob *__restrict__ jobs, double2 *Z, long ldz, long zstride, int mp,
                            double2 *X, long ldx, long xstride, int ncx)
{
    const int m = 3 * jobs[blockIdx.z].n;
    const int j = blockIdx.y;
    if (j < mp) {
        double2 *c = Z + blockIdx.z * zstride + (long)j * ldz;
        for (int i = (j < m ? m : 0) + threadIdx.x; i < mp; i += blockDim.x)
            c[i] = make_double2(i == j ? 1.0 : 0.0, 0.0);
    } else if (j - mp < ncx) {
        double2 *c = X + blockIdx.z * xstride + (long)(j - mp) * ldx;
        for (int i = m + threadIdx.x; i < mp; i += blockDim.x) c[i] = make_double2(0.0, 0.0);
    }
}

// Solutions of a grouped batch back to their blocks: rows [0, 3 n) of the
// ncol columns of matrix q go to X_b, b = jobs[q].dst.
__global__ void k_copy_out_group(const BlockJob *__restrict__ jobs, const double2 *__restrict__ S, long lds,
                                 long sstride, double2 *D, long ldd, long dstride)
{
    const BlockJob jb = jobs[blockIdx.z];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * jb.n) return;
    D[jb.dst * dstride + (long)blockIdx.y * ldd + i] = S[blockIdx.z * sstride + (long)blockIdx.y * lds + i];
}

}  // namespace midas

// E_dev: E is device memory (the fields stay in HBM) rather than host memory.
// Releases a LuWork when its scope ends.
template <typename T> struct LuRelease {
    LuWork<T> &w;
    ~LuRelease() { w.release(); }
};

// Records the launches of enqueue() on st into a hipGraph and launches it as
// one.  The executable graph of batch `slot` stays in the context: the next
// call whose batch records the same launches (other buffers and arguments)
// updates it in place instead of instantiating a new one.
template <typename F> static int run_as_graph(midas_ctx *ctx, size_t slot, F &&enqueue, hipStream_t st)
{
    MIDAS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int r = enqueue();
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (r || e != hipSuccess) {
        if (g) hipGraphDestroy(g);
        if (r) return r;
        MIDAS_HIP(e);
    }
    hipGraphExec_t &x = ctx->cbfm_graphs[slot];
    if (x) {
        hipGraphExecUpdateResult res = hipGraphExecUpdateError;
        hipGraphNode_t bad = nullptr;
        if (hipGraphExecUpdate(x, g, &bad, &res) != hipSuccess || res != hipGraphExecUpdateSuccess) {
            (void)hipGetLastError();
            hipGraphExecDestroy(x);
            x = nullptr;
        }
    }
    hipError_t ie = hipSuccess;
    if (!x) ie = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (ie != hipSuccess) x = nullptr;
    MIDAS_HIP(ie);
    MIDAS_HIP(hipGraphLaunch(x, st));
    return 0;
}

static int compute_efields_cbfm(midas_ctx *ctx, const midas_cells_t *cells, const midas_blocks_t *blocks, double k0,
                                int quirks, int ntr_cbf, const double *th_cbf, const double *ph_cbf, int iws, int ntr,
                                const double *tr_th, const double *tr_ph, int use_aca, int kmax, double eps_aca,
                                double threshold, int mode, double *E, long lde, bool E_dev, int *K_out,
                                double *C_job, long ldc, int kcap, int sr, double res_sr, double *fsr_out,
                                long *nnz_out, midas_solve_info_t *info)
{
    MIDAS_CHECK(ctx && cells && blocks && E && th_cbf && ph_cbf && tr_th && tr_ph, -1, "cbfm: bad arguments");
    const int N = cells->n, nb = blocks->nblocks;
    MIDAS_CHECK(lde >= 3L * N, -1, "cbfm: lde %ld < 3n", lde);
    MIDAS_CHECK(!C_job || (ldc >= 3L * N && kcap > 0), -1, "cbfm: C_job needs ldc >= 3n and kcap > 0");
    MIDAS_CHECK(N > 0 && nb > 0 && blocks->block_start && blocks->ext_start, -1, "cbfm: empty cells or blocks");
    MIDAS_CHECK(ntr_cbf > 0 && ntr > 0, -1, "cbfm: ntr_cbf %d, ntr %d", ntr_cbf, ntr);
    MIDAS_CHECK(mode >= 0 && mode <= 5, -1, "cbfm: unknown solver mode %d", mode);
    MIDAS_CHECK(iws >= 1 && iws <= 3, -1, "cbfm: CBF excitation type %d (1 plane, 2/3 spherical waves)", iws);
    if (use_aca)
        MIDAS_CHECK(kmax >= 1 && kmax <= ACA_KMAX_SUPPORTED && eps_aca > 0, -1, "cbfm: kmax %d (1..%d), eps %g", kmax,
                    ACA_KMAX_SUPPORTED, eps_aca);
    const int *bs = blocks->block_start, *es = blocks->ext_start;
    MIDAS_CHECK(bs[0] == 0 && bs[nb] == N && es[0] == 0, -1, "cbfm: block_start must run 0..n, ext_start from 0");
    for (int b = 0; b < nb; ++b) {
        MIDAS_CHECK(bs[b + 1] > bs[b] && es[b + 1] >= es[b], -1, "cbfm: block %d empty or tables not increasing", b);
        for (int e = es[b]; e < es[b + 1]; ++e) {
            const int c = blocks->ext_cells[e];
            MIDAS_CHECK(c >= 0 && c < N && (c < bs[b] || c >= bs[b + 1]), -1,
                        "cbfm: extension cell %d of block %d outside the cells or inside the block", c, b);
        }
    }
    MIDAS_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->st;
    const bool ref32 = (quirks & MIDAS_Q_FP32_BLOCKS) != 0;
    MIDAS_CHECK(threshold > 0 || cells->eps, -1, "cbfm: the threshold rule needs eps");
    const double thr = threshold > 0 ? threshold : cbf_threshold_rule(cells, use_aca);
    const long N3 = 3L * N;
    const int T2c = 2 * ntr_cbf, T2 = 2 * ntr;
    // batching budgets (midas_set_cbfm_budgets; 0 = default)
    auto pick_budget = [&](int i, double dflt) { return ctx->cbfm_budget[i] > 0 ? ctx->cbfm_budget[i] : dflt; };
    const double lu_budget = pick_budget(0, 24e9), strip_budget = pick_budget(1, (double)(8L << 30)),
                 dense_budget = pick_budget(2, (double)(4L << 30)), aca_budget = pick_budget(3, 4.0e9);

    // MIDAS_VERBOSE=1 prints the time of each phase (synchronising between them)
    const bool verbose = std::getenv("MIDAS_VERBOSE") && std::atoi(std::getenv("MIDAS_VERBOSE")) > 0;
    std::unique_ptr<Timer> tm(verbose ? new Timer(st) : nullptr);
    auto phase = [&](const char *name) {
        if (tm) std::fprintf(stderr, "[cbfm] %-28s %10.2f ms\n", name, tm->ms());
    };
    // every device buffer of the call is the context's, grown when needed
    DevCellsOwned &dc = ctx->call_cells;
    TRY(dc.from_host(cells, k0, quirks, st));
    DBuf &dthc = pooled(ctx, "thc"), &dphc = pooled(ctx, "phc"), &dtt = pooled(ctx, "tt"), &dtp = pooled(ctx, "tp");
    TRY(upload_keep(dthc, th_cbf, sizeof(double) * ntr_cbf, st));
    TRY(upload_keep(dphc, ph_cbf, sizeof(double) * ntr_cbf, st));
    TRY(upload_keep(dtt, tr_th, sizeof(double) * ntr, st));
    TRY(upload_keep(dtp, tr_ph, sizeof(double) * ntr, st));

    // ---- extended cell lists of every block (block cells first)
    std::vector<int> nbk(nb), sb(nb), ioff(nb + 1, 0);
    std::vector<long> coff(nb + 1, 0);
    int smax = 0, nmax = 0;
    for (int b = 0; b < nb; ++b) {
        nbk[b] = bs[b + 1] - bs[b];
        sb[b] = nbk[b] + es[b + 1] - es[b];
        ioff[b + 1] = ioff[b] + sb[b];
        coff[b + 1] = coff[b] + 3L * nbk[b] * std::min(3 * sb[b], T2c);
        smax = std::max(smax, sb[b]);
        nmax = std::max(nmax, nbk[b]);
    }
    std::vector<int> allidx(ioff[nb]);
    for (int b = 0; b < nb; ++b) {
        for (int i = 0; i < nbk[b]; ++i) allidx[ioff[b] + i] = bs[b] + i;
        for (int e = es[b]; e < es[b + 1]; ++e) allidx[ioff[b] + nbk[b] + (e - es[b])] = blocks->ext_cells[e];
    }
    DBuf &didx = pooled(ctx, "idx");
    TRY(upload_keep(didx, allidx.data(), sizeof(int) * allidx.size(), st));

    // ---- 1. characteristic basis functions
    // 1a. block solutions X_b = Z_b^-1 E_b (FP64 LU).  All blocks are solved
    // as one batch: each Z_b is embedded in an mp x mp matrix (mp = 3 max s_b)
    // with a unit diagonal on the padding and E_b padded with zero rows, so the
    // padded solution rows are zero and every LU kernel launch serves the whole
    // batch.  X_b is kept at stride mp * 2T_c, leading dimension mp.
    const long mmax = 3L * smax;
    const long mp = mmax;
    int rpad = 1;
    for (int b = 0; b < nb; ++b) rpad = std::max(rpad, std::min(3 * sb[b], T2c));
    DBuf &Call = pooled(ctx, "C"), &Xall = pooled(ctx, "X");
    TRY(Call.reserve(sizeof(double2) * std::max<long>(coff[nb], 1)));
    TRY(Xall.reserve(sizeof(double2) * mp * T2c * nb));
    auto Xb = [&](int b) { return Xall.as<double2>() + (long)b * mp * T2c; };
    int worst_info = 0;
    // SR: the fSR of every block from one pass of norms (setfSR)
    std::vector<double> fsr(nb, 0.0);
    if (sr) {
        DBuf &stats = pooled(ctx, "srstats");
        TRY(stats.reserve(sizeof(SrStats) * nb));
        MIDAS_HIP(hipMemsetAsync(stats.p, 0, sizeof(SrStats) * nb, st));
        for (int b = 0; b < nb; ++b) {
            CellSet S{sb[b], 0, didx.as<int>() + ioff[b]};
            hipLaunchKernelGGL(k_sr_stats, dim3((sb[b] + 63) / 64, sb[b]), dim3(64), 0, st, dc.v, S, k0,
                               host_pi(quirks), FSR_TESTS, stats.as<SrStats>() + b);
        }
        MIDAS_HIP(hipGetLastError());
        std::vector<SrStats> hs(nb);
        TRY(copy_sync(hs.data(), stats.p, sizeof(SrStats) * nb, hipMemcpyDeviceToHost, st));
        for (int b = 0; b < nb; ++b) {
            const double fn = std::sqrt(hs[b].fn2);
            double res = 1e3;
            int j = 0;
            unsigned long long nz = 0;
            while (j < NFSR && res > res_sr) {
                fsr[b] = FSR_TESTS.f[j];
                nz = hs[b].nnz[j];
                res = std::fabs(fn - std::sqrt(hs[b].g2[j])) / fn * 100.0;
                ++j;
            }
            if (fsr_out) fsr_out[b] = fsr[b];
            if (nnz_out) nnz_out[b] = (long)nz;
        }
    }
    {
        // Blocks sorted by extended size (largest first) and solved in batches,
        // each padded only to its own largest block (uneven blocks would
        // otherwise all pay for the largest), within the LU budget (24 GB unless set) for the
        // batch's matrices.  Solutions are staged per batch (the batched solve
        // needs one stride) and copied to X_b.
        struct Chunk { int pos, nc; long mpc; };
        std::vector<int> order(nb);
        for (int b = 0; b < nb; ++b) order[b] = b;
        std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return sb[a] > sb[c]; });
        std::vector<Chunk> chunks;
        long zmax = 0, smax_c = 0, tmax = 0, xmax = 0, pmax_c = 0;
        const int nb_lu = tune(TUNE_LU_NB_FP64);
        // the block systems: FP64 LU, or a mixed-precision LU + FP64 refinement
        // (single-term split modes and c32: the batched trailing updates).
        // Each refinement sweep costs a residual of 8 m^2 T2c FP64 flops, so
        // the mixed LU (about 5x cheaper than the FP64 one, 8/3 m^3) pays
        // only when the blocks are much taller than the excitation count:
        // mode 5 takes f16 for m >= 16 T2c (C5: m / T2c ~ 18.5 gained 7 % of
        // the block phase; a0006 at ~9 lost 60 %), else the FP64 LU.
        const int bm_set = tune(TUNE_CBFM_BLOCK_MODE);
        const int nb_mixed = 256;
        BlockMixedWs &bmw = ctx->cbfm_mixed;
        for (int pos = 0; pos < nb;) {
            const long mpc = 3L * sb[order[pos]];
            const long zb = (long)sizeof(double2) * mpc * mpc;
            const int nc = (int)std::max(1L, std::min((long)(nb - pos), (long)(lu_budget / (double)zb)));
            chunks.push_back({pos, nc, mpc});
            zmax = std::max(zmax, zb * nc);
            pmax_c = std::max(pmax_c, mpc * nc);
            tmax = std::max(tmax, getrs_tmp_elems((int)mpc, nb_lu, T2c) * nc);
            xmax = std::max(xmax, mpc * T2c * nc);
            smax_c = std::max(smax_c, (long)nc);
            pos += nc;
        }
        DBuf &Zall = pooled(ctx, "Zb"), &ipiv = pooled(ctx, "ipiv"), &perm = pooled(ctx, "perm"),
             &tmp = pooled(ctx, "getrs_tmp"), &Xc = pooled(ctx, "Xc");
        TRY(Zall.reserve(zmax));
        TRY(ipiv.reserve(sizeof(int) * pmax_c));
        TRY(perm.reserve(sizeof(int) * pmax_c));
        TRY(tmp.reserve(sizeof(double2) * tmax));
        TRY(Xc.reserve(sizeof(double2) * xmax));
        MIDAS_HIP(hipMemsetAsync(Xall.p, 0, sizeof(double2) * mp * T2c * nb, st));
        std::vector<int> hinfo(smax_c);
        DBuf &djob = pooled(ctx, "jobs"), &l21f = pooled(ctx, "l21");
        TRY(djob.reserve(sizeof(BlockJob) * smax_c));
        const BlockJob *jd = djob.as<BlockJob>();
        const int *idx = didx.as<int>();
        const bool use_graph = tune(TUNE_CBFM_GRAPH) != 0;
        if (ctx->cbfm_graphs.size() < chunks.size()) ctx->cbfm_graphs.resize(chunks.size(), nullptr);
        for (size_t ci = 0; ci < chunks.size(); ++ci) {
            const Chunk &ch = chunks[ci];
            const int nc = ch.nc;
            const long mpc = ch.mpc;
            MIDAS_CHECK(mpc + T2c <= 65535, -1, "cbfm: block order %ld + %d excitations too large for one launch", mpc,
                        T2c);
            // the batch's jobs: cells, SR factor, spherical-wave frame
            std::vector<BlockJob> hj(nc);
            int ncell = 0;
            for (int q = 0; q < nc; ++q) {
                const int b = order[ch.pos + q];
                BlockJob &j = hj[q];
                j = BlockJob{sb[b], ioff[b], b, 0, sr ? fsr[b] : 1.0, 0.0, 0.0, 0.0, 0.0};
                if (iws != 1) {
                    // spherical waves around the extended block (type_iws 2 and 3 both call
                    // Incident_Field_Spherical, Compute_EFields_CBFME.f90:363-369)
                    double cen[3];
                    sph_source_frame(cells->x, cells->y, cells->z, allidx.data() + ioff[b], sb[b], k0, quirks, cen,
                                     &j.rs);
                    j.cx = cen[0];
                    j.cy = cen[1];
                    j.cz = cen[2];
                }
                ncell = std::max(ncell, sb[b]);
            }
            TRY(copy_sync(djob.p, hj.data(), sizeof(BlockJob) * nc, hipMemcpyHostToDevice, st));
            const int block_mode = bm_set == 5 ? (mpc >= 16L * T2c ? MIDAS_SOLVE_F16_IR : MIDAS_SOLVE_FP64)
                                   : bm_set == 3 ? MIDAS_SOLVE_F16_IR : bm_set;
            const bool fp64 = block_mode == MIDAS_SOLVE_FP64;
            LuWork<double> &w = ctx->cbfm_lu;
            TrailingGemm<double> tg;
            if (fp64) {
                w.share_stream(ctx->ps, ctx->pev[0], ctx->pev[1]);
                TRY(w.alloc((int)mpc, nb_lu, nc, mpc * mpc));
                w.defer_fault_check = true;   // checked after the batch (no synchronisation inside a graph)
                tg.batch = nc;
                tg.stride = mpc * mpc;
                // pivot search in the panels' diagonal blocks, kept only where
                // certified (max |L21| <= 1/sqrt 2: partial pivoting's pivots);
                // otherwise the batch is built and factored again below
                w.diag_pivot = tune(TUNE_LU_PIVOT_SEARCH) >= 1;
                if (w.diag_pivot) {
                    const size_t sl = sizeof(double2) * (size_t)mpc * nb_lu * nc;
                    TRY(l21f.reserve(sl + 256));
                    w.l21 = l21f.as<z64>();
                    w.lmax = (unsigned *)((char *)l21f.p + sl);
                }
            }
            double2 *Z = Zall.as<double2>(), *X = Xc.as<double2>();
            // One batch: grouped block fills, SR masks, padding, excitations;
            // with the FP64 LU also the batched factorisation, its solve and
            // the copy of the solutions to their blocks.
            auto enqueue = [&]() -> int {
                TRY(launch_fill_group(dc.v, jd, idx, nc, ncell, k0, quirks, ref32, Z, mpc, mpc * mpc, st));
                if (sr)
                    hipLaunchKernelGGL(k_sr_mask_group, dim3((3 * ncell + 255) / 256, 3 * ncell, nc), dim3(256), 0, st,
                                       jd, idx, dc.v, Z, mpc, mpc * mpc);
                hipLaunchKernelGGL(k_pad_group, dim3(1, (int)mpc + T2c, nc), dim3(256), 0, st, jd, Z, mpc, mpc * mpc,
                                   (int)mpc, X, mpc, mpc * T2c, T2c);
                MIDAS_HIP(hipGetLastError());
                TRY(launch_rhs_group(dc.v, jd, idx, nc, ncell, ntr_cbf, dthc.as<double>(), dphc.as<double>(), k0,
                                     quirks, ref32, iws != 1, X, mpc, mpc * T2c, st));
                if (!fp64) return 0;
                TRY(getrf<double>(Z, mpc, (int)mpc, ipiv.as<int>(), w, st, tg));
                TRY(ipiv_to_perm_dev(ipiv.as<int>(), perm.as<int>(), (int)mpc, nc, st));
                TRY(getrs<double>(Z, mpc, (int)mpc, perm.as<int>(), w, X, mpc, T2c, tmp.as<double2>(), st, mpc * T2c));
                hipLaunchKernelGGL(k_copy_out_group, dim3((int)((mpc + 255) / 256), T2c, nc), dim3(256), 0, st, jd, X,
                                   mpc, mpc * T2c, Xall.as<double2>(), mp, mp * T2c);
                MIDAS_HIP(hipGetLastError());
                return 0;
            };
            if (use_graph)
                TRY(run_as_graph(ctx, ci, enqueue, st));
            else
                TRY(enqueue());
            if (fp64) {
                TRY(lu_fault_check<double>(w, st));
                if (w.diag_pivot) {
                    unsigned lm = 0;
                    TRY(copy_sync(&lm, w.lmax, sizeof(unsigned), hipMemcpyDeviceToHost, st));
                    if (!((double)__builtin_bit_cast(float, lm) <= 0.70710678)) {
                        // not certified: the batch again, with the whole-column search
                        w.diag_pivot = false;
                        TRY(enqueue());
                        TRY(lu_fault_check<double>(w, st));
                    }
                }
                TRY(copy_sync(hinfo.data(), w.info, sizeof(int) * nc, hipMemcpyDeviceToHost, st));
            } else {
                TRY(block_solve_mixed(ctx, block_mode, nc, (int)mpc, T2c, Z, X, ipiv.as<int>(), perm.as<int>(), hinfo,
                                      bmw, nb_mixed, nb_lu, st));
                hipLaunchKernelGGL(k_copy_out_group, dim3((int)((mpc + 255) / 256), T2c, nc), dim3(256), 0, st, jd, X,
                                   mpc, mpc * T2c, Xall.as<double2>(), mp, mp * T2c);
                MIDAS_HIP(hipGetLastError());
            }
            for (int q = 0; q < nc; ++q)
                if (hinfo[q] > 0 && worst_info == 0) worst_info = hinfo[q];
        }
    }
    phase(sr ? "SR masks + block fills + LU solves" : "block fills + LU solves");
    // 1b. left singular vectors of X_b with sigma_d / sigma_1 >= threshold.
    std::vector<int> K(nb);
    if (thr >= CBF_GRAM_MIN_THRESHOLD) {
        // Gram route: G_b = X^H X (T2c <= m_b, then U = X W S^-1) or X X^H
        // (T2c > m_b, then U = W); every block's G zero-padded to rpad,
        // reduced to its numerical rank r_b by pivoted Cholesky G = L L^H,
        // and B = L^H L (r x r, r = max r_b, zero-padded) decomposed by one
        // launch of the batched Hermitian eigensolver (midas_heev.hip):
        // B w = lambda w gives G (L w) = lambda (L w), |L w|^2 = lambda.
        const long gs = (long)rpad * rpad;
        DBuf &G = pooled(ctx, "gram"), &Lc = pooled(ctx, "chol"), &dr = pooled(ctx, "rank"),
             &dinfo = pooled(ctx, "info"), &Wk = pooled(ctx, "wk"), &Wr = pooled(ctx, "wr"), &HW = pooled(ctx, "heev"),
             &Bm = pooled(ctx, "lhl"), &D = pooled(ctx, "eig");
        TRY(G.reserve(sizeof(double2) * gs * nb));
        TRY(Lc.reserve(sizeof(double2) * gs * nb));
        TRY(dr.reserve(sizeof(int) * nb));
        TRY(dinfo.reserve(sizeof(int) * nb));
        TRY(Wk.reserve(sizeof(double2) * (size_t)T2c * T2c));
        MIDAS_HIP(hipMemsetAsync(G.p, 0, sizeof(double2) * gs * nb, st));
        bool all_tall = true;
        for (int b = 0; b < nb; ++b) all_tall = all_tall && T2c <= 3 * sb[b];
        if (all_tall) {
            // X_b^H X_b of every block in one batched launch: the rows of X_b
            // past 3 s_b are zero, so the sum over all mp rows adds exact zeros
            TRY(launch_cgemm<double>(T2c, T2c, (int)mp, ONE, Xb(0), mp, 2, Xb(0), mp, 0, ZERO, G.p, rpad, st, nb,
                                     mp * T2c, mp * T2c, gs));
        } else {
            for (int b = 0; b < nb; ++b) {
                const int m = 3 * sb[b];
                double2 *Gb = G.as<double2>() + b * gs;
                if (T2c <= m)
                    TRY(launch_cgemm<double>(T2c, T2c, m, ONE, Xb(b), mp, 2, Xb(b), mp, 0, ZERO, Gb, rpad, st));
                else
                    TRY(launch_cgemm<double>(m, m, T2c, ONE, Xb(b), mp, 0, Xb(b), mp, 2, ZERO, Gb, rpad, st));
            }
        }
        TRY(launch_pchol_batched(rpad, G.as<double2>(), rpad, gs, Lc.as<double2>(), rpad, gs, dr.as<int>(), nb,
                                 CBF_GRAM_RANK_TOL, st));
        std::vector<int> rk(nb);
        TRY(copy_sync(rk.data(), dr.p, sizeof(int) * nb, hipMemcpyDeviceToHost, st));
        const int rr = *std::max_element(rk.begin(), rk.end());
        std::vector<double> hd((size_t)std::max(rr, 1) * nb, 0.0);
        if (rr > 0) {
            const long bs2 = (long)rr * rr;
            TRY(Bm.reserve(sizeof(double2) * bs2 * nb));
            TRY(D.reserve(sizeof(double) * (size_t)rr * nb));
            TRY(HW.reserve(heev_work_bytes(rr, nb)));
            TRY(Wr.reserve(sizeof(double2) * (size_t)rr * T2c));
            TRY(launch_cgemm<double>(rr, rr, rpad, ONE, Lc.p, rpad, 2, Lc.p, rpad, 0, ZERO, Bm.p, rr, st, nb, gs, gs,
                                     bs2));
            // (large orders compute only the pairs at or above a quarter of
            // the threshold's eigenvalue ratio, a margin over the count below)
            TRY(launch_heev_batched(rr, Bm.as<double2>(), rr, bs2, D.as<double>(), rr, dinfo.as<int>(), nb, HW.p, st,
                                    0.25 * thr * thr));
            std::vector<int> hi(nb);
            MIDAS_HIP(hipMemcpyAsync(hi.data(), dinfo.p, sizeof(int) * nb, hipMemcpyDeviceToHost, st));
            TRY(copy_sync(hd.data(), D.p, sizeof(double) * hd.size(), hipMemcpyDeviceToHost, st));
            for (int b = 0; b < nb; ++b)
                MIDAS_CHECK(hi[b] == 0, -2, "cbfm: eigensolver failed on block %d (info %d)", b, hi[b]);
        }
        for (int b = 0; b < nb; ++b) {
            const int m = 3 * sb[b], r = std::min(std::min(m, T2c), rr);
            const double *lam = hd.data() + (size_t)b * std::max(rr, 1);   // ascending
            const double s1 = rr > 0 ? std::sqrt(std::max(lam[rr - 1], 0.0)) : 0.0;
            int kb = 0;
            for (int d = 0; d < r; ++d)
                if (s1 > 0 && std::sqrt(std::max(lam[rr - 1 - d], 0.0)) / s1 >= thr) ++kb;
            K[b] = kb;
            if (!kb) continue;
            const double2 *Wb = Bm.as<double2>() + (long)b * rr * rr;
            const double *Db = D.as<double>() + (size_t)b * rr;
            const double2 *Lb = Lc.as<double2>() + b * gs;
            if (T2c <= m) {
                // W_G S^-1 = L W Lambda^-1, then U = X (W_G S^-1)
                TRY(launch_pick_eigvecs(Wb, rr, Db, rr, rr, kb, 2, Wr.as<double2>(), rr, st));
                TRY(launch_cgemm<double>(T2c, kb, rr, ONE, Lb, rpad, 0, Wr.p, rr, 0, ZERO, Wk.p, T2c, st));
                TRY(launch_cgemm<double>(3 * nbk[b], kb, T2c, ONE, Xb(b), mp, 0, Wk.p, T2c, 0, ZERO,
                                         Call.as<double2>() + coff[b], 3L * nbk[b], st));
            } else {
                // U = L W Lambda^-1/2 (rows of the block's own cells)
                TRY(launch_pick_eigvecs(Wb, rr, Db, rr, rr, kb, 1, Wr.as<double2>(), rr, st));
                TRY(launch_cgemm<double>(3 * nbk[b], kb, rr, ONE, Lb, rpad, 0, Wr.p, rr, 0, ZERO,
                                         Call.as<double2>() + coff[b], 3L * nbk[b], st));
            }
        }
    } else {
        // thresholds below what the Gram route resolves: one-by-one SVD
        // (rocSOLVER gesvd; not on the default path)
        RocHandle rh;
        MIDAS_ROC(rocblas_create_handle(&rh.h));
        MIDAS_ROC(rocblas_set_stream(rh.h, st));
        DBuf Usv, Sv, Esup, dinfo;
        const long rmax = std::min<long>(mmax, T2c);
        TRY(Usv.alloc(sizeof(double2) * mmax * rmax));
        TRY(Sv.alloc(sizeof(double) * rmax));
        TRY(Esup.alloc(sizeof(double) * rmax));
        TRY(dinfo.alloc(sizeof(int)));
        std::vector<double> hs(rmax);
        for (int b = 0; b < nb; ++b) {
            const int m = 3 * sb[b], r = std::min(m, T2c);
            MIDAS_ROC(rocsolver_zgesvd(rh.h, rocblas_svect_singular, rocblas_svect_none, m, T2c,
                                       (rocblas_double_complex *)Xb(b), (int)mp, Sv.as<double>(),
                                       (rocblas_double_complex *)Usv.p, m, nullptr, 1, Esup.as<double>(),
                                       rocblas_outofplace, dinfo.as<int>()));
            int sinfo = 0;
            MIDAS_HIP(hipMemcpyAsync(&sinfo, dinfo.p, sizeof(int), hipMemcpyDeviceToHost, st));
            TRY(copy_sync(hs.data(), Sv.p, sizeof(double) * r, hipMemcpyDeviceToHost, st));
            MIDAS_CHECK(sinfo == 0, -2, "cbfm: SVD of block %d did not converge (info %d)", b, sinfo);
            int kb = 0;
            for (int d = 0; d < r; ++d)
                if (hs[0] > 0 && hs[d] / hs[0] >= thr) ++kb;
            K[b] = kb;
            TRY(copy2d(Usv.as<double2>(), m, Call.as<double2>() + coff[b], 3L * nbk[b], 3 * nbk[b], kb, st));
        }
    }
    phase("CBF singular subspaces");
    std::vector<long> offs(nb + 1, 0);
    int kbmax = 0;
    for (int b = 0; b < nb; ++b) {
        offs[b + 1] = offs[b] + K[b];
        kbmax = std::max(kbmax, K[b]);
    }
    const long Kt = offs[nb];
    MIDAS_CHECK(Kt > 0, -3, "cbfm: no basis function passed the threshold %g", thr);
    auto Cb = [&](int b) { return Call.as<double2>() + coff[b]; };

    // ---- 2. reduced matrix
    DBuf &Zc = pooled(ctx, "Zc");
    TRY(Zc.reserve(sizeof(double2) * (size_t)Kt * Kt));
    double2 *zc_ = Zc.as<double2>();
    const CellSet ALL{N, 0, nullptr};
    DBuf &djobs = pooled(ctx, "gjobs");
    std::vector<GJob> jobs;
    if (!use_aca) {
        // T(:, block j) = Z(:, block j) C_j for a chunk of block columns (one
        // strip fill and one GEMM each), then every C_i^T T(rows of i, j) as
        // one grouped launch
        const long tbudget = (long)strip_budget;
        const int tcols = (int)std::max<long>(kbmax, std::min<long>(Kt, tbudget / ((long)sizeof(double2) * N3)));
        DBuf &Zs = pooled(ctx, "strip"), &Tt = pooled(ctx, "strip_t");
        TRY(Zs.reserve(sizeof(double2) * N3 * 3 * nmax));
        TRY(Tt.reserve(sizeof(double2) * N3 * std::max(tcols, 1)));
        int j0 = 0;
        while (j0 < nb) {
            int j1 = j0, used = 0;
            std::vector<int> tcol(nb, 0);
            while (j1 < nb && (j1 == j0 || used + K[j1] <= tcols)) {
                tcol[j1] = used;
                used += K[j1];
                ++j1;
            }
            jobs.clear();
            for (int j = j0; j < j1; ++j) {
                if (!K[j]) continue;
                TRY(launch_fill(dc.v, ALL, CellSet{nbk[j], bs[j], nullptr}, k0, quirks, ref32, Zs.as<double2>(), N3,
                                nullptr, 0, st));
                double2 *Tj = Tt.as<double2>() + (long)tcol[j] * N3;
                TRY(launch_cgemm<double>(N3, K[j], 3 * nbk[j], ONE, Zs.p, N3, 0, Cb(j), 3L * nbk[j], 0, ZERO, Tj, N3,
                                         st));
                for (int i = 0; i < nb; ++i)
                    if (K[i])
                        add_gemm_tiles(jobs, Cb(i), 3 * nbk[i], 1, Tj + 3L * bs[i], (int)N3,
                                       zc_ + offs[i] + offs[j] * Kt, (int)Kt, K[i], K[j], 3 * nbk[i]);
            }
            TRY(run_ggemm(jobs, djobs, st));
            j0 = j1;
        }
    } else {
        // dense products C_i^T Z_ij C_j for the diagonal and for ACA fallbacks:
        // the blocks Z_ij are filled side by side, then M = C_i^T Z_ij and
        // Zc_ij = M C_j for all of them in two grouped launches
        std::vector<std::pair<int, int>> dense;
        // reused across batches (allocating multi-GB buffers per batch costs
        // milliseconds each and serialises the stream)
        DBuf &Zd = pooled(ctx, "dense_z"), &Md = pooled(ctx, "dense_m");
        auto dense_pairs = [&](hipStream_t sp) -> int {
            const long zbudget = (long)dense_budget;
            size_t q0 = 0;
            while (q0 < dense.size()) {
                size_t q1 = q0;
                long zused = 0, mused = 0;
                std::vector<long> zo, mo;
                while (q1 < dense.size()) {
                    const int i = dense[q1].first, j = dense[q1].second;
                    const long zb = 9L * nbk[i] * nbk[j];
                    if (q1 > q0 && (zused + zb) * (long)sizeof(double2) > zbudget) break;
                    zo.push_back(zused);
                    mo.push_back(mused);
                    zused += zb;
                    mused += (long)K[i] * 3 * nbk[j];
                    ++q1;
                }
                TRY(Zd.reserve(sizeof(double2) * zused));
                TRY(Md.reserve(sizeof(double2) * std::max(mused, 1L)));
                jobs.clear();
                for (size_t q = q0; q < q1; ++q) {
                    const int i = dense[q].first, j = dense[q].second;
                    double2 *Zq = Zd.as<double2>() + zo[q - q0];
                    TRY(launch_fill(dc.v, CellSet{nbk[i], bs[i], nullptr}, CellSet{nbk[j], bs[j], nullptr}, k0, quirks,
                                    ref32, Zq, 3L * nbk[i], nullptr, 0, sp));
                    add_gemm_tiles(jobs, Cb(i), 3 * nbk[i], 1, Zq, 3 * nbk[i], Md.as<double2>() + mo[q - q0], K[i],
                                   K[i], 3 * nbk[j], 3 * nbk[i]);
                }
                TRY(run_ggemm(jobs, djobs, sp));
                jobs.clear();
                for (size_t q = q0; q < q1; ++q) {
                    const int i = dense[q].first, j = dense[q].second;
                    add_gemm_tiles(jobs, Md.as<double2>() + mo[q - q0], K[i], 0, Cb(j), 3 * nbk[j],
                                   zc_ + offs[i] + offs[j] * Kt, (int)Kt, K[i], K[j], 3 * nbk[j]);
                }
                TRY(run_ggemm(jobs, djobs, sp));
                q0 = q1;
            }
            dense.clear();
            return 0;
        };
        for (int j = 0; j < nb; ++j)
            if (K[j]) dense.push_back({j, j});
        TRY(dense_pairs(st));
        // off-diagonal pairs by ACA, a batch of block columns per launch; then
        // Zc_ij = (C_i^T U)(V C_j) through two grouped launches.  Two batches
        // are in flight: the ACA of batch b runs on the context stream while
        // the products of batch b - 1 run on a second stream (their buffers
        // alternate), so the latency-bound ACA and the MFMA products overlap.
        const double budget = aca_budget;   // bytes of U, Vt and the small factors per batch
        struct AcaBatch {
            std::vector<AcaJob> jobs;
            std::vector<int> pi, pj;
            long uo = 0, vo = 0;
        };
        std::vector<AcaBatch> batches;
        for (int j0 = 0, j1; j0 < nb; j0 = j1) {
            AcaBatch bt;
            j1 = j0;
            while (j1 < nb) {
                long add_u = 0, add_v = 0;
                std::vector<AcaJob> col;
                std::vector<int> ci;
                for (int i = 0; i < nb; ++i) {
                    if (i == j1 || !K[i] || !K[j1]) continue;
                    AcaJob jb{3 * nbk[i], 3 * nbk[j1], bs[i], bs[j1], bt.uo + add_u, bt.vo + add_v};
                    add_u += 3L * nbk[i] * kmax;
                    add_v += 3L * nbk[j1] * kmax;
                    col.push_back(jb);
                    ci.push_back(i);
                }
                if (j1 > j0 && (double)(bt.uo + add_u + bt.vo + add_v) * sizeof(double2) > budget) break;
                bt.jobs.insert(bt.jobs.end(), col.begin(), col.end());
                for (int i : ci) {
                    bt.pi.push_back(i);
                    bt.pj.push_back(j1);
                }
                bt.uo += add_u;
                bt.vo += add_v;
                ++j1;
            }
            if (!bt.jobs.empty()) batches.push_back(std::move(bt));
        }
        // buffers of the two parities, sized once for the largest batch (a
        // buffer that grew while a kernel of the other stream read it would be
        // freed under it)
        size_t nj_max = 1;
        long uo_max = 1, vo_max = 1;
        for (const AcaBatch &bt : batches) {
            nj_max = std::max(nj_max, bt.jobs.size());
            uo_max = std::max(uo_max, bt.uo);
            vo_max = std::max(vo_max, bt.vo);
        }
        // the second stream and its events are the context's (made once)
        struct Pipe {
            hipStream_t s2 = nullptr;
            hipEvent_t ev[5] = {};   // start, rank[2], products[2]
            int *hk[2] = {};         // pinned (the context's): the ranks of a batch
            ~Pipe()
            {
                if (s2) hipStreamSynchronize(s2);
            }
        } pp;
        AcaBufs &ab = ctx->cbfm_aca;
        DBuf(&dj)[2] = ab.dj, (&Ub)[2] = ab.Ub, (&Vb)[2] = ab.Vb, (&dk)[2] = ab.dk, (&Gb)[2] = ab.Gb,
            (&LR)[2] = ab.LR, (&dgj)[2] = ab.dgj;
        if (!batches.empty()) {
            pp.s2 = ctx->s2;
            for (int q = 0; q < 5; ++q) pp.ev[q] = ctx->aev[q];
            TRY(ab.reserve_ranks(nj_max));
            for (int q = 0; q < 2; ++q) {
                pp.hk[q] = ab.hk[q];
                TRY(dj[q].reserve(sizeof(AcaJob) * nj_max));
                TRY(Ub[q].reserve(sizeof(double2) * uo_max));
                TRY(Vb[q].reserve(sizeof(double2) * vo_max));
                TRY(dk[q].reserve(sizeof(int) * nj_max));
                TRY(Gb[q].reserve(aca_gram_bytes(kmax, (int)nj_max)));
            }
            // the second stream starts after the diagonal products
            MIDAS_HIP(hipEventRecord(pp.ev[0], st));
            MIDAS_HIP(hipStreamWaitEvent(pp.s2, pp.ev[0], 0));
        }
        size_t prof_at[2] = {0, 0};
        // the products of batch b (parity q) on the second stream
        auto products = [&](int b) -> int {
            const int q = b & 1;
            const AcaBatch &bt = batches[b];
            MIDAS_HIP(hipEventSynchronize(pp.ev[1 + q]));
            const int *hk = pp.hk[q];
            if (profile().recs.size() > prof_at[q]) {
                // algorithmic bytes of the launch, known once the ranks are:
                // rank step s of an m x n pair reads the s - 1 previous
                // columns of U and rows of V and writes the new pair,
                // ~16 (m + n) (s + 2) bytes (complex FP64)
                double by = 0.0;
                for (size_t r = 0; r < bt.jobs.size(); ++r) {
                    const double steps = std::min(hk[r], kmax);
                    by += 16.0 * (bt.jobs[r].m + bt.jobs[r].n) * (steps * (steps + 1) / 2 + 2 * steps);
                }
                profile().recs[prof_at[q]].bytes = by;
            }
            // L_ij = C_i^T U (K_i x k) and R_ij = V C_j (k x K_j) in one scratch
            std::vector<long> lo(bt.jobs.size()), ro(bt.jobs.size());
            long lr = 0;
            for (size_t r = 0; r < bt.jobs.size(); ++r) {
                const int kk = std::min(hk[r], kmax);
                if (kk >= kmax) continue;
                lo[r] = lr;
                lr += (long)K[bt.pi[r]] * kk;
                ro[r] = lr;
                lr += (long)kk * K[bt.pj[r]];
            }
            TRY(LR[q].reserve(sizeof(double2) * std::max(lr, 1L)));
            jobs.clear();
            for (size_t r = 0; r < bt.jobs.size(); ++r) {
                const int i = bt.pi[r], j = bt.pj[r], kk = std::min(hk[r], kmax);
                if (kk >= kmax) continue;
                const double2 *U = Ub[q].as<double2>() + bt.jobs[r].u_off, *Vt = Vb[q].as<double2>() + bt.jobs[r].v_off;
                add_gemm_tiles(jobs, Cb(i), 3 * nbk[i], 1, U, 3 * nbk[i], LR[q].as<double2>() + lo[r], K[i], K[i], kk,
                               3 * nbk[i]);
                add_gemm_tiles(jobs, Vt, 3 * nbk[j], 1, Cb(j), 3 * nbk[j], LR[q].as<double2>() + ro[r], kk, kk, K[j],
                               3 * nbk[j]);
            }
            TRY(run_ggemm(jobs, dgj[q], pp.s2));
            jobs.clear();
            for (size_t r = 0; r < bt.jobs.size(); ++r) {
                const int i = bt.pi[r], j = bt.pj[r], kk = std::min(hk[r], kmax);
                if (kk >= kmax) continue;
                add_gemm_tiles(jobs, LR[q].as<double2>() + lo[r], K[i], 0, LR[q].as<double2>() + ro[r], kk,
                               zc_ + offs[i] + offs[j] * Kt, (int)Kt, K[i], K[j], kk);
            }
            TRY(run_ggemm(jobs, dgj[q], pp.s2));
            // not rank-deficient enough: dense blocks (Compute_EFields_CBFME_ACA.f90:609-620)
            for (size_t r = 0; r < bt.jobs.size(); ++r)
                if (std::min(hk[r], kmax) >= kmax) dense.push_back({bt.pi[r], bt.pj[r]});
            TRY(dense_pairs(pp.s2));
            MIDAS_HIP(hipEventRecord(pp.ev[3 + q], pp.s2));
            return 0;
        };
        for (size_t b = 0; b <= batches.size(); ++b) {
            if (b < batches.size()) {
                const int q = b & 1;
                const AcaBatch &bt = batches[b];
                // parity q's buffers were last read by the products of batch b - 2
                if (b >= 2) MIDAS_HIP(hipStreamWaitEvent(st, pp.ev[3 + q], 0));
                MIDAS_HIP(hipMemcpyAsync(dj[q].p, bt.jobs.data(), sizeof(AcaJob) * bt.jobs.size(),
                                         hipMemcpyHostToDevice, st));
                prof_at[q] = profile().recs.size();
                {
                    ProfScope ps(st, PROF_ACA, 0.0, 0.0);
                    TRY(launch_aca_cells(dc.v, ref32, k0, host_pi(quirks), dj[q].as<AcaJob>(), (int)bt.jobs.size(),
                                         Ub[q].as<double2>(), Vb[q].as<double2>(), kmax, eps_aca, dk[q].as<int>(),
                                         Gb[q].as<double2>(), st));
                }
                MIDAS_HIP(hipMemcpyAsync(pp.hk[q], dk[q].p, sizeof(int) * bt.jobs.size(), hipMemcpyDeviceToHost, st));
                MIDAS_HIP(hipEventRecord(pp.ev[1 + q], st));
            }
            if (b >= 1) TRY(products((int)b - 1));
        }
        if (!batches.empty()) {
            // the reduced matrix is complete when the second stream is
            MIDAS_HIP(hipEventRecord(pp.ev[0], pp.s2));
            MIDAS_HIP(hipStreamWaitEvent(st, pp.ev[0], 0));
        }
    }
    phase(use_aca ? "reduced matrix (ACA)" : "reduced matrix (dense)");
    // ---- 3. excitation, reduced solve, fields
    DBuf &Einc = pooled(ctx, "Einc"), &Vc = pooled(ctx, "Vc"), &Al = pooled(ctx, "alpha"), &dE = pooled(ctx, "E");
    TRY(Einc.reserve(sizeof(double2) * N3 * T2));
    TRY(Vc.reserve(sizeof(double2) * (size_t)Kt * T2));
    TRY(Al.reserve(sizeof(double2) * (size_t)Kt * T2));
    TRY(launch_rhs(dc.v, ALL, ntr, dtt.as<double>(), dtp.as<double>(), k0, quirks, ref32, Einc.as<double2>(), N3,
                   st));
    // reduced excitation V_i = C_i^T E_i of every block as one grouped launch
    jobs.clear();
    for (int i = 0; i < nb; ++i)
        if (K[i])
            add_gemm_tiles(jobs, Cb(i), 3 * nbk[i], 1, Einc.as<double2>() + 3L * bs[i], (int)N3,
                           Vc.as<double2>() + offs[i], (int)Kt, K[i], T2, 3 * nbk[i]);
    TRY(run_ggemm(jobs, djobs, st));
    midas_solve_info_t inf{};
    int r = solve_dev(ctx, (int)Kt, T2, zc_, Kt, nullptr, 0, Vc.as<double2>(), Kt, Al.as<double2>(), Kt, mode, nullptr,
                      &inf, true);
    if (r < 0) return r;
    phase("reduced solve");
    TRY(dE.reserve(sizeof(double2) * N3 * T2));
    MIDAS_HIP(hipMemsetAsync(dE.p, 0, sizeof(double2) * N3 * T2, st));
    // E_i = C_i alpha_i of every block as one grouped launch
    jobs.clear();
    for (int i = 0; i < nb; ++i)
        if (K[i])
            add_gemm_tiles(jobs, Cb(i), 3 * nbk[i], 0, Al.as<double2>() + offs[i], (int)Kt,
                           dE.as<double2>() + 3L * bs[i], (int)N3, 3 * nbk[i], T2, K[i]);
    TRY(run_ggemm(jobs, djobs, st));
    MIDAS_HIP(hipMemcpy2DAsync(E, sizeof(double2) * lde, dE.p, sizeof(double2) * N3, sizeof(double2) * N3, T2,
                               E_dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
    if (C_job) {
        MIDAS_CHECK(kbmax <= kcap, -1, "cbfm: a block has %d CBFs, more than kcap = %d", kbmax, kcap);
        std::vector<double2> zero((size_t)ldc, make_double2(0.0, 0.0));
        for (int c = 0; c < kcap; ++c) std::copy(zero.begin(), zero.end(), (double2 *)C_job + (size_t)c * ldc);
        for (int b = 0; b < nb; ++b)
            if (K[b])
                MIDAS_HIP(hipMemcpy2DAsync((double2 *)C_job + 3L * bs[b], sizeof(double2) * ldc, Cb(b),
                                           sizeof(double2) * 3 * nbk[b], sizeof(double2) * 3 * nbk[b], K[b],
                                           hipMemcpyDeviceToHost, st));
    }
    MIDAS_HIP(hipStreamSynchronize(st));
    phase("fields + copy out");
    if (tm) std::fprintf(stderr, "[cbfm] K total %ld over %d blocks\n", Kt, nb);
    if (K_out)
        for (int b = 0; b < nb; ++b) K_out[b] = K[b];
    if (info) *info = inf;
    return inf.info > 0 ? inf.info : worst_info;
}

extern "C" int midas_compute_efields_cbfm(midas_ctx *ctx, const midas_cells_t *cells, const midas_blocks_t *blocks,
                                          double k0, int quirks, int ntr_cbf, const double *th_cbf,
                                          const double *ph_cbf, int iws, int ntr, const double *tr_th,
                                          const double *tr_ph, int use_aca, int kmax, double eps_aca,
                                          double threshold, int mode, double *E, long lde, int *K_out,
                                          double *C_job, long ldc, int kcap, int sr, double res_sr,
                                          double *fsr_out, long *nnz_out, midas_solve_info_t *info)
{
    return compute_efields_cbfm(ctx, cells, blocks, k0, quirks, ntr_cbf, th_cbf, ph_cbf, iws, ntr, tr_th, tr_ph,
                                use_aca, kmax, eps_aca, threshold, mode, E, lde, false, K_out, C_job, ldc, kcap, sr,
                                res_sr, fsr_out, nnz_out, info);
}

extern "C" int midas_dev_compute_efields_cbfm(midas_ctx *ctx, const midas_cells_t *cells,
                                              const midas_blocks_t *blocks, double k0, int quirks, int ntr_cbf,
                                              const double *th_cbf, const double *ph_cbf, int iws, int ntr,
                                              const double *tr_th, const double *tr_ph, int use_aca, int kmax,
                                              double eps_aca, double threshold, int mode, double *E_dev, long lde,
                                              int *K_out, double *C_job, long ldc, int kcap, int sr, double res_sr,
                                              double *fsr_out, long *nnz_out, midas_solve_info_t *info)
{
    return compute_efields_cbfm(ctx, cells, blocks, k0, quirks, ntr_cbf, th_cbf, ph_cbf, iws, ntr, tr_th, tr_ph,
                                use_aca, kmax, eps_aca, threshold, mode, E_dev, lde, true, K_out, C_job, ldc, kcap,
                                sr, res_sr, fsr_out, nnz_out, info);
}
