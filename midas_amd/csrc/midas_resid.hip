// One pass over the FP64 matrix for a few right-hand sides (gfx950).
//
//   residual  R = B - Z X   (the FP64 residual of iterative refinement, as the
//                            reference's error estimate reads Z again,
//                            Compute_EFields_MoM.f90:154-196)
//   row sums  s_i = sum_j |Z_ij|   (||Z||_inf, PZLANGE 'I')
//
// With 2 right-hand sides a GEMM tiled for MFMA wastes most of each tile and
// reads Z at a fraction of HBM speed; both operations are one stream over the
// 16 n^2 bytes of Z.  The grid is 2-D -- 512-row strips x 256-column chunks --
// so every CU streams its own part of Z; each workgroup keeps its X chunk in
// LDS and writes per-chunk partial sums, which a second small kernel adds up
// in a fixed order (the result does not depend on scheduling).
#include "midas_common.h"

namespace midas {

constexpr int RS_ROWS = 512, RS_COLS = 256, RS_T = 256;

template <int NR>
__global__ __launch_bounds__(RS_T) void k_zx_partial(const double2 *__restrict__ Z, long ldz, int n,
                                                     const double2 *__restrict__ X, long ldx,
                                                     double2 *__restrict__ P)
{
    const int c0 = blockIdx.y * RS_COLS;
    const int c1 = min(n, c0 + RS_COLS);
    __shared__ double2 xs[RS_COLS][NR];
    for (int e = threadIdx.x; e < RS_COLS * NR; e += RS_T) {
        const int c = e / NR, q = e % NR;
        xs[c][q] = (c0 + c < n) ? X[(long)(c0 + c) + (long)q * ldx] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    const int r0 = blockIdx.x * RS_ROWS + threadIdx.x, r1 = r0 + RS_T;
    const bool v0 = r0 < n, v1 = r1 < n;
    z64 a0[NR], a1[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) a0[q] = a1[q] = z64{0.0, 0.0};
    const double2 *col = Z + (long)c0 * ldz;
#pragma unroll 4
    for (int c = c0; c < c1; ++c, col += ldz) {
        const double2 z0 = v0 ? col[r0] : make_double2(0.0, 0.0);
        const double2 z1 = v1 ? col[r1] : make_double2(0.0, 0.0);
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            const double2 x = xs[c - c0][q];
            a0[q].r += z0.x * x.x - z0.y * x.y;
            a0[q].i += z0.x * x.y + z0.y * x.x;
            a1[q].r += z1.x * x.x - z1.y * x.y;
            a1[q].i += z1.x * x.y + z1.y * x.x;
        }
    }
    double2 *p = P + (long)blockIdx.y * n * NR;
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        if (v0) p[(long)r0 * NR + q] = make_double2(a0[q].r, a0[q].i);
        if (v1) p[(long)r1 * NR + q] = make_double2(a1[q].r, a1[q].i);
    }
}

// R(:, q) = B(:, q) - sum over chunks of P, chunks in order
template <int NR>
__global__ void k_zx_reduce(const double2 *__restrict__ P, int nchunk, int n, const double2 *__restrict__ B, long ldb,
                            double2 *__restrict__ R, long ldr)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        double sr = 0.0, si = 0.0;
        for (int c = 0; c < nchunk; ++c) {
            const double2 v = P[((long)c * n + r) * NR + q];
            sr += v.x;
            si += v.y;
        }
        const double2 b = B[(long)r + (long)q * ldb];
        R[(long)r + (long)q * ldr] = make_double2(b.x - sr, b.y - si);
    }
}

__global__ __launch_bounds__(RS_T) void k_rowabs_partial(const double2 *__restrict__ Z, long ldz, int n,
                                                         double *__restrict__ P)
{
    const int c0 = blockIdx.y * RS_COLS;
    const int c1 = min(n, c0 + RS_COLS);
    const int r0 = blockIdx.x * RS_ROWS + threadIdx.x, r1 = r0 + RS_T;
    const bool v0 = r0 < n, v1 = r1 < n;
    double s0 = 0.0, s1 = 0.0;
    const double2 *col = Z + (long)c0 * ldz;
#pragma unroll 4
    for (int c = c0; c < c1; ++c, col += ldz) {
        if (v0) {
            const double2 z = col[r0];
            s0 += sqrt(z.x * z.x + z.y * z.y);
        }
        if (v1) {
            const double2 z = col[r1];
            s1 += sqrt(z.x * z.x + z.y * z.y);
        }
    }
    double *p = P + (long)blockIdx.y * n;
    if (v0) p[r0] = s0;
    if (v1) p[r1] = s1;
}

__global__ void k_rowabs_reduce(const double *__restrict__ P, int nchunk, int n, double *__restrict__ rowsum)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += P[(long)c * n + r];
    rowsum[r] = s;
}

size_t zx_scratch_bytes(int n, int nrhs) { return (size_t)((n + RS_COLS - 1) / RS_COLS) * n * nrhs * sizeof(double2); }

// R = B - Z X for nrhs <= 4; scratch of zx_scratch_bytes(n, nrhs).
int launch_residual_few(const double2 *Z, long ldz, int n, const double2 *X, long ldx, int nrhs, const double2 *B,
                        long ldb, double2 *R, long ldr, void *scratch, hipStream_t st)
{
    const int nchunk = (n + RS_COLS - 1) / RS_COLS;
    const dim3 g((n + RS_ROWS - 1) / RS_ROWS, nchunk);
    double2 *P = (double2 *)scratch;
    switch (nrhs) {
    case 1:
        hipLaunchKernelGGL(k_zx_partial<1>, g, dim3(RS_T), 0, st, Z, ldz, n, X, ldx, P);
        hipLaunchKernelGGL(k_zx_reduce<1>, dim3((n + 255) / 256), dim3(256), 0, st, P, nchunk, n, B, ldb, R, ldr);
        break;
    case 2:
        hipLaunchKernelGGL(k_zx_partial<2>, g, dim3(RS_T), 0, st, Z, ldz, n, X, ldx, P);
        hipLaunchKernelGGL(k_zx_reduce<2>, dim3((n + 255) / 256), dim3(256), 0, st, P, nchunk, n, B, ldb, R, ldr);
        break;
    case 3:
        hipLaunchKernelGGL(k_zx_partial<3>, g, dim3(RS_T), 0, st, Z, ldz, n, X, ldx, P);
        hipLaunchKernelGGL(k_zx_reduce<3>, dim3((n + 255) / 256), dim3(256), 0, st, P, nchunk, n, B, ldb, R, ldr);
        break;
    case 4:
        hipLaunchKernelGGL(k_zx_partial<4>, g, dim3(RS_T), 0, st, Z, ldz, n, X, ldx, P);
        hipLaunchKernelGGL(k_zx_reduce<4>, dim3((n + 255) / 256), dim3(256), 0, st, P, nchunk, n, B, ldb, R, ldr);
        break;
    default:
        MIDAS_CHECK(false, -1, "residual_few: nrhs %d outside 1..4", nrhs);
    }
    MIDAS_HIP(hipGetLastError());
    return 0;
}

// rowsum[i] = sum_j |Z_ij|; scratch of zx_scratch_bytes(n, 1).
int launch_rowabs(const double2 *Z, long ldz, int n, double *rowsum, void *scratch, hipStream_t st)
{
    const int nchunk = (n + RS_COLS - 1) / RS_COLS;
    hipLaunchKernelGGL(k_rowabs_partial, dim3((n + RS_ROWS - 1) / RS_ROWS, nchunk), dim3(RS_T), 0, st, Z, ldz, n,
                       (double *)scratch);
    hipLaunchKernelGGL(k_rowabs_reduce, dim3((n + 255) / 256), dim3(256), 0, st, (const double *)scratch, nchunk, n,
                       rowsum);
    MIDAS_HIP(hipGetLastError());
    return 0;
}

__global__ void k_iota1(int *id, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) id[i] = i + 1;
}

int launch_iota1(int *id, int n, hipStream_t st)
{
    hipLaunchKernelGGL(k_iota1, dim3((n + 255) / 256), dim3(256), 0, st, id, n);
    MIDAS_HIP(hipGetLastError());
    return 0;
}

}  // namespace midas
