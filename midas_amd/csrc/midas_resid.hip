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
#include "midas_elem.h"

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

// ---------------------------------------------------------------------------
// The same two operations without the matrix: every entry is recomputed from
// the cells (compute_Green_s_tr_single_element, Compute_EFields_MoM.f90:433-523,
// as k_fill<false> computes it: the same dyad64 in FP64), so the MoM mixed
// solve need not store its 16 n^2-byte FP64 matrix for the refinement.  A
// (receiver, source) cell pair's 3 x 3 block is a I + c r r^T (a, c of the
// source cell's kappa chi), so its product with a 3-vector x is
// a x + c r (r . x): one complex exp per pair serves all right-hand sides.
// Thread = receiver cell (three rows); a workgroup takes MF_CH source cells
// (constants and X rows staged in LDS, read as broadcasts) and writes per-chunk
// partial sums, added in order by k_zx_reduce / k_rowabs_reduce.
// ---------------------------------------------------------------------------
constexpr int MF_T = 256, MF_CH = 256;

template <int NR>
__global__ __launch_bounds__(MF_T) void k_zx_free(DevCells cells, int nc, double k0, double inv4pi,
                                                  const double2 *__restrict__ X, long ldx, double2 *__restrict__ P)
{
    __shared__ double sx[MF_CH], sy[MF_CH], sz[MF_CH];
    __shared__ double2 skc[MF_CH], sdg[MF_CH];   // kappa chi, 1 - Znnpp
    __shared__ int sid[MF_CH];
    __shared__ double2 sX[MF_CH][3 * NR];
    const int q0 = blockIdx.y * MF_CH, nq = min(MF_CH, nc - q0);
    for (int e = threadIdx.x; e < nq; e += MF_T) {
        const int g = q0 + e;
        sx[e] = cells.x[g];
        sy[e] = cells.y[g];
        sz[e] = cells.z[g];
        skc[e] = make_double2(cells.kchi[2 * g], cells.kchi[2 * g + 1]);
        sdg[e] = make_double2(1.0 - cells.znnpp[2 * g], -cells.znnpp[2 * g + 1]);
        sid[e] = cells.id[g];
#pragma unroll
        for (int r = 0; r < NR; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) sX[e][3 * r + c] = X[3L * g + c + (long)r * ldx];
    }
    __syncthreads();
    const int p = blockIdx.x * MF_T + threadIdx.x;
    if (p >= nc) return;
    const double px = cells.x[p], py = cells.y[p], pz = cells.z[p];
    const int pid = cells.id[p];
    z64 acc[3 * NR];
#pragma unroll
    for (int e = 0; e < 3 * NR; ++e) acc[e] = z64{0.0, 0.0};
    for (int j = 0; j < nq; ++j) {
        if (sid[j] == pid) {
            const z64 d = {sdg[j].x, sdg[j].y};
#pragma unroll
            for (int e = 0; e < 3 * NR; ++e) acc[e] = cadd(acc[e], cmul(d, z64{sX[j][e].x, sX[j][e].y}));
            continue;
        }
        const double rx = px - sx[j], ry = py - sy[j], rz = pz - sz[j];
        const Dyad64 d = dyad64(rx, ry, rz, k0, inv4pi, z64{skc[j].x, skc[j].y});
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const z64 x0 = {sX[j][3 * r].x, sX[j][3 * r].y}, x1 = {sX[j][3 * r + 1].x, sX[j][3 * r + 1].y},
                      x2 = {sX[j][3 * r + 2].x, sX[j][3 * r + 2].y};
            const z64 s = {rx * x0.r + ry * x1.r + rz * x2.r, rx * x0.i + ry * x1.i + rz * x2.i};
            const z64 cs = cmul(d.c, s);
            acc[3 * r] = cadd(acc[3 * r], cadd(cmul(d.a, x0), cscale(cs, rx)));
            acc[3 * r + 1] = cadd(acc[3 * r + 1], cadd(cmul(d.a, x1), cscale(cs, ry)));
            acc[3 * r + 2] = cadd(acc[3 * r + 2], cadd(cmul(d.a, x2), cscale(cs, rz)));
        }
    }
    const long n = 3L * nc;
    double2 *P0 = P + (long)blockIdx.y * n * NR;
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) P0[(3L * p + c) * NR + r] = make_double2(acc[3 * r + c].r, acc[3 * r + c].i);
}

__global__ __launch_bounds__(MF_T) void k_rowabs_free(DevCells cells, int nc, double k0, double inv4pi,
                                                      double *__restrict__ P)
{
    __shared__ double sx[MF_CH], sy[MF_CH], sz[MF_CH];
    __shared__ double2 skc[MF_CH], sdg[MF_CH];
    __shared__ int sid[MF_CH];
    const int q0 = blockIdx.y * MF_CH, nq = min(MF_CH, nc - q0);
    for (int e = threadIdx.x; e < nq; e += MF_T) {
        const int g = q0 + e;
        sx[e] = cells.x[g];
        sy[e] = cells.y[g];
        sz[e] = cells.z[g];
        skc[e] = make_double2(cells.kchi[2 * g], cells.kchi[2 * g + 1]);
        sdg[e] = make_double2(1.0 - cells.znnpp[2 * g], -cells.znnpp[2 * g + 1]);
        sid[e] = cells.id[g];
    }
    __syncthreads();
    const int p = blockIdx.x * MF_T + threadIdx.x;
    if (p >= nc) return;
    const double px = cells.x[p], py = cells.y[p], pz = cells.z[p];
    const int pid = cells.id[p];
    double s[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < nq; ++j) {
        if (sid[j] == pid) {
            const double a = sqrt(sdg[j].x * sdg[j].x + sdg[j].y * sdg[j].y);
#pragma unroll
            for (int c = 0; c < 3; ++c) s[c] += a;
            continue;
        }
        const double rx = px - sx[j], ry = py - sy[j], rz = pz - sz[j];
        const Dyad64 d = dyad64(rx, ry, rz, k0, inv4pi, z64{skc[j].x, skc[j].y});
#pragma unroll
        for (int po = 0; po < 3; ++po)
#pragma unroll
            for (int ps = 0; ps < 3; ++ps) {
                const z64 e = dyad_entry(d, po, ps);
                s[po] += sqrt(e.r * e.r + e.i * e.i);
            }
    }
    double *P0 = P + (long)blockIdx.y * 3L * nc;
#pragma unroll
    for (int c = 0; c < 3; ++c) P0[3L * p + c] = s[c];
}

size_t mf_scratch_bytes(int ncells, int nrhs)
{
    return (size_t)((ncells + MF_CH - 1) / MF_CH) * 3 * ncells * (nrhs > 0 ? nrhs : 1) * sizeof(double2);
}

// R = B - Z X (nrhs <= 4) with Z recomputed from the cells; scratch of
// mf_scratch_bytes(ncells, nrhs).
int launch_residual_free(const DevCells &cells, int nc, double k0, double pi, const double2 *X, long ldx, int nrhs,
                         const double2 *B, long ldb, double2 *R, long ldr, void *scratch, hipStream_t st)
{
    const int nchunk = (nc + MF_CH - 1) / MF_CH, n = 3 * nc;
    const dim3 g((nc + MF_T - 1) / MF_T, nchunk);
    double2 *P = (double2 *)scratch;
    const double inv4pi = 0.25 / pi;
    switch (nrhs) {
#define MF_CASE(NR)                                                                                                  \
    case NR:                                                                                                         \
        hipLaunchKernelGGL(k_zx_free<NR>, g, dim3(MF_T), 0, st, cells, nc, k0, inv4pi, X, ldx, P);                   \
        hipLaunchKernelGGL(k_zx_reduce<NR>, dim3((n + 255) / 256), dim3(256), 0, st, P, nchunk, n, B, ldb, R, ldr); \
        break;
        MF_CASE(1)
        MF_CASE(2)
        MF_CASE(3)
        MF_CASE(4)
#undef MF_CASE
    default:
        MIDAS_CHECK(false, -1, "residual_free: nrhs %d outside 1..4", nrhs);
    }
    MIDAS_HIP(hipGetLastError());
    return 0;
}

// rowsum[i] = sum_j |Z_ij| with Z recomputed from the cells; scratch of
// mf_scratch_bytes(ncells, 1).
int launch_rowabs_free(const DevCells &cells, int nc, double k0, double pi, double *rowsum, void *scratch,
                       hipStream_t st)
{
    const int nchunk = (nc + MF_CH - 1) / MF_CH, n = 3 * nc;
    hipLaunchKernelGGL(k_rowabs_free, dim3((nc + MF_T - 1) / MF_T, nchunk), dim3(MF_T), 0, st, cells, nc, k0,
                       0.25 / pi, (double *)scratch);
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
