// Adaptive cross approximation (partially pivoted ACA) on gfx950.
//
//  Calcul_MatrixZij_ACA  (Calcul_ACA_matrices.f90:6-219): off-diagonal block
//      Z(rows of cells [r0, r0+nr), columns of cells [c0, c0+nc)) whose rows
//      and columns are generated on the fly (computeBlockRow / computeBlockCol,
//      Green_functions.f90:905-1012, 1119-1225; FP32 intermediates when the
//      FP32-blocks quirk is set, else the FP64 element).
//  Calcul_GenMatrix_ACA  (Calcul_ACA_matrices.f90:222-424): the same iteration
//      over an explicit matrix.
//
// The iteration is sequential in the rank k (<= kmax, 50 by default) and each
// step touches O((m + n) k) data, so one ACA is latency-bound.  The kernel runs
// one ACA per workgroup and a launch carries a whole list of block pairs (the
// CBFM assembly submits every pair of a block column at once), so the chip is
// filled by independent ACAs rather than by splitting one.
//
// Semantics kept from the reference:
//   - the first pivot row is local row 0;
//   - pivots are maxloc of |.| over rows (columns) not chosen before, first
//     index on ties (exact ties are common on voxel lattices);
//   - v_k = (row - sum U(i_k, l) V(l, :)) / pivot, u_k = col - sum V(l, j_k) U(:, l);
//   - ||Z_k||_F^2 = ||Z_{k-1}||_F^2 + 2 sum_l |u_l . u_k| |v_k . v_l| + ||u_k||^2 ||v_k||^2
//     with NON-conjugated inner products (as the reference writes them);
//   - stop after term k (k >= 2) when ||u_k|| ||v_k|| / ||Z_k||_F <= eps; the
//     returned count is k, or kmax + 1 when the tolerance is never met.
//
// Storage: U is (m x kmax) column-major (the reference's Matrix_U); V is kept
// transposed, Vt (n x kmax), so that both factors are read with unit stride
// across the workgroup; the C-ABI transposes Vt to the reference's
// Matrix_V (kmax x n) on the way out.
#include "midas_aca.h"
#include "midas_elem.h"
#include "midas_host.h"

#include <climits>

namespace midas {

constexpr int ACA_T = 256;            // threads per ACA
constexpr int ACA_W = ACA_T / 64;     // waves per ACA
constexpr int ACA_KMAX = ACA_KMAX_SUPPORTED;

// ---------------------------------------------------------------------------
// entry generators: group g covers entries 3g .. 3g+2 of a row or column
// ---------------------------------------------------------------------------
template <bool REF32> struct GenCells {
    DevCells c;
    double k0, pi;
    __device__ void dyad(int o, int s, z64 e[9]) const
    {
        double rx = c.x[o] - c.x[s], ry = c.y[o] - c.y[s], rz = c.z[o] - c.z[s];
        if (REF32) {
            dyad_ref32(rx, ry, rz, k0, pi, c.kappa[s], z64{c.che[2 * s], c.che[2 * s + 1]}, e);
        } else {
            Dyad64 d = dyad64(rx, ry, rz, k0, 0.25 / pi, z64{c.kchi[2 * s], c.kchi[2 * s + 1]});
#pragma unroll
            for (int ps = 0; ps < 3; ++ps)
#pragma unroll
                for (int po = 0; po < 3; ++po) e[3 * ps + po] = dyad_entry(d, po, ps);
        }
    }
    // row i of the job, entries 3g..3g+2 (source cell c0 + g)
    __device__ void row3(const AcaJob &j, int i, int g, z64 out[3]) const
    {
        z64 e[9];
        dyad(j.r0 + i / 3, j.c0 + g, e);
        const int po = i % 3;   // picked by selects: a runtime index would put e[] in scratch
#pragma unroll
        for (int ps = 0; ps < 3; ++ps) out[ps] = po == 0 ? e[3 * ps] : po == 1 ? e[3 * ps + 1] : e[3 * ps + 2];
    }
    // column jc of the job, entries 3g..3g+2 (receiver cell r0 + g)
    __device__ void col3(const AcaJob &j, int jc, int g, z64 out[3]) const
    {
        z64 e[9];
        dyad(j.r0 + g, j.c0 + jc / 3, e);
        const int ps = jc % 3;
#pragma unroll
        for (int po = 0; po < 3; ++po) out[po] = ps == 0 ? e[po] : ps == 1 ? e[3 + po] : e[6 + po];
    }
};

struct GenMatrix {
    const double2 *A;
    long lda;
    __device__ void row3(const AcaJob &j, int i, int g, z64 out[3]) const
    {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const int c = 3 * g + q;
            double2 v = c < j.n ? A[(long)c * lda + i] : make_double2(0.0, 0.0);
            out[q] = {v.x, v.y};
        }
    }
    __device__ void col3(const AcaJob &j, int jc, int g, z64 out[3]) const
    {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const int r = 3 * g + q;
            double2 v = r < j.m ? A[(long)jc * lda + r] : make_double2(0.0, 0.0);
            out[q] = {v.x, v.y};
        }
    }
};

// ---------------------------------------------------------------------------
// workgroup reductions
// ---------------------------------------------------------------------------
__device__ inline void amax_merge(double &v, int &i, double v2, int i2)
{
    if (v2 > v || (v2 == v && i2 < i)) {
        v = v2;
        i = i2;
    }
}
__device__ inline void wave_amax(double &v, int &i)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        double v2 = __shfl_xor(v, o, 64);
        int i2 = __shfl_xor(i, o, 64);
        amax_merge(v, i, v2, i2);
    }
}
__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ inline double zabs(z64 a) { return sqrt(a.r * a.r + a.i * a.i); }
__device__ inline z64 zdiv(z64 a, z64 b)
{
    double d = b.r * b.r + b.i * b.i;
    return {(a.r * b.r + a.i * b.i) / d, (a.i * b.r - a.r * b.i) / d};
}

// acc[q] = sum_l coef[l] * x[l * ld + q] for q < nq (nq <= 3 entries at
// unit stride): one pass over l with the three sums as independent chains,
// each summed in ascending l.  The same pass marks bit q of the result when
// entry j0 + q is among the previous pivots piv[0 .. kk-1].
__device__ inline int dot3(const double2 *coef, const int *piv, const double2 *x, int ld, int kk, int j0, int nq,
                           z64 acc[3])
{
    int used = 0;
#pragma unroll
    for (int q = 0; q < 3; ++q) acc[q] = {0.0, 0.0};
    // unrolled so that the loads of several l are in flight at once (the sums
    // still run in ascending l)
    if (nq == 3) {
#pragma unroll 2
        for (int l = 0; l < kk; ++l) {
            const double2 c = coef[l];
            const int d = piv[l] - j0;
            used |= (unsigned)d < 3u ? 1 << d : 0;
            const double2 *row = x + (long)l * ld;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const double2 y = row[q];
                acc[q].r += c.x * y.x - c.y * y.y;
                acc[q].i += c.x * y.y + c.y * y.x;
            }
        }
    } else {
#pragma unroll 2
        for (int l = 0; l < kk; ++l) {
            const double2 c = coef[l];
            const int d = piv[l] - j0;
            used |= (unsigned)d < 3u ? 1 << d : 0;
            const double2 *row = x + (long)l * ld;
#pragma unroll
            for (int q = 0; q < 3; ++q)
                if (q < nq) {
                    const double2 y = row[q];
                    acc[q].r += c.x * y.x - c.y * y.y;
                    acc[q].i += c.x * y.y + c.y * y.x;
                }
        }
    }
    return used;
}

template <class G>
__global__ __launch_bounds__(ACA_T) void k_aca(G gen, const AcaJob *__restrict__ jobs, double2 *__restrict__ U,
                                               double2 *__restrict__ Vt, int kmax, double eps, int *__restrict__ kout)
{
    const AcaJob jb = jobs[blockIdx.x];
    const int m = jb.m, n = jb.n;
    const int gm = (m + 2) / 3, gn = (n + 2) / 3;
    double2 *u = U + jb.u_off;    // u(i, l) = u[l*m + i]
    double2 *v = Vt + jb.v_off;   // v(j, l) = v[l*n + j]
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;

    __shared__ int Ir[ACA_KMAX + 1], Jc[ACA_KMAX + 1];
    __shared__ double2 coef[ACA_KMAX];
    __shared__ double2 dpart[ACA_KMAX], epart[ACA_KMAX];
    __shared__ double sred[4][ACA_W];
    __shared__ int ired[2][ACA_W];
    __shared__ double s_cross;

    if (t == 0) Ir[0] = 0;
    __syncthreads();
    double nf_prev = 0.0;
    int result = kmax + 1;
    for (int k = 1; k <= kmax; ++k) {
        const int kk = k - 1;             // 0-based index of the new term
        const int ik = Ir[kk];
        // ---- (1) residual row ik into v(:, kk); arg max over unused columns
        if (t < kk) {
            double2 x = u[(long)t * m + ik];
            coef[t] = x;
        }
        __syncthreads();
        double bv = -1.0;
        int bi = INT_MAX;
        for (int g = t; g < gn; g += ACA_T) {
            z64 e[3];
            gen.row3(jb, ik, g, e);
            z64 acc[3];
            const int nq = min(3, n - 3 * g);
            const int used = dot3(coef, Jc, v + 3 * g, n, kk, 3 * g, nq, acc);
#pragma unroll
            for (int q = 0; q < 3; ++q)
                if (q < nq) {
                    const int j = 3 * g + q;
                    z64 r = {e[q].r - acc[q].r, e[q].i - acc[q].i};
                    v[(long)kk * n + j] = make_double2(r.r, r.i);
                    amax_merge(bv, bi, (used >> q) & 1 ? 0.0 : zabs(r), j);
                }
        }
        wave_amax(bv, bi);
        if (lane == 0) {
            sred[0][w] = bv;
            ired[0][w] = bi;
        }
        __syncthreads();
        bv = sred[0][0];
        bi = ired[0][0];
        for (int q = 1; q < ACA_W; ++q) amax_merge(bv, bi, sred[0][q], ired[0][q]);
        const int jk = bi;
        const double2 pv = v[(long)kk * n + jk];
        const z64 piv = {pv.x, pv.y};
        if (t == 0) Jc[kk] = jk;
        if (t < kk) coef[t] = v[(long)t * n + jk];
        __syncthreads();
        // ---- (2) scale v(:, kk) by the pivot; ||v_k||^2
        double nv2 = 0.0;
        for (int g = t; g < gn; g += ACA_T) {
            const int nq = min(3, n - 3 * g);
            double2 x[3];   // the three loads issued before the divisions
#pragma unroll
            for (int q = 0; q < 3; ++q) x[q] = q < nq ? v[(long)kk * n + 3 * g + q] : make_double2(0.0, 0.0);
#pragma unroll
            for (int q = 0; q < 3; ++q)
                if (q < nq) {
                    z64 y = zdiv(z64{x[q].x, x[q].y}, piv);
                    v[(long)kk * n + 3 * g + q] = make_double2(y.r, y.i);
                    nv2 += y.r * y.r + y.i * y.i;
                }
        }
        // ---- (3) residual column jk into u(:, kk); ||u_k||^2; arg max over unused rows
        double nu2 = 0.0;
        double rv = -1.0;
        int ri = INT_MAX;
        for (int g = t; g < gm; g += ACA_T) {
            z64 e[3];
            gen.col3(jb, jk, g, e);
            z64 acc[3];
            const int nq = min(3, m - 3 * g);
            int used = dot3(coef, Ir, u + 3 * g, m, kk, 3 * g, nq, acc);
            const int dk = Ir[kk] - 3 * g;   // the current pivot row counts as used too
            used |= (unsigned)dk < 3u ? 1 << dk : 0;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const int i = 3 * g + q;
                if (q >= nq) continue;
                z64 r = {e[q].r - acc[q].r, e[q].i - acc[q].i};
                u[(long)kk * m + i] = make_double2(r.r, r.i);
                nu2 += r.r * r.r + r.i * r.i;
                amax_merge(rv, ri, (used >> q) & 1 ? 0.0 : zabs(r), i);
            }
        }
        nu2 = wave_sum(nu2);
        nv2 = wave_sum(nv2);
        wave_amax(rv, ri);
        if (lane == 0) {
            sred[1][w] = nu2;
            sred[2][w] = nv2;
            sred[3][w] = rv;
            ired[1][w] = ri;
        }
        __syncthreads();
        // ---- (4) cross terms sum_l |u_l . u_k| |v_k . v_l| (non-conjugated).
        // Wave w owns l = w, w + ACA_W, ...; it takes two of them per pass so
        // that their eight wave reductions run interleaved.
        for (int l0 = w; l0 < kk; l0 += 2 * ACA_W) {
            const int l1 = l0 + ACA_W;
            const bool two = l1 < kk;   // uniform across the wave
            double x[8] = {0, 0, 0, 0, 0, 0, 0, 0};   // du0, dv0, du1, dv1 (re, im)
#pragma unroll 2
            for (int i = lane; i < m; i += 64) {
                const double2 b = u[(long)kk * m + i], a = u[(long)l0 * m + i];
                x[0] += a.x * b.x - a.y * b.y;
                x[1] += a.x * b.y + a.y * b.x;
                if (two) {
                    const double2 a1 = u[(long)l1 * m + i];
                    x[4] += a1.x * b.x - a1.y * b.y;
                    x[5] += a1.x * b.y + a1.y * b.x;
                }
            }
#pragma unroll 2
            for (int j = lane; j < n; j += 64) {
                const double2 a = v[(long)kk * n + j], b = v[(long)l0 * n + j];
                x[2] += a.x * b.x - a.y * b.y;
                x[3] += a.x * b.y + a.y * b.x;
                if (two) {
                    const double2 b1 = v[(long)l1 * n + j];
                    x[6] += a.x * b1.x - a.y * b1.y;
                    x[7] += a.x * b1.y + a.y * b1.x;
                }
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1)
#pragma unroll
                for (int q = 0; q < 8; ++q) x[q] += __shfl_xor(x[q], o, 64);
            if (lane == 0) {
                dpart[l0] = make_double2(x[0], x[1]);
                epart[l0] = make_double2(x[2], x[3]);
                if (two) {
                    dpart[l1] = make_double2(x[4], x[5]);
                    epart[l1] = make_double2(x[6], x[7]);
                }
            }
        }
        __syncthreads();
        if (w == 0) {
            double term = 0.0;
            if (lane < kk) {
                const double2 du = dpart[lane], dv = epart[lane];
                term = zabs(z64{du.x, du.y}) * zabs(z64{dv.x, dv.y});
            }
            term = wave_sum(term);
            if (lane == 0) s_cross = term;
        }
        __syncthreads();
        double snu = 0.0, snv = 0.0;
        rv = sred[3][0];
        ri = ired[1][0];
        for (int q = 0; q < ACA_W; ++q) {
            snu += sred[1][q];
            snv += sred[2][q];
        }
        for (int q = 1; q < ACA_W; ++q) amax_merge(rv, ri, sred[3][q], ired[1][q]);
        const double nu = sqrt(snu), nvv = sqrt(snv);
        const double nf = sqrt(nf_prev * nf_prev + 2.0 * s_cross + nu * nu * nvv * nvv);
        nf_prev = nf;
        if (k >= 2 && nu * nvv / nf <= eps) {
            result = k;
            break;
        }
        if (t == 0) Ir[k] = ri;
        __syncthreads();
    }
    if (t == 0) kout[blockIdx.x] = result;
}

// Vt (n x kmax, ld n) -> V (kmax x n, ld ldv)
__global__ void k_vt_to_v(const double2 *Vt, int n, int kmax, double2 *V, long ldv)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int l = blockIdx.y;
    if (j < n && l < kmax) V[(long)j * ldv + l] = Vt[(long)l * n + j];
}

int launch_aca_cells(const DevCells &cells, bool ref32, double k0, double pi, const AcaJob *jobs, int njobs,
                     double2 *U, double2 *Vt, int kmax, double eps, int *kout, hipStream_t st)
{
    MIDAS_CHECK(kmax >= 1 && kmax <= ACA_KMAX, -1, "aca: kmax %d outside 1..%d", kmax, ACA_KMAX);
    if (njobs <= 0) return 0;
    if (ref32)
        hipLaunchKernelGGL(k_aca<GenCells<true>>, dim3(njobs), dim3(ACA_T), 0, st, GenCells<true>{cells, k0, pi},
                           jobs, U, Vt, kmax, eps, kout);
    else
        hipLaunchKernelGGL(k_aca<GenCells<false>>, dim3(njobs), dim3(ACA_T), 0, st, GenCells<false>{cells, k0, pi},
                           jobs, U, Vt, kmax, eps, kout);
    MIDAS_HIP(hipGetLastError());
    return 0;
}

int launch_aca_matrix(const double2 *A, long lda, const AcaJob *jobs, int njobs, double2 *U, double2 *Vt, int kmax,
                      double eps, int *kout, hipStream_t st)
{
    MIDAS_CHECK(kmax >= 1 && kmax <= ACA_KMAX, -1, "aca: kmax %d outside 1..%d", kmax, ACA_KMAX);
    hipLaunchKernelGGL(k_aca<GenMatrix>, dim3(njobs), dim3(ACA_T), 0, st, GenMatrix{A, lda}, jobs, U, Vt, kmax, eps,
                       kout);
    MIDAS_HIP(hipGetLastError());
    return 0;
}

int launch_vt_to_v(const double2 *Vt, int n, int kmax, double2 *V, long ldv, hipStream_t st)
{
    hipLaunchKernelGGL(k_vt_to_v, dim3((n + 255) / 256, kmax), dim3(256), 0, st, Vt, n, kmax, V, ldv);
    MIDAS_HIP(hipGetLastError());
    return 0;
}

}  // namespace midas

// ===========================================================================
// extern "C"
// ===========================================================================
extern "C" {

int midas_aca_block(midas_ctx *ctx, const midas_cells_t *cells, double k0, int quirks, int r0, int nr, int c0, int nc,
                    int kmax, double eps, double *U, double *V, int *k_out)
{
    MIDAS_CHECK(ctx && cells && U && V && k_out, -1, "aca_block: bad arguments");
    MIDAS_CHECK(nr > 0 && nc > 0 && r0 >= 0 && c0 >= 0 && r0 + nr <= cells->n && c0 + nc <= cells->n, -1,
                "aca_block: cell ranges [%d,%d) x [%d,%d) outside 0..%d", r0, r0 + nr, c0, c0 + nc, cells->n);
    MIDAS_CHECK(r0 + nr <= c0 || c0 + nc <= r0, -1, "aca_block: row and column cell ranges overlap");
    MIDAS_CHECK(kmax >= 1 && kmax <= ACA_KMAX && eps > 0, -1, "aca_block: kmax %d (1..%d), eps %g", kmax, ACA_KMAX,
                eps);
    MIDAS_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->st;
    DevCellsOwned dc;
    TRY(dc.from_host(cells, k0, quirks, st));
    const int m = 3 * nr, n = 3 * nc;
    AcaJob job{m, n, r0, c0, 0, 0};
    DBuf dj, du, dvt, dv, dk;
    TRY(upload(dj, &job, sizeof(job), st));
    TRY(du.alloc(sizeof(double2) * (size_t)m * kmax));
    TRY(dvt.alloc(sizeof(double2) * (size_t)n * kmax));
    TRY(dv.alloc(sizeof(double2) * (size_t)n * kmax));
    TRY(dk.alloc(sizeof(int)));
    MIDAS_HIP(hipMemsetAsync(du.p, 0, du.bytes, st));
    MIDAS_HIP(hipMemsetAsync(dvt.p, 0, dvt.bytes, st));
    TRY(launch_aca_cells(dc.v, (quirks & MIDAS_Q_FP32_BLOCKS) != 0, k0, host_pi(quirks), dj.as<AcaJob>(), 1,
                         du.as<double2>(), dvt.as<double2>(), kmax, eps, dk.as<int>(), st));
    TRY(launch_vt_to_v(dvt.as<double2>(), n, kmax, dv.as<double2>(), kmax, st));
    MIDAS_HIP(hipMemcpyAsync(U, du.p, du.bytes, hipMemcpyDeviceToHost, st));
    MIDAS_HIP(hipMemcpyAsync(V, dv.p, dv.bytes, hipMemcpyDeviceToHost, st));
    TRY(copy_sync(k_out, dk.p, sizeof(int), hipMemcpyDeviceToHost, st));
    return 0;
}

int midas_aca_matrix(midas_ctx *ctx, int m, int n, const double *A, long lda, int kmax, double eps, double *U,
                     double *V, int *k_out)
{
    MIDAS_CHECK(ctx && A && U && V && k_out && m > 0 && n > 0 && lda >= m, -1, "aca_matrix: bad arguments");
    MIDAS_CHECK(kmax >= 1 && kmax <= ACA_KMAX && eps > 0, -1, "aca_matrix: kmax %d (1..%d), eps %g", kmax, ACA_KMAX,
                eps);
    MIDAS_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->st;
    AcaJob job{m, n, 0, 0, 0, 0};
    DBuf da, dj, du, dvt, dv, dk;
    TRY(da.alloc(sizeof(double2) * (size_t)m * n));
    MIDAS_HIP(hipMemcpy2DAsync(da.p, sizeof(double2) * m, A, sizeof(double2) * lda, sizeof(double2) * m, n,
                               hipMemcpyHostToDevice, st));
    TRY(upload(dj, &job, sizeof(job), st));
    TRY(du.alloc(sizeof(double2) * (size_t)m * kmax));
    TRY(dvt.alloc(sizeof(double2) * (size_t)n * kmax));
    TRY(dv.alloc(sizeof(double2) * (size_t)n * kmax));
    TRY(dk.alloc(sizeof(int)));
    MIDAS_HIP(hipMemsetAsync(du.p, 0, du.bytes, st));
    MIDAS_HIP(hipMemsetAsync(dvt.p, 0, dvt.bytes, st));
    TRY(launch_aca_matrix(da.as<double2>(), m, dj.as<AcaJob>(), 1, du.as<double2>(), dvt.as<double2>(), kmax, eps,
                          dk.as<int>(), st));
    TRY(launch_vt_to_v(dvt.as<double2>(), n, kmax, dv.as<double2>(), kmax, st));
    MIDAS_HIP(hipMemcpyAsync(U, du.p, du.bytes, hipMemcpyDeviceToHost, st));
    MIDAS_HIP(hipMemcpyAsync(V, dv.p, dv.bytes, hipMemcpyDeviceToHost, st));
    TRY(copy_sync(k_out, dk.p, sizeof(int), hipMemcpyDeviceToHost, st));
    return 0;
}

}  // extern "C"
