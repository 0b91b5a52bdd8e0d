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

// The cross terms of step k need u_l . u_k and v_k . v_l for every l < k.
// They come from the non-conjugated Gram matrices Gu(l, l') = u_l . u_l',
// Gv(l, l') = v_l . v_l' (per job in global memory, kmax x kmax each) and the
// products of the generated row / column with the previous terms, summed in
// the same passes that read those terms for the residuals:
//   u_l . u_k = u_l . col - sum_l' v_l'(jk) Gu(l, l'),
//   v_k . v_l = (row . v_l - sum_l' u_l'(ik) Gv(l, l')) / pivot,
// so no step reads the factors a second time.  Per-wave partial sums are
// added in a fixed order, so the result does not depend on scheduling.
template <class G>
__global__ __launch_bounds__(ACA_T) void k_aca(G gen, const AcaJob *__restrict__ jobs, double2 *__restrict__ U,
                                               double2 *__restrict__ Vt, int kmax, double eps, int *__restrict__ kout,
                                               double2 *__restrict__ gram)
{
    const AcaJob jb = jobs[blockIdx.x];
    const int m = jb.m, n = jb.n;
    const int gm = (m + 2) / 3, gn = (n + 2) / 3;
    double2 *u = U + jb.u_off;    // u(i, l) = u[l*m + i]
    double2 *v = Vt + jb.v_off;   // v(j, l) = v[l*n + j]
    double2 *Gu = gram + (long)blockIdx.x * 2 * kmax * kmax, *Gv = Gu + (long)kmax * kmax;
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;

    __shared__ int Ir[ACA_KMAX + 1], Jc[ACA_KMAX + 1];
    __shared__ double2 cu[ACA_KMAX], cv[ACA_KMAX];   // u_l(ik), v_l(jk)
    __shared__ double2 part[ACA_W][ACA_KMAX];       // per-wave row . v_l / col . u_l
    __shared__ double2 brow[ACA_KMAX];              // row . v_l
    __shared__ double sred[4][ACA_W];
    __shared__ double2 zred[2][ACA_W];              // u_k . u_k, v_k . v_k (non-conjugated)
    __shared__ int ired[2][ACA_W];
    __shared__ double s_cross;

    if (t == 0) Ir[0] = 0;
    __syncthreads();
    double nf_prev = 0.0;
    int result = kmax + 1;
    for (int k = 1; k <= kmax; ++k) {
        const int kk = k - 1;             // 0-based index of the new term
        const int ik = Ir[kk];
        // ---- (1) residual row ik into v(:, kk); arg max over unused columns;
        // row . v_l for l < kk
        if (t < kk) cu[t] = u[(long)t * m + ik];
        __syncthreads();
        double bv = -1.0;
        int bi = INT_MAX;
        for (int g0 = 0; g0 < gn; g0 += ACA_T) {   // uniform trip count: the wave sums below
            const int g = g0 + t;
            const bool ok = g < gn;
            const int nq = ok ? min(3, n - 3 * g) : 0;
            z64 e[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
            if (ok) gen.row3(jb, ik, g, e);
            z64 acc[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
            int used = 0;
#pragma unroll 2
            for (int l = 0; l < kk; ++l) {
                const double2 c = cu[l];
                const int d = Jc[l] - 3 * g;
                used |= (unsigned)d < 3u ? 1 << d : 0;
                double px = 0.0, py = 0.0;
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const double2 y = q < nq ? v[(long)l * n + 3 * g + q] : make_double2(0.0, 0.0);
                    acc[q].r += c.x * y.x - c.y * y.y;
                    acc[q].i += c.x * y.y + c.y * y.x;
                    px += e[q].r * y.x - e[q].i * y.y;
                    py += e[q].r * y.y + e[q].i * y.x;
                }
                px = wave_sum(px);
                py = wave_sum(py);
                if (lane == 0) {
                    const double2 o = g0 ? part[w][l] : make_double2(0.0, 0.0);
                    part[w][l] = make_double2(o.x + px, o.y + py);
                }
            }
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
        if (t < kk) {
            cv[t] = v[(long)t * n + jk];
            double2 b = part[0][t];
            for (int q = 1; q < ACA_W; ++q) b = make_double2(b.x + part[q][t].x, b.y + part[q][t].y);
            brow[t] = b;
        }
        __syncthreads();
        // ---- (2) scale v(:, kk) by the pivot; ||v_k||^2 and v_k . v_k
        double nv2 = 0.0, vvx = 0.0, vvy = 0.0;
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
                    vvx += y.r * y.r - y.i * y.i;
                    vvy += 2.0 * y.r * y.i;
                }
        }
        // ---- (3) residual column jk into u(:, kk); ||u_k||^2, u_k . u_k; arg max
        // over unused rows; col . u_l for l < kk
        double nu2 = 0.0, uux = 0.0, uuy = 0.0;
        double rv = -1.0;
        int ri = INT_MAX;
        for (int g0 = 0; g0 < gm; g0 += ACA_T) {
            const int g = g0 + t;
            const bool ok = g < gm;
            const int nq = ok ? min(3, m - 3 * g) : 0;
            z64 e[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
            if (ok) gen.col3(jb, jk, g, e);
            z64 acc[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
            int used = 0;
#pragma unroll 2
            for (int l = 0; l < kk; ++l) {
                const double2 c = cv[l];
                const int d = Ir[l] - 3 * g;
                used |= (unsigned)d < 3u ? 1 << d : 0;
                double px = 0.0, py = 0.0;
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const double2 y = q < nq ? u[(long)l * m + 3 * g + q] : make_double2(0.0, 0.0);
                    acc[q].r += c.x * y.x - c.y * y.y;
                    acc[q].i += c.x * y.y + c.y * y.x;
                    px += e[q].r * y.x - e[q].i * y.y;
                    py += e[q].r * y.y + e[q].i * y.x;
                }
                px = wave_sum(px);
                py = wave_sum(py);
                if (lane == 0) {
                    const double2 o = g0 ? part[w][l] : make_double2(0.0, 0.0);
                    part[w][l] = make_double2(o.x + px, o.y + py);
                }
            }
            const int dk = Ir[kk] - 3 * g;   // the current pivot row counts as used too
            used |= (unsigned)dk < 3u ? 1 << dk : 0;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const int i = 3 * g + q;
                if (q >= nq) continue;
                z64 r = {e[q].r - acc[q].r, e[q].i - acc[q].i};
                u[(long)kk * m + i] = make_double2(r.r, r.i);
                nu2 += r.r * r.r + r.i * r.i;
                uux += r.r * r.r - r.i * r.i;
                uuy += 2.0 * r.r * r.i;
                amax_merge(rv, ri, (used >> q) & 1 ? 0.0 : zabs(r), i);
            }
        }
        nu2 = wave_sum(nu2);
        nv2 = wave_sum(nv2);
        uux = wave_sum(uux);
        uuy = wave_sum(uuy);
        vvx = wave_sum(vvx);
        vvy = wave_sum(vvy);
        wave_amax(rv, ri);
        if (lane == 0) {
            sred[1][w] = nu2;
            sred[2][w] = nv2;
            sred[3][w] = rv;
            ired[1][w] = ri;
            zred[0][w] = make_double2(uux, uuy);
            zred[1][w] = make_double2(vvx, vvy);
        }
        __syncthreads();
        // ---- (4) cross terms sum_l |u_l . u_k| |v_k . v_l| from the Gram rows
        // (kk <= 63 < 64: wave 0 holds them all); the Gram matrices take row
        // and column kk
        if (w == 0) {
            double term = 0.0;
            if (lane < kk) {
                const int l = lane;
                double2 a = part[0][l];
                for (int q = 1; q < ACA_W; ++q) a = make_double2(a.x + part[q][l].x, a.y + part[q][l].y);
                double2 b = brow[l];
                const double2 *gu = Gu + (long)l * kmax, *gv = Gv + (long)l * kmax;
                for (int l2 = 0; l2 < kk; ++l2) {
                    const double2 x = gu[l2], y = gv[l2], c = cv[l2], d = cu[l2];
                    a.x -= c.x * x.x - c.y * x.y;
                    a.y -= c.x * x.y + c.y * x.x;
                    b.x -= d.x * y.x - d.y * y.y;
                    b.y -= d.x * y.y + d.y * y.x;
                }
                const z64 sv = zdiv(z64{b.x, b.y}, piv);
                Gu[(long)kk * kmax + l] = a;
                Gu[(long)l * kmax + kk] = a;
                Gv[(long)kk * kmax + l] = make_double2(sv.r, sv.i);
                Gv[(long)l * kmax + kk] = make_double2(sv.r, sv.i);
                term = zabs(z64{a.x, a.y}) * zabs(sv);
            }
            term = wave_sum(term);
            if (lane == 0) {
                s_cross = term;
                double2 su = zred[0][0], svv = zred[1][0];
                for (int q = 1; q < ACA_W; ++q) {
                    su = make_double2(su.x + zred[0][q].x, su.y + zred[0][q].y);
                    svv = make_double2(svv.x + zred[1][q].x, svv.y + zred[1][q].y);
                }
                Gu[(long)kk * kmax + kk] = su;
                Gv[(long)kk * kmax + kk] = svv;
            }
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
                     double2 *U, double2 *Vt, int kmax, double eps, int *kout, double2 *gram, hipStream_t st)
{
    MIDAS_CHECK(kmax >= 1 && kmax <= ACA_KMAX && gram, -1, "aca: kmax %d outside 1..%d or no Gram buffer", kmax,
                ACA_KMAX);
    if (njobs <= 0) return 0;
    if (ref32)
        hipLaunchKernelGGL(k_aca<GenCells<true>>, dim3(njobs), dim3(ACA_T), 0, st, GenCells<true>{cells, k0, pi},
                           jobs, U, Vt, kmax, eps, kout, gram);
    else
        hipLaunchKernelGGL(k_aca<GenCells<false>>, dim3(njobs), dim3(ACA_T), 0, st, GenCells<false>{cells, k0, pi},
                           jobs, U, Vt, kmax, eps, kout, gram);
    MIDAS_HIP(hipGetLastError());
    return 0;
}

int launch_aca_matrix(const double2 *A, long lda, const AcaJob *jobs, int njobs, double2 *U, double2 *Vt, int kmax,
                      double eps, int *kout, double2 *gram, hipStream_t st)
{
    MIDAS_CHECK(kmax >= 1 && kmax <= ACA_KMAX && gram, -1, "aca: kmax %d outside 1..%d or no Gram buffer", kmax,
                ACA_KMAX);
    hipLaunchKernelGGL(k_aca<GenMatrix>, dim3(njobs), dim3(ACA_T), 0, st, GenMatrix{A, lda}, jobs, U, Vt, kmax, eps,
                       kout, gram);
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
    DBuf dj, du, dvt, dv, dk, dg;
    TRY(upload(dj, &job, sizeof(job), st));
    TRY(du.alloc(sizeof(double2) * (size_t)m * kmax));
    TRY(dvt.alloc(sizeof(double2) * (size_t)n * kmax));
    TRY(dv.alloc(sizeof(double2) * (size_t)n * kmax));
    TRY(dk.alloc(sizeof(int)));
    TRY(dg.alloc(aca_gram_bytes(kmax, 1)));
    MIDAS_HIP(hipMemsetAsync(du.p, 0, du.bytes, st));
    MIDAS_HIP(hipMemsetAsync(dvt.p, 0, dvt.bytes, st));
    TRY(launch_aca_cells(dc.v, (quirks & MIDAS_Q_FP32_BLOCKS) != 0, k0, host_pi(quirks), dj.as<AcaJob>(), 1,
                         du.as<double2>(), dvt.as<double2>(), kmax, eps, dk.as<int>(), dg.as<double2>(), st));
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
    DBuf da, dj, du, dvt, dv, dk, dg;
    TRY(dg.alloc(aca_gram_bytes(kmax, 1)));
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
                          dk.as<int>(), dg.as<double2>(), st));
    TRY(launch_vt_to_v(dvt.as<double2>(), n, kmax, dv.as<double2>(), kmax, st));
    MIDAS_HIP(hipMemcpyAsync(U, du.p, du.bytes, hipMemcpyDeviceToHost, st));
    MIDAS_HIP(hipMemcpyAsync(V, dv.p, dv.bytes, hipMemcpyDeviceToHost, st));
    TRY(copy_sync(k_out, dk.p, sizeof(int), hipMemcpyDeviceToHost, st));
    return 0;
}

}  // extern "C"
