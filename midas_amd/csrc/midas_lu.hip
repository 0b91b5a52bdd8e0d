// Dense complex LU with partial pivoting and triangular solves (gfx950).
//
// Replaces ScaLAPACK PZGESV (Compute_EFields_MoM.f90:173, Compute_EFields_CBFME.f90:912)
// on one GPU: right-looking blocked LU, column-major, interleaved complex.
//
//   outer panel  nb columns : factored in inner blocks of IB columns;
//   inner column step       : one fused kernel per column -- scale + rank-1 update
//                             for the previous column, then the pivot search of
//                             this column; the last workgroup to finish reduces
//                             the per-workgroup maxima, swaps the pivot row across
//                             the outer panel and records ipiv (LAPACK getf2 order);
//   inner block update      : IBxIB unit-lower solve + GEMM on the rest of the panel;
//   end of a panel          : L11^-1 of its diagonal block (one wave per column);
//   outside the panel       : row swaps (laswp), U12 = L11^-1 A12 as one GEMM with
//                             that inverse, trailing GEMM A22 -= L21 U12.
//
// T = double is the FP64 baseline; T = float holds the complex-FP32 factors of
// the mixed-precision solvers, whose trailing GEMM runs either on FP32 MFMA or
// on the FP16/BF16 split kernels (midas_gemm16.hip).
#include "midas_common.h"
#include "midas_lu.h"

namespace midas {

constexpr int STEP_THREADS = 256;
constexpr int STEP_ROWS = 512;    // rows per workgroup in the column-step kernel

// ---------------------------------------------------------------------------
// Fused column step.  On entry column j-1 has been pivoted (row j-1 holds the
// pivot row).  Rows [j, n): scale column j-1 by 1/pivot, update columns
// [j, jend) with the rank-1 term, then find this workgroup's max |A(i, j)|.
// The last workgroup reduces, swaps rows j <-> p over columns [pc0, pc1),
// stores ipiv[j] = p and 1/pivot for the next step.  With j == jend only the
// scaling of column j-1 is done.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(STEP_THREADS) void k_getf2_step(typename cvec<T>::t *A, long lda, int n, int j,
                                                             int jend, int pc0, int pc1, int *ipiv,
                                                             LuWork<T> w)
{
    typedef typename cvec<T>::t V;
    __shared__ T smax[STEP_THREADS];
    __shared__ int sidx[STEP_THREADS];
    __shared__ V urow[64];
    __shared__ int last;
    const int tid = threadIdx.x;
    // matrix of the batch
    const int zb = blockIdx.y;
    A += zb * w.sA;
    ipiv += (long)zb * n;
    T *pmax = w.pmax + (long)zb * w.maxblk;
    int *pidx = w.pidx + (long)zb * w.maxblk;
    int *counter = w.counter + zb;
    cplx<T> *pinvp = w.pinv + zb;
    const int r0 = j + blockIdx.x * STEP_ROWS;
    const int r1 = min(n, r0 + STEP_ROWS);
    const int ncol = jend - j;   // columns to update
    cplx<T> pinv = {0, 0};
    if (j > w.jb) {
        pinv = pinvp[0];
        for (int c = tid; c < ncol; c += STEP_THREADS) urow[c] = A[(long)(j - 1) + (long)(j + c) * lda];
    }
    __syncthreads();
    T best = (T)-1;
    int bidx = 0x7fffffff;
    for (int i = r0 + tid; i < r1; i += STEP_THREADS) {
        if (j > w.jb) {
            V a = A[(long)i + (long)(j - 1) * lda];
            cplx<T> l = cmul(cplx<T>{a.x, a.y}, pinv);
            A[(long)i + (long)(j - 1) * lda] = V{l.r, l.i};
            for (int c = 0; c < ncol; ++c) {
                V *p = A + (long)i + (long)(j + c) * lda;
                V v = *p;
                cplx<T> u = {urow[c].x, urow[c].y};
                cplx<T> r = csub(cplx<T>{v.x, v.y}, cmul(l, u));
                *p = V{r.r, r.i};
            }
        }
        if (ncol > 0) {
            V v = A[(long)i + (long)j * lda];
            // |Re| + |Im|, the complex pivot measure of LAPACK i?amax (dcabs1)
            T m2 = fabs(v.x) + fabs(v.y);
            if (m2 > best || (m2 == best && i < bidx)) { best = m2; bidx = i; }
        }
    }
    if (ncol == 0) return;
    smax[tid] = best;
    sidx[tid] = bidx;
    __syncthreads();
    for (int s = STEP_THREADS / 2; s > 0; s >>= 1) {
        if (tid < s) {
            T o = smax[tid + s];
            int oi = sidx[tid + s];
            if (o > smax[tid] || (o == smax[tid] && oi < sidx[tid])) { smax[tid] = o; sidx[tid] = oi; }
        }
        __syncthreads();
    }
    if (tid == 0) {
        pmax[blockIdx.x] = smax[0];
        pidx[blockIdx.x] = sidx[0];
        __threadfence();
        int done = atomicAdd(counter, 1);
        last = (done == (int)gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    // last workgroup: global argmax over the per-workgroup candidates
    // (first index on ties, as i?amax)
    {
        T b = (T)-1;
        int bi = 0x7fffffff;
        for (int q = tid; q < (int)gridDim.x; q += STEP_THREADS) {
            T v = ((volatile T *)pmax)[q];
            int vi = ((volatile int *)pidx)[q];
            if (v > b || (v == b && vi < bi)) { b = v; bi = vi; }
        }
        smax[tid] = b;
        sidx[tid] = bi;
        __syncthreads();
        for (int s2 = STEP_THREADS / 2; s2 > 0; s2 >>= 1) {
            if (tid < s2) {
                T o = smax[tid + s2];
                int oi = sidx[tid + s2];
                if (o > smax[tid] || (o == smax[tid] && oi < sidx[tid])) { smax[tid] = o; sidx[tid] = oi; }
            }
            __syncthreads();
        }
        if (tid == 0) {
            if (sidx[0] == 0x7fffffff) sidx[0] = j;
            ipiv[j] = sidx[0];
            *counter = 0;
        }
    }
    __syncthreads();
    const int p = sidx[0];
    if (p != j) {
        for (int c = pc0 + tid; c < pc1; c += STEP_THREADS) {
            V *a = A + (long)j + (long)c * lda;
            V *b = A + (long)p + (long)c * lda;
            V t = *a;
            *a = *b;
            *b = t;
        }
    }
    __syncthreads();
    if (tid == 0) {
        V d = A[(long)j + (long)j * lda];
        cplx<T> dv = {d.x, d.y};
        if (d.x == (T)0 && d.y == (T)0) {
            // exactly singular: record the first zero pivot (LAPACK INFO), keep going
            atomicCAS(w.info + zb, 0, j + 1);
            pinvp[0] = cplx<T>{0, 0};
        } else {
            pinvp[0] = cinv(dv);
        }
    }
}

// Forward substitution with the unit-lower IBxIB block L(jb:jend, jb:jend) on
// rows [jb, jend) of columns [c0, c1): one thread per column.
template <typename T>
__global__ void k_trsm_lunit_small(typename cvec<T>::t *A, long lda, long sA, int jb, int ib, int c0, int c1)
{
    typedef typename cvec<T>::t V;
    A += blockIdx.y * sA;
    int c = c0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= c1) return;
    V *x = A + (long)jb + (long)c * lda;
    const V *L = A + (long)jb + (long)jb * lda;
    for (int r = 1; r < ib; ++r) {
        cplx<T> acc = {x[r].x, x[r].y};
        for (int q = 0; q < r; ++q) {
            V l = L[(long)r + (long)q * lda];
            acc = csub(acc, cmul(cplx<T>{l.x, l.y}, cplx<T>{x[q].x, x[q].y}));
        }
        x[r] = V{acc.r, acc.i};
    }
}

// Apply the row interchanges ipiv[k0..k1) to columns [c0, c1) (and [c2, c3)).
template <typename T>
__global__ void k_laswp(typename cvec<T>::t *A, long lda, long sA, const int *ipiv, int n, int k0, int k1, int c0,
                        int c1, int c2, int c3)
{
    typedef typename cvec<T>::t V;
    A += blockIdx.y * sA;
    ipiv += (long)blockIdx.y * n;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int w0 = c1 - c0;
    int c = t < w0 ? c0 + t : c2 + (t - w0);
    if (t >= w0 + (c3 - c2)) return;
    V *col = A + (long)c * lda;
    for (int j = k0; j < k1; ++j) {
        int p = ipiv[j];
        if (p != j) {
            V a = col[j];
            col[j] = col[p];
            col[p] = a;
        }
    }
}

// Inverse of one triangular bs x bs diagonal block D (bs <= 64 R), one wave per
// column of the inverse.  Column-oriented substitution: lane l owns rows
// l, l+64, ..; at step q the solved x_q is broadcast and every row updates with
// column q of D (a unit-stride read).  The columns of D are loaded QC at a time
// ahead of the dependent chain of broadcasts, so the loop waits for memory once
// per QC steps.  LOWER: unit lower L, x = L^-1 e_c; otherwise upper U with its
// diagonal, x = U^-1 e_c.  Inv is ldi x bs.
constexpr int TRINV_QC = 8;
template <typename T, bool LOWER, int R>
__global__ __launch_bounds__(256) void k_trinv_wave(const typename cvec<T>::t *D, long lda, long sD, int bs,
                                                    typename cvec<T>::t *Inv, long ldi, long sI)
{
    typedef typename cvec<T>::t V;
    D += blockIdx.y * sD;
    Inv += blockIdx.y * sI;
    const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (c >= bs) return;   // wave-uniform
    cplx<T> x[R];
#pragma unroll
    for (int t = 0; t < R; ++t) x[t] = cplx<T>{(T)((lane + 64 * t) == c ? 1 : 0), (T)0};
    auto bcast = [&](int q) {
        cplx<T> v = {(T)0, (T)0};
#pragma unroll
        for (int t = 0; t < R; ++t)
            if (t == (q >> 6)) v = x[t];
        return cplx<T>{__shfl(v.r, q & 63, 64), __shfl(v.i, q & 63, 64)};
    };
    V col[TRINV_QC][R];
    if (LOWER) {
        for (int q0 = c; q0 < bs - 1; q0 += TRINV_QC) {
#pragma unroll
            for (int u = 0; u < TRINV_QC; ++u)
#pragma unroll
                for (int t = 0; t < R; ++t) {
                    const int q = q0 + u, r = lane + 64 * t;
                    col[u][t] = (q < bs - 1 && r > q && r < bs) ? D[(long)r + (long)q * lda] : V{0, 0};
                }
#pragma unroll
            for (int u = 0; u < TRINV_QC; ++u) {
                const int q = q0 + u;
                if (q >= bs - 1) break;
                const cplx<T> xq = bcast(q);
#pragma unroll
                for (int t = 0; t < R; ++t) x[t] = csub(x[t], cmul(cplx<T>{col[u][t].x, col[u][t].y}, xq));
            }
        }
    } else {
        V dg[TRINV_QC];
        for (int q0 = c; q0 >= 0; q0 -= TRINV_QC) {
#pragma unroll
            for (int u = 0; u < TRINV_QC; ++u) {
                const int q = q0 - u;
                dg[u] = q >= 0 ? D[(long)q + (long)q * lda] : V{1, 0};
#pragma unroll
                for (int t = 0; t < R; ++t) {
                    const int r = lane + 64 * t;
                    col[u][t] = (q >= 0 && r < q) ? D[(long)r + (long)q * lda] : V{0, 0};
                }
            }
#pragma unroll
            for (int u = 0; u < TRINV_QC; ++u) {
                const int q = q0 - u;
                if (q < 0) break;
                // x_q /= U(q, q) on the owning lane, then broadcast
                const cplx<T> dq = cinv(cplx<T>{dg[u].x, dg[u].y});
#pragma unroll
                for (int t = 0; t < R; ++t)
                    if (lane + 64 * t == q) x[t] = cmul(x[t], dq);
                const cplx<T> xq = bcast(q);
#pragma unroll
                for (int t = 0; t < R; ++t) x[t] = csub(x[t], cmul(cplx<T>{col[u][t].x, col[u][t].y}, xq));
            }
        }
    }
#pragma unroll
    for (int t = 0; t < R; ++t) {
        const int r = lane + 64 * t;
        if (r < bs) Inv[(long)r + (long)c * ldi] = V{x[t].r, x[t].i};
    }
}

template <typename T, bool LOWER>
static int trinv(const typename cvec<T>::t *D, long lda, int bs, typename cvec<T>::t *Inv, long ldi, hipStream_t st,
                 int batch = 1, long sD = 0, long sI = 0)
{
    MIDAS_CHECK(bs <= 512, -1, "trinv: block %d above 512", bs);
    const dim3 g((bs + 3) / 4, batch), b(256);
    if (bs <= 128)
        hipLaunchKernelGGL((k_trinv_wave<T, LOWER, 2>), g, b, 0, st, D, lda, sD, bs, Inv, ldi, sI);
    else if (bs <= 256)
        hipLaunchKernelGGL((k_trinv_wave<T, LOWER, 4>), g, b, 0, st, D, lda, sD, bs, Inv, ldi, sI);
    else
        hipLaunchKernelGGL((k_trinv_wave<T, LOWER, 8>), g, b, 0, st, D, lda, sD, bs, Inv, ldi, sI);
    MIDAS_HIP(hipGetLastError());
    return 0;
}

// Copy a (rows x cols) block between column-major arrays.
template <typename T>
__global__ void k_copy2d(const typename cvec<T>::t *S, long lds, typename cvec<T>::t *D, long ldd, int rows, int cols,
                         long sS = 0, long sD = 0)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    int c = blockIdx.y;
    S += blockIdx.z * sS;
    D += blockIdx.z * sD;
    if (r < rows && c < cols) D[(long)r + (long)c * ldd] = S[(long)r + (long)c * lds];
}

template <typename T> static cplx<T> cone() { return {(T)1, (T)0}; }
template <typename T> static cplx<T> cmone() { return {(T)-1, (T)0}; }
template <typename T> static cplx<T> czero() { return {(T)0, (T)0}; }

// Panel [k, k+nb) of an n x n matrix.
template <typename T>
static int panel(typename cvec<T>::t *A, long lda, int n, int k, int nb, int *ipiv, LuWork<T> &w, hipStream_t st)
{
    for (int jb = k; jb < k + nb; jb += LU_IB) {
        int jend = min(jb + LU_IB, k + nb);
        w.jb = jb;
        for (int j = jb; j <= jend; ++j) {
            int rows = n - j;
            int nblk = rows > 0 ? (rows + STEP_ROWS - 1) / STEP_ROWS : 1;
            MIDAS_CHECK(nblk <= w.maxblk, -2, "lu: step workspace too small");
            hipLaunchKernelGGL(k_getf2_step<T>, dim3(nblk, w.batch), dim3(STEP_THREADS), 0, st, A, lda, n, j,
                               j < jend ? jend : j, k, k + nb, ipiv, w);
        }
        MIDAS_HIP(hipGetLastError());
        // rest of the outer panel: columns [jend, k+nb)
        int rc = k + nb - jend;
        if (rc > 0) {
            hipLaunchKernelGGL(k_trsm_lunit_small<T>, dim3((rc + 63) / 64, w.batch), dim3(64), 0, st, A, lda, w.sA, jb,
                               jend - jb, jend, k + nb);
            MIDAS_HIP(hipGetLastError());
            if (n - jend > 0) {
                int r = launch_cgemm<T>(n - jend, rc, jend - jb, cmone<T>(), A + jend + (long)jb * lda, lda, 0,
                                        A + jb + (long)jend * lda, lda, 0, cone<T>(), A + jend + (long)jend * lda,
                                        lda, st, w.batch, w.sA, w.sA, w.sA);
                if (r) return r;
            }
        }
    }
    // inverse of the unit-lower diagonal block L11 of this panel: U12 = L11^-1 A12
    // becomes one GEMM on the main stream, and getrs reuses it
    return trinv<T, true>(A + k + (long)k * lda, lda, nb, (typename cvec<T>::t *)w.linv + (long)(k / w.nb) * w.nb * w.nb,
                          w.nb, st, w.batch, w.sA, w.inv_stride());
}

// U12(:, c0:c1) = L11^-1 A12(:, c0:c1) through the workspace, for panel k.
template <typename T>
static int u12_solve(typename cvec<T>::t *A, long lda, int k, int nb, int c0, int c1, LuWork<T> &w, hipStream_t st)
{
    typedef typename cvec<T>::t V;
    const int nc = c1 - c0;
    if (nc <= 0) return 0;
    const V *Li = (const V *)w.linv + (long)(k / w.nb) * w.nb * w.nb;
    V *tmp = (V *)w.u12;
    const long st12 = (long)w.nb * w.n;
    int r = launch_cgemm<T>(nb, nc, nb, cone<T>(), Li, w.nb, 0, A + k + (long)c0 * lda, lda, 0, czero<T>(), tmp, nb,
                            st, w.batch, w.inv_stride(), w.sA, st12);
    if (r) return r;
    hipLaunchKernelGGL(k_copy2d<T>, dim3((nb + 255) / 256, nc, w.batch), dim3(256), 0, st, tmp, (long)nb,
                       A + k + (long)c0 * lda, lda, nb, nc, st12, w.sA);
    MIDAS_HIP(hipGetLastError());
    return 0;
}

// Look-ahead: the panel of step k+1 is factored on a second stream while the
// main stream applies the rest of step k's trailing update.  Per step, on the
// main stream: wait for panel k, row swaps outside it, U12 = L11^-1 A12,
// update of the next panel's columns (then signal the panel stream), update
// of the remaining columns.  The panel stream only touches the next panel's
// columns, which the remaining update never reads or writes.
template <typename T>
int getrf(typename cvec<T>::t *A, long lda, int n, int *ipiv, LuWork<T> &w, hipStream_t st, TrailingGemm<T> tg)
{
    typedef typename cvec<T>::t V;
    hipStream_t ps = w.pstream;
    MIDAS_CHECK(w.batch == 1 || tg.batch == w.batch, -1, "getrf: trailing GEMM not batched like the factorisation");
    MIDAS_HIP(hipMemsetAsync(w.counter, 0, sizeof(int) * w.batch, st));
    MIDAS_HIP(hipMemsetAsync(w.info, 0, sizeof(int) * w.batch, st));
    MIDAS_HIP(hipEventRecord(w.ev_next, st));
    MIDAS_HIP(hipStreamWaitEvent(ps, w.ev_next, 0));
    int r = panel<T>(A, lda, n, 0, min(w.nb, n), ipiv, w, ps);
    if (r) return r;
    MIDAS_HIP(hipEventRecord(w.ev_panel, ps));
    for (int k = 0; k < n; k += w.nb) {
        int nb = min(w.nb, n - k);
        MIDAS_HIP(hipStreamWaitEvent(st, w.ev_panel, 0));
        // swaps outside the panel: columns [0, k) and [k+nb, n)
        int ncols = k + (n - k - nb);
        if (ncols > 0) {
            hipLaunchKernelGGL(k_laswp<T>, dim3((ncols + 255) / 256, w.batch), dim3(256), 0, st, A, lda, w.sA, ipiv, n,
                               k, k + nb, 0, k, k + nb, n);
            MIDAS_HIP(hipGetLastError());
        }
        int nr = n - k - nb;
        if (nr <= 0) continue;
        // The next panel's columns first (U12 = L11^-1 A12, then A22 -= L21 U12),
        // so that the panel stream can start on them; then the rest.
        const int nb1 = min(w.nb, nr);
        r = u12_solve<T>(A, lda, k, nb, k + nb, k + nb + nb1, w, st);
        if (r) return r;
        r = tg(nr, nb1, nb, A + (k + nb) + (long)k * lda, lda, A + k + (long)(k + nb) * lda, lda,
               A + (k + nb) + (long)(k + nb) * lda, lda, st);
        if (r) return r;
        MIDAS_HIP(hipEventRecord(w.ev_next, st));
        // The rest of the update is enqueued before the panel: a panel is
        // hundreds of small launches, and the host takes milliseconds to issue
        // them; issued first, the big update runs while they are issued.
        if (nr > nb1) {
            r = u12_solve<T>(A, lda, k, nb, k + nb + nb1, n, w, st);
            if (r) return r;
            r = tg(nr, nr - nb1, nb, A + (k + nb) + (long)k * lda, lda, A + k + (long)(k + nb + nb1) * lda, lda,
                   A + (k + nb) + (long)(k + nb + nb1) * lda, lda, st);
            if (r) return r;
        }
        MIDAS_HIP(hipStreamWaitEvent(ps, w.ev_next, 0));
        r = panel<T>(A, lda, n, k + nb, nb1, ipiv, w, ps);
        if (r) return r;
        MIDAS_HIP(hipEventRecord(w.ev_panel, ps));
    }
    MIDAS_HIP(hipStreamWaitEvent(st, w.ev_panel, 0));
    // inverses of the diagonal blocks of U, used by getrs (those of L come from the panels)
    for (int k = 0; k < n; k += w.nb) {
        r = trinv<T, false>(A + k + (long)k * lda, lda, min(w.nb, n - k), (V *)w.uinv + (long)(k / w.nb) * w.nb * w.nb,
                            w.nb, st, w.batch, w.sA, w.inv_stride());
        if (r) return r;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Triangular solves with the factors (getrs), nrhs columns of B (ldb).
// ---------------------------------------------------------------------------
// B <- P B with the final permutation perm (B2[i] = B[perm[i]]); batch on z.
template <typename T>
__global__ void k_permute_rows(const typename cvec<T>::t *B, long ldb, long sB, typename cvec<T>::t *D, long ldd,
                               long sD, const int *perm, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int c = blockIdx.y;
    B += blockIdx.z * sB;
    D += blockIdx.z * sD;
    perm += (long)blockIdx.z * n;
    if (i < n) D[(long)i + (long)c * ldd] = B[(long)perm[i] + (long)c * ldb];
}

// perm from LAPACK interchanges on the device, one thread per matrix
__global__ void k_ipiv_to_perm(const int *ipiv, int *perm, int n, int batch)
{
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    if (z >= batch) return;
    ipiv += (long)z * n;
    perm += (long)z * n;
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int i = 0; i < n; ++i) {
        const int p = ipiv[i];
        if (p != i) {
            const int t = perm[i];
            perm[i] = perm[p];
            perm[p] = t;
        }
    }
}

int ipiv_to_perm_dev(const int *ipiv, int *perm, int n, int batch, hipStream_t st)
{
    hipLaunchKernelGGL(k_ipiv_to_perm, dim3((batch + 63) / 64), dim3(64), 0, st, ipiv, perm, n, batch);
    MIDAS_HIP(hipGetLastError());
    return 0;
}

// Solve with the factors of getrf for every matrix of the batch: B (n x nrhs,
// ldb, stride sB per matrix); tmp needs max(n, nb) x nrhs per matrix.
template <typename T>
int getrs(const typename cvec<T>::t *A, long lda, int n, const int *perm, const LuWork<T> &w,
          typename cvec<T>::t *B, long ldb, int nrhs, typename cvec<T>::t *tmp, hipStream_t st, long sB)
{
    typedef typename cvec<T>::t V;
    const int nb = w.nb, bt = w.batch;
    const long sA = w.sA, sI = w.inv_stride(), sT = (long)max(n, nb) * nrhs;
    if (n <= 0 || nrhs <= 0) return 0;
    MIDAS_CHECK(nrhs <= 65535, -1, "getrs: nrhs too large for one launch");
    hipLaunchKernelGGL(k_permute_rows<T>, dim3((n + 255) / 256, nrhs, bt), dim3(256), 0, st, B, ldb, sB, tmp, (long)n,
                       sT, perm, n);
    hipLaunchKernelGGL(k_copy2d<T>, dim3((n + 255) / 256, nrhs, bt), dim3(256), 0, st, tmp, (long)n, B, ldb, n, nrhs,
                       sT, sB);
    MIDAS_HIP(hipGetLastError());
    // forward: B_k <- Linv_k B_k ; B_{>k} -= L_{>k,k} B_k
    for (int k = 0; k < n; k += nb) {
        int b = min(nb, n - k);
        const V *Li = (const V *)w.linv + (long)(k / nb) * nb * nb;
        int r = launch_cgemm<T>(b, nrhs, b, cone<T>(), Li, nb, 0, B + k, ldb, 0, czero<T>(), tmp, nb, st, bt, sI, sB,
                                sT);
        if (r) return r;
        hipLaunchKernelGGL(k_copy2d<T>, dim3((b + 255) / 256, nrhs, bt), dim3(256), 0, st, tmp, (long)nb, B + k, ldb,
                           b, nrhs, sT, sB);
        MIDAS_HIP(hipGetLastError());
        if (n - k - b > 0) {
            r = launch_cgemm<T>(n - k - b, nrhs, b, cmone<T>(), A + (k + b) + (long)k * lda, lda, 0, B + k, ldb, 0,
                                cone<T>(), B + k + b, ldb, st, bt, sA, sB, sB);
            if (r) return r;
        }
    }
    // backward: B_k <- Uinv_k B_k ; B_{<k} -= U_{<k,k} B_k
    for (int k = ((n - 1) / nb) * nb; k >= 0; k -= nb) {
        int b = min(nb, n - k);
        const V *Ui = (const V *)w.uinv + (long)(k / nb) * nb * nb;
        int r = launch_cgemm<T>(b, nrhs, b, cone<T>(), Ui, nb, 0, B + k, ldb, 0, czero<T>(), tmp, nb, st, bt, sI, sB,
                                sT);
        if (r) return r;
        hipLaunchKernelGGL(k_copy2d<T>, dim3((b + 255) / 256, nrhs, bt), dim3(256), 0, st, tmp, (long)nb, B + k, ldb,
                           b, nrhs, sT, sB);
        MIDAS_HIP(hipGetLastError());
        if (k > 0) {
            r = launch_cgemm<T>(k, nrhs, b, cmone<T>(), A + (long)k * lda, lda, 0, B + k, ldb, 0, cone<T>(), B, ldb, st,
                                bt, sA, sB, sB);
            if (r) return r;
        }
    }
    return 0;
}

// Final row permutation from LAPACK-style interchanges: perm[i] is the
// original row that ends up in position i.
void ipiv_to_perm(const int *ipiv, int n, int *perm)
{
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int i = 0; i < n; ++i) {
        int p = ipiv[i];
        if (p != i) {
            int t = perm[i];
            perm[i] = perm[p];
            perm[p] = t;
        }
    }
}

template <typename T>
int LuWork<T>::alloc(int n_, int nb_, int batch_, long sA_)
{
    n = n_;
    nb = nb_;
    batch = batch_;
    sA = sA_;
    maxblk = (n + STEP_ROWS - 1) / STEP_ROWS + 1;
    MIDAS_HIP(hipMalloc(&pmax, sizeof(T) * maxblk * batch));
    MIDAS_HIP(hipMalloc(&pidx, sizeof(int) * maxblk * batch));
    MIDAS_HIP(hipMalloc(&counter, sizeof(int) * 2 * batch));
    info = counter + batch;
    MIDAS_HIP(hipMalloc(&pinv, sizeof(cplx<T>) * batch));
    MIDAS_HIP(hipMalloc(&linv, sizeof(cplx<T>) * inv_stride() * batch));
    MIDAS_HIP(hipMalloc(&uinv, sizeof(cplx<T>) * inv_stride() * batch));
    MIDAS_HIP(hipMalloc(&u12, sizeof(cplx<T>) * (size_t)nb * n * batch));
    MIDAS_HIP(hipStreamCreate(&pstream));
    MIDAS_HIP(hipEventCreateWithFlags(&ev_panel, hipEventDisableTiming));
    MIDAS_HIP(hipEventCreateWithFlags(&ev_next, hipEventDisableTiming));
    return 0;
}

template <typename T>
void LuWork<T>::release()
{
    hipFree(pmax);
    hipFree(pidx);
    hipFree(counter);
    hipFree(pinv);
    hipFree(linv);
    hipFree(uinv);
    hipFree(u12);
    if (pstream) {
        hipStreamSynchronize(pstream);
        hipStreamDestroy(pstream);
    }
    if (ev_panel) hipEventDestroy(ev_panel);
    if (ev_next) hipEventDestroy(ev_next);
    pstream = nullptr;
    ev_panel = ev_next = nullptr;
    pmax = nullptr;
    pidx = counter = info = nullptr;
    pinv = nullptr;
    linv = uinv = u12 = nullptr;
}

template struct LuWork<double>;
template struct LuWork<float>;
template int getrf<double>(double2 *, long, int, int *, LuWork<double> &, hipStream_t, TrailingGemm<double>);
template int getrf<float>(float2 *, long, int, int *, LuWork<float> &, hipStream_t, TrailingGemm<float>);
template int getrs<double>(const double2 *, long, int, const int *, const LuWork<double> &, double2 *, long, int,
                           double2 *, hipStream_t, long);
template int getrs<float>(const float2 *, long, int, const int *, const LuWork<float> &, float2 *, long, int,
                          float2 *, hipStream_t, long);

}  // namespace midas
