// Complex GEMM on CDNA4 matrix cores (gfx950).
//
//   C = alpha * op(A) * op(B) + beta * C      op in {N, T}   (no conjugation:
//   the CBFM products are plain transposes, Compute_EFields_CBFME.f90:666-669)
//
// Interleaved complex operands, column-major.  The product is formed from four
// real MFMA products per complex multiply-add (4M): Cr += Ar Br - Ai Bi,
// Ci += Ar Bi + Ai Br, with planar (re / im) tiles staged through LDS.
//   double: v_mfma_f64_16x16x4_f64   (ZGEMM with a transposed operand: the CBFM
//           Gram matrices and projections; products without transposes go to
//           k_zgemm_nn, midas_zgemm.hip)
//   float : v_mfma_f32_16x16x4_f32   (CGEMM; complex-FP32 LU trailing update)
// A 256-thread workgroup computes a 64x64 tile; each wave owns a 32x32
// quadrant as 2x2 MFMA 16x16 tiles.  K advances 16 complex values per stage;
// the next stage is prefetched into registers while the current one computes.
// A batch of equal-shape products runs on blockIdx.z with element strides.
#include "midas_common.h"
#include "midas_lu.h"
#include "midas_tune.h"

namespace midas {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename T> struct mfma16;
// Operand layout (both): lane l supplies A(l%16, l/16) and B(l/16, l%16).
// Output layout (measured on gfx950): column l%16; row 4*(l/16) + q for the
// f32 instruction but (l/16) + 4*q for the f64 one.
template <> struct mfma16<double> {
    typedef d4 acc;
    __device__ static inline acc op(double a, double b, acc c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    __device__ static inline int row(int lane, int q) { return (lane >> 4) + 4 * q; }
};
template <> struct mfma16<float> {
    typedef f4 acc;
    __device__ static inline acc op(float a, float b, acc c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    __device__ static inline int row(int lane, int q) { return 4 * (lane >> 4) + q; }
};

constexpr int GB_M = 64, GB_N = 64, GB_K = 16, GPAD = 4;

// Load one complex element of op(X) (rows x cols), zero outside.
template <typename T>
__device__ inline typename cvec<T>::t ldx(const typename cvec<T>::t *X, long ldx_, int transp, int r, int c, int rows,
                                         int cols)
{
    typedef typename cvec<T>::t V;
    if (r >= rows || c >= cols) return V{0, 0};
    return transp ? X[(long)c + (long)r * ldx_] : X[(long)r + (long)c * ldx_];
}

template <typename T>
__global__ __launch_bounds__(256) void k_cgemm(int m, int n, int k, cplx<T> alpha, const typename cvec<T>::t *A,
                                               long lda, long sA, int tA, const typename cvec<T>::t *B, long ldb,
                                               long sB, int tB, cplx<T> beta, typename cvec<T>::t *C, long ldc,
                                               long sC)
{
    typedef typename cvec<T>::t V;
    typedef typename mfma16<T>::acc ACC;
    __shared__ T Ar[GB_K][GB_M + GPAD], Ai[GB_K][GB_M + GPAD];
    __shared__ T Br[GB_K][GB_N + GPAD], Bi[GB_K][GB_N + GPAD];

    A += blockIdx.z * sA;
    B += blockIdx.z * sB;
    C += blockIdx.z * sC;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i0 = blockIdx.x * GB_M, j0 = blockIdx.y * GB_N;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;

    ACC accr[2][2], acci[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            accr[a][b] = ACC{0, 0, 0, 0};
            acci[a][b] = ACC{0, 0, 0, 0};
        }

    // Per-thread load coordinates for one stage (4 complex of A, 4 of B).
    // op(A) is m x k; for tA == 0 rows are contiguous, else k is contiguous.
    V ra[4], rb[4];
    auto gload = [&](int k0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            int ar, ac, br, bc;
            if (!tA) { ar = tid & 63; ac = (tid >> 6) + 4 * q; }
            else { ac = tid & 15; ar = (tid >> 4) + 16 * q; }
            ra[q] = ldx<T>(A, lda, tA, i0 + ar, k0 + ac, m, k);
            if (!tB) { br = tid & 15; bc = (tid >> 4) + 16 * q; }
            else { bc = tid & 63; br = (tid >> 6) + 4 * q; }
            rb[q] = ldx<T>(B, ldb, tB, k0 + br, j0 + bc, k, n);
        }
    };
    auto sstore = [&]() {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            int ar, ac, br, bc;
            if (!tA) { ar = tid & 63; ac = (tid >> 6) + 4 * q; }
            else { ac = tid & 15; ar = (tid >> 4) + 16 * q; }
            Ar[ac][ar] = ra[q].x;
            Ai[ac][ar] = tA == 2 ? -ra[q].y : ra[q].y;
            if (!tB) { br = tid & 15; bc = (tid >> 4) + 16 * q; }
            else { bc = tid & 63; br = (tid >> 6) + 4 * q; }
            Br[br][bc] = rb[q].x;
            Bi[br][bc] = tB == 2 ? -rb[q].y : rb[q].y;
        }
    };

    gload(0);
    for (int k0 = 0; k0 < k; k0 += GB_K) {
        __syncthreads();
        sstore();
        __syncthreads();
        if (k0 + GB_K < k) gload(k0 + GB_K);
#pragma unroll
        for (int kk = 0; kk < GB_K; kk += 4) {
            const int kr = kk + (lane >> 4);
            T far[2], fai[2], fbr[2], fbi[2], fbn[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                far[t] = Ar[kr][wm + 16 * t + (lane & 15)];
                fai[t] = Ai[kr][wm + 16 * t + (lane & 15)];
                fbr[t] = Br[kr][wn + 16 * t + (lane & 15)];
                fbi[t] = Bi[kr][wn + 16 * t + (lane & 15)];
                fbn[t] = -fbi[t];
            }
            // independent accumulators back to back (no MFMA waits on the previous one)
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) accr[a][b] = mfma16<T>::op(far[a], fbr[b], accr[a][b]);
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) acci[a][b] = mfma16<T>::op(far[a], fbi[b], acci[a][b]);
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) accr[a][b] = mfma16<T>::op(fai[a], fbn[b], accr[a][b]);
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) acci[a][b] = mfma16<T>::op(fai[a], fbr[b], acci[a][b]);
        }
    }
    // Epilogue: 16x16 MFMA output, lane holds 4 rows (mfma16<T>::row) of column lane%16.
    const bool bz = (beta.r == (T)0 && beta.i == (T)0);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                int row = i0 + wm + 16 * a + mfma16<T>::row(lane, q);
                int col = j0 + wn + 16 * b + (lane & 15);
                if (row < m && col < n) {
                    cplx<T> v = {accr[a][b][q], acci[a][b][q]};
                    cplx<T> r = cmul(alpha, v);
                    V *pc = C + (long)row + (long)col * ldc;
                    if (!bz) {
                        V c0 = *pc;
                        r = cadd(r, cmul(beta, cplx<T>{c0.x, c0.y}));
                    }
                    *pc = V{r.r, r.i};
                }
            }
}

template <typename T>
int launch_cgemm(int m, int n, int k, cplx<T> alpha, const void *A, long lda, int tA, const void *B, long ldb,
                 int tB, cplx<T> beta, void *C, long ldc, hipStream_t st, int batch, long sA, long sB, long sC)
{
    typedef typename cvec<T>::t V;
    if (m <= 0 || n <= 0 || batch <= 0) return 0;
    // without transposes: the faster k_zgemm_nn kernels (FP64, and FP32 unless
    // cgemm_tile selects this kernel)
    if (tA == 0 && tB == 0) {
        if constexpr (sizeof(T) == 8)
            return launch_zgemm_nn(m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, st,
                                   tune(TUNE_ZGEMM_TILE));
        else if (tune(TUNE_CGEMM_TILE) < 5)
            return launch_cgemm_nn(m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, st,
                                   tune(TUNE_CGEMM_TILE));
    }
    MIDAS_CHECK((n + GB_N - 1) / GB_N <= 65535, -1, "cgemm: n too large for one launch");
    if (k <= 0) {
        // C = beta C only: run with k=0 (loop body skipped)
    }
    dim3 grd((m + GB_M - 1) / GB_M, (n + GB_N - 1) / GB_N, batch);
    hipLaunchKernelGGL(k_cgemm<T>, grd, dim3(256), 0, st, m, n, k, alpha, (const V *)A, lda, sA, tA, (const V *)B,
                       ldb, sB, tB, beta, (V *)C, ldc, sC);
    MIDAS_HIP(hipGetLastError());
    return 0;
}

template int launch_cgemm<double>(int, int, int, z64, const void *, long, int, const void *, long, int, z64, void *,
                                  long, hipStream_t, int, long, long, long);
template int launch_cgemm<float>(int, int, int, c32, const void *, long, int, const void *, long, int, c32, void *,
                                 long, hipStream_t, int, long, long, long);

}  // namespace midas

// ---------------------------------------------------------------------------
// matrix-core lane layout probe (tests)
// ---------------------------------------------------------------------------
namespace midas {
typedef _Float16 h8p __attribute__((ext_vector_type(8)));
typedef __bf16 b8p __attribute__((ext_vector_type(8)));

// A(i,k) = (i + 2k) % 7 - 3 and B(k,j) = (3k + j) % 5 - 2 placed with the
// operand layouts the kernels assume; D read back with the assumed output layout.
__device__ inline float pa(int i, int k) { return (float)((i + 2 * k) % 7 - 3); }
__device__ inline float pb(int k, int j) { return (float)((3 * k + j) % 5 - 2); }

__global__ void k_mfma_probe(int which, double *out)
{
    const int l = threadIdx.x;
    f4 acc = {0, 0, 0, 0};
    if (which == 0) {
        int i = l & 15, k = l >> 4;
        d4 a64 = {0, 0, 0, 0};
        a64 = __builtin_amdgcn_mfma_f64_16x16x4f64((double)pa(i, k), (double)pb(k, l & 15), a64, 0, 0, 0);
        for (int q = 0; q < 4; ++q) out[mfma16<double>::row(l, q) + 16 * (l & 15)] = a64[q];
        return;
    } else if (which == 1) {
        int i = l & 15, k = l >> 4;
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(pa(i, k), pb(k, l & 15), acc, 0, 0, 0);
    } else if (which >= 4) {
        // 32x32x16: lane l supplies A(l%32, 8*(l/32)+t) and B(8*(l/32)+t, l%32);
        // D(8*(q/4) + 4*(l/32) + q%4, l%32) lands in register q
        typedef float f16x __attribute__((ext_vector_type(16)));
        f16x d = {0};
        int i = l & 31, kb = 8 * (l >> 5);
        if (which == 4) {
            h8p a, b;
            for (int t = 0; t < 8; ++t) {
                a[t] = (_Float16)pa(i, kb + t);
                b[t] = (_Float16)pb(kb + t, l & 31);
            }
            d = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, d, 0, 0, 0);
        } else {
            b8p a, b;
            for (int t = 0; t < 8; ++t) {
                a[t] = (__bf16)pa(i, kb + t);
                b[t] = (__bf16)pb(kb + t, l & 31);
            }
            d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, d, 0, 0, 0);
        }
        for (int q = 0; q < 16; ++q) out[(8 * (q >> 2) + 4 * (l >> 5) + (q & 3)) + 32 * (l & 31)] = d[q];
        return;
    } else {
        int i = l & 15, kb = 8 * (l >> 4);
        if (which == 2) {
            h8p a, b;
            for (int t = 0; t < 8; ++t) {
                a[t] = (_Float16)pa(i, kb + t);
                b[t] = (_Float16)pb(kb + t, l & 15);
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, acc, 0, 0, 0);
        } else {
            b8p a, b;
            for (int t = 0; t < 8; ++t) {
                a[t] = (__bf16)pa(i, kb + t);
                b[t] = (__bf16)pb(kb + t, l & 15);
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
        }
    }
    for (int q = 0; q < 4; ++q) out[(4 * (l >> 4) + q) + 16 * (l & 15)] = acc[q];
}

int mfma_probe(int which, double *dout, hipStream_t st)
{
    hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, st, which, dout);
    MIDAS_HIP(hipGetLastError());
    return 0;
}
}  // namespace midas
