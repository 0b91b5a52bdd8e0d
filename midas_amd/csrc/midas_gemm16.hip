// Split-precision trailing update of the complex-FP32 LU on FP16 / BF16
// matrix cores (gfx950 v_mfma_f32_16x16x32_{f16,bf16}).
//
//   C(m x n) -= A(m x k) B(k x n)        C, A, B complex FP32 (interleaved)
//
// A = L21 has |Re|,|Im| <= 1 (partial pivoting bounds |Re|+|Im| of every
// multiplier by 1).  B = U12 is scaled per column by an exact power of two so
// that its entries are <= 1 in magnitude; the epilogue undoes the scale.
// Each operand element x is split as hi = rn16(x), lo = rn16(x - hi):
//   terms = 1 :  A_hi B_hi
//   terms = 3 :  A_hi B_hi + A_hi B_lo + A_lo B_hi        (~FP32-level products)
// The split is done while staging tiles into LDS, so no converted copies of
// the panels are kept in HBM.  Accumulation is FP32 in the MFMA accumulators;
// the complex product uses 4 real MFMA products (4M) per term pair.
#include "midas_common.h"
#include "midas_lu.h"
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

namespace midas {

typedef float f4v __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef short s8 __attribute__((ext_vector_type(8)));

constexpr int H_M = 64, H_N = 64, H_K = 32, H_PAD = 8;

template <bool BF> struct half_t;
template <> struct half_t<false> {
    typedef _Float16 t;
    __device__ static inline t cvt(float x) { return (_Float16)x; }
    __device__ static inline float back(t x) { return (float)x; }
};
template <> struct half_t<true> {
    typedef __bf16 t;
    __device__ static inline t cvt(float x) { return (__bf16)x; }
    __device__ static inline float back(t x) { return (float)x; }
};

template <bool BF> __device__ inline f4v mma(s8 a, s8 b, f4v c);
template <> __device__ inline f4v mma<false>(s8 a, s8 b, f4v c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(h8, a), __builtin_bit_cast(h8, b), c, 0, 0, 0);
}
template <> __device__ inline f4v mma<true>(s8 a, s8 b, f4v c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8, a), __builtin_bit_cast(b8, b), c, 0, 0, 0);
}

__device__ inline s8 neg8(s8 v)
{
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    u4 u = __builtin_bit_cast(u4, v);
    u ^= (u4){0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
    return __builtin_bit_cast(s8, u);
}

// Per-column power-of-two scale of B: s_j = 2^-e with max(|Re|,|Im|) of column j in [2^(e-1), 2^e).
__global__ void k_colscale(const float2 *B, long ldb, int k, int n, float *scale)
{
    int j = blockIdx.x;
    float m = 0.f;
    for (int r = threadIdx.x; r < k; r += blockDim.x) {
        float2 v = B[(long)r + (long)j * ldb];
        m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
    }
    __shared__ float red[256];
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        int e = 0;
        float mm = red[0];
        if (mm > 0.f && isfinite(mm)) frexpf(mm, &e);
        scale[j] = ldexpf(1.0f, -e);
    }
}

template <bool BF, int TERMS>
__global__ __launch_bounds__(256) void k_trail16(int m, int n, int k, const float2 *__restrict__ A, long lda,
                                                 const float2 *__restrict__ B, long ldb, float2 *__restrict__ C,
                                                 long ldc, const float *__restrict__ bscale)
{
    typedef typename half_t<BF>::t H;
    constexpr int NP = TERMS == 1 ? 1 : 2;   // planes per component (hi, lo)
    // A planes [p][re/im][row][k], B planes [p][re/im][col][k]; k contiguous
    __shared__ __attribute__((aligned(16))) H As[NP][2][H_M][H_K + H_PAD];
    __shared__ __attribute__((aligned(16))) H Bs[NP][2][H_N][H_K + H_PAD];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i0 = blockIdx.x * H_M, j0 = blockIdx.y * H_N;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;

    f4v accr[2][2], acci[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            accr[a][b] = f4v{0, 0, 0, 0};
            acci[a][b] = f4v{0, 0, 0, 0};
        }

    // staging: A tile 64 x 32 complex = 2048 values, 8 per thread; B tile 32 x 64 the same
    float2 ra[8], rb[8];
    float bs[8];
    auto gload = [&](int k0) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            int r = tid & 63, kk = (tid >> 6) + 4 * q;   // A: rows contiguous
            int gi = i0 + r, gk = k0 + kk;
            ra[q] = (gi < m && gk < k) ? A[(long)gi + (long)gk * lda] : float2{0.f, 0.f};
            int kb = tid & 31, c = (tid >> 5) + 8 * q;     // B: k contiguous
            int gj = j0 + c, gkb = k0 + kb;
            rb[q] = (gj < n && gkb < k) ? B[(long)gkb + (long)gj * ldb] : float2{0.f, 0.f};
        }
    };
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        int gj = j0 + (tid >> 5) + 8 * q;
        bs[q] = gj < n ? bscale[gj] : 1.0f;
    }
    auto sstore = [&]() {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            int r = tid & 63, kk = (tid >> 6) + 4 * q;
            H hr = half_t<BF>::cvt(ra[q].x), hi = half_t<BF>::cvt(ra[q].y);
            As[0][0][r][kk] = hr;
            As[0][1][r][kk] = hi;
            if (NP == 2) {
                As[NP - 1][0][r][kk] = half_t<BF>::cvt(ra[q].x - half_t<BF>::back(hr));
                As[NP - 1][1][r][kk] = half_t<BF>::cvt(ra[q].y - half_t<BF>::back(hi));
            }
            int kb = tid & 31, c = (tid >> 5) + 8 * q;
            float xr = rb[q].x * bs[q], xi = rb[q].y * bs[q];
            H br = half_t<BF>::cvt(xr), bi = half_t<BF>::cvt(xi);
            Bs[0][0][c][kb] = br;
            Bs[0][1][c][kb] = bi;
            if (NP == 2) {
                Bs[NP - 1][0][c][kb] = half_t<BF>::cvt(xr - half_t<BF>::back(br));
                Bs[NP - 1][1][c][kb] = half_t<BF>::cvt(xi - half_t<BF>::back(bi));
            }
        }
    };

    gload(0);
    for (int k0 = 0; k0 < k; k0 += H_K) {
        __syncthreads();
        sstore();
        __syncthreads();
        if (k0 + H_K < k) gload(k0 + H_K);
        const int ko = 8 * (lane >> 4);
        s8 fa[NP][2][2], fb[NP][2][2], fbn[NP][2];
#pragma unroll
        for (int p = 0; p < NP; ++p)
#pragma unroll
            for (int t = 0; t < 2; ++t) {
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    fa[p][c][t] = *(const s8 *)&As[p][c][wm + 16 * t + (lane & 15)][ko];
                    fb[p][c][t] = *(const s8 *)&Bs[p][c][wn + 16 * t + (lane & 15)][ko];
                }
                fbn[p][t] = neg8(fb[p][1][t]);
            }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                // term pairs (pa, pb): (0,0) [, (0,1), (1,0)]
#pragma unroll
                for (int tp = 0; tp < TERMS; ++tp) {
                    const int pa = (tp == 2) ? NP - 1 : 0;
                    const int pb = (tp == 1) ? NP - 1 : 0;
                    accr[a][b] = mma<BF>(fa[pa][0][a], fb[pb][0][b], accr[a][b]);
                    accr[a][b] = mma<BF>(fa[pa][1][a], fbn[pb][b], accr[a][b]);
                    acci[a][b] = mma<BF>(fa[pa][0][a], fb[pb][1][b], acci[a][b]);
                    acci[a][b] = mma<BF>(fa[pa][1][a], fb[pb][0][b], acci[a][b]);
                }
            }
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            int col = j0 + wn + 16 * b + (lane & 15);
            if (col >= n) continue;
            float inv = 1.0f / bscale[col];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                int row = i0 + wm + 16 * a + 4 * (lane >> 4) + q;
                if (row < m) {
                    float2 *pc = C + (long)row + (long)col * ldc;
                    float2 c0 = *pc;
                    *pc = float2{c0.x - accr[a][b][q] * inv, c0.y - acci[a][b][q] * inv};
                }
            }
        }
}

int launch_trail16(int mode, int m, int n, int k, const float2 *A, long lda, const float2 *B, long ldb, float2 *C,
                   long ldc, float *scale, hipStream_t st)
{
    if (m <= 0 || n <= 0 || k <= 0) return 0;
    MIDAS_CHECK((n + H_N - 1) / H_N <= 65535, -1, "trail16: n too large");
    hipLaunchKernelGGL(k_colscale, dim3(n), dim3(256), 0, st, B, ldb, k, n, scale);
    dim3 grd((m + H_M - 1) / H_M, (n + H_N - 1) / H_N);
    switch (mode) {
    case TRAIL_F16_1: hipLaunchKernelGGL((k_trail16<false, 1>), grd, dim3(256), 0, st, m, n, k, A, lda, B, ldb, C, ldc, scale); break;
    case TRAIL_F16_3: hipLaunchKernelGGL((k_trail16<false, 3>), grd, dim3(256), 0, st, m, n, k, A, lda, B, ldb, C, ldc, scale); break;
    case TRAIL_BF16_1: hipLaunchKernelGGL((k_trail16<true, 1>), grd, dim3(256), 0, st, m, n, k, A, lda, B, ldb, C, ldc, scale); break;
    case TRAIL_BF16_3: hipLaunchKernelGGL((k_trail16<true, 3>), grd, dim3(256), 0, st, m, n, k, A, lda, B, ldb, C, ldc, scale); break;
    default: MIDAS_CHECK(false, -1, "trail16: bad mode %d", mode);
    }
    MIDAS_HIP(hipGetLastError());
    return 0;
}

template <>
int TrailingGemm<double>::operator()(int m, int n, int k, const void *A, long lda, const void *B, long ldb, void *C,
                                     long ldc, hipStream_t st) const
{
    ProfScope ps(st, PROF_TRAIL, 8.0 * m * n * k, 16.0 * (2.0 * m * n + (double)m * k + (double)k * n));
    return launch_cgemm<double>(m, n, k, z64{-1.0, 0.0}, A, lda, 0, B, ldb, 0, z64{1.0, 0.0}, C, ldc, st);
}

template <>
int TrailingGemm<float>::operator()(int m, int n, int k, const void *A, long lda, const void *B, long ldb, void *C,
                                    long ldc, hipStream_t st) const
{
    ProfScope ps(st, PROF_TRAIL, 8.0 * m * n * k, 8.0 * (2.0 * m * n + (double)m * k + (double)k * n));
    if (mode == TRAIL_NATIVE)
        return launch_cgemm<float>(m, n, k, c32{-1.f, 0.f}, A, lda, 0, B, ldb, 0, c32{1.f, 0.f}, C, ldc, st);
    return launch_trail16(mode, m, n, k, (const float2 *)A, lda, (const float2 *)B, ldb, (float2 *)C, ldc,
                          (float *)scratch, st);
}

}  // namespace midas
