// Complex GEMM C = alpha A B + beta C (no transposes) on the matrix cores of
// gfx950: FP64 (v_mfma_f64_16x16x4_f64) for the FP64 LU's trailing update
// (plain baseline and CBFM block LUs, strided batch) and the refinement
// residuals R = B - Z X with many right-hand sides; FP32
// (v_mfma_f32_16x16x4_f32) for the complex-FP32 LU's other products
// (U12 = L11^-1 A12) and its many-RHS substitutions.
//
// A 512-thread workgroup owns a 128 x 128 tile of C; each of its 8 waves a
// 64 x 32 part as 4 x 2 MFMA tiles of 16 x 16, two accumulators per tile (real
// and imaginary part: Cr += Ar Br - Ai Bi, Ci += Ar Bi + Ai Br, four real
// products per complex multiply-add, the minus sign taken by the FP64
// instruction's negate modifier on its A operand, by a negated register for
// FP32), so 32 independent MFMAs per k-step of 4.
// The 128 accumulator registers and the rest of a lane's state fit in the 256
// registers that two waves per SIMD leave each; one workgroup per CU.  Stages
// of ZK k are staged through LDS as complex values (k-major rows of 128, one
// b128 read per fragment, rows padded by one value so that the column-strided
// stores of B spread over the banks) in two buffers; the next stage's global
// loads are issued into registers before the current stage computes, one
// barrier per stage.  Loads outside the matrices
// read a clamped in-range address and are replaced by zero, so the loop has no
// branches.
#include "midas_common.h"
#include "midas_lu.h"

namespace midas {

typedef double zd4 __attribute__((ext_vector_type(4)));

constexpr int ZK = 8;   // k per LDS stage

// neg bits of the f64 MFMA (the blgp operand on gfx950): 1 negates A
constexpr int NEG_A = 1;

typedef float zf4 __attribute__((ext_vector_type(4)));
template <typename T> struct ZMf;
template <> struct ZMf<double> {
    typedef double2 V;
    typedef zd4 acc;
    __device__ static acc op(double a, double b, acc c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    __device__ static acc op_nega(double a, double b, acc c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, NEG_A);
    }
    __device__ static int row(int lane, int q) { return (lane >> 4) + 4 * q; }
};
template <> struct ZMf<float> {
    typedef float2 V;
    typedef zf4 acc;
    __device__ static acc op(float a, float b, acc c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    __device__ static acc op_nega(float a, float b, acc c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(-a, b, c, 0, 0, 0);
    }
    __device__ static int row(int lane, int q) { return 4 * (lane >> 4) + q; }
};

// Workgroup tile TM x TN, wave part WM x WN, OCC workgroups per CU.
template <typename T, int TM, int TN, int WM, int WN, int OCC>
__global__ __launch_bounds__((TM / WM) * (TN / WN) * 64, OCC) void k_zgemm_nn(
    int m, int n, int k, cplx<T> alpha, const typename ZMf<T>::V *__restrict__ A, long lda, long sA,
    const typename ZMf<T>::V *__restrict__ B, long ldb, long sB, cplx<T> beta, typename ZMf<T>::V *__restrict__ C,
    long ldc, long sC)
{
    typedef typename ZMf<T>::V V;
    typedef typename ZMf<T>::acc ACC;
    constexpr int ZTH = (TM / WM) * (TN / WN) * 64;
    constexpr int ZTA = WM / 16, ZTB = WN / 16;   // a wave's MFMA tiles
    constexpr int ZPA = TM * ZK / ZTH;            // A values per thread per stage
    constexpr int ZPB = TN * ZK / ZTH;            // B values per thread per stage
    static_assert(ZPA * ZTH == TM * ZK && ZPB * ZTH == TN * ZK && ZTH % TM == 0 && ZTH % ZK == 0, "tile shares");
    struct ZStage {
        V a[ZK][TM + 1], b[ZK][TN + 1];
    };
    __shared__ ZStage S[2];
    A += blockIdx.z * sA;
    B += blockIdx.z * sB;
    C += blockIdx.z * sC;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int i0 = blockIdx.x * TM, j0 = blockIdx.y * TN;
    const int wm = (wave % (TM / WM)) * WM, wn = (wave / (TM / WM)) * WN;

    ACC cr[ZTA][ZTB], ci[ZTA][ZTB];
#pragma unroll
    for (int a = 0; a < ZTA; ++a)
#pragma unroll
        for (int b = 0; b < ZTB; ++b) {
            cr[a][b] = ACC{0, 0, 0, 0};
            ci[a][b] = ACC{0, 0, 0, 0};
        }

    // this thread's share of a stage: A row ar_ at k ak_ + (ZTH / TM) i,
    // B column bc_ + (ZTH / ZK) i at k bk_
    const int ar_ = tid % TM, ak_ = tid / TM, bk_ = tid % ZK, bc_ = tid / ZK;
    const bool rin = i0 + ar_ < m;
    const long arow = rin ? i0 + ar_ : m - 1;
    V ga[ZPA], gb[ZPB];
    // loads from clamped addresses; the values outside the matrices are
    // replaced by zero when stored (sstore), so no wait sits between the loads
    // and the products
    auto gload = [&](int k0) {
#pragma unroll
        for (int i = 0; i < ZPA; ++i) {
            const int kk = k0 + ak_ + (ZTH / TM) * i;
            const long kc = kk < k ? kk : k - 1;
            ga[i] = A[arow + kc * lda];
        }
#pragma unroll
        for (int i = 0; i < ZPB; ++i) {
            const int kb = k0 + bk_, c = j0 + bc_ + (ZTH / ZK) * i;
            const long kc = kb < k ? kb : k - 1, cc = c < n ? c : n - 1;
            gb[i] = B[kc + cc * ldb];
        }
    };
    // (selects on the two doubles, not on the double2: a select of aggregates
    // would send the staging registers to scratch memory)
    auto sstore = [&](ZStage &st, int k0) {
#pragma unroll
        for (int i = 0; i < ZPA; ++i) {
            const int kk = k0 + ak_ + (ZTH / TM) * i;
            const bool in = rin && kk < k;
            st.a[ak_ + (ZTH / TM) * i][ar_] = V{in ? ga[i].x : T(0), in ? ga[i].y : T(0)};
        }
#pragma unroll
        for (int i = 0; i < ZPB; ++i) {
            const int kb = k0 + bk_, c = j0 + bc_ + (ZTH / ZK) * i;
            const bool in = kb < k && c < n;
            st.b[bk_][bc_ + (ZTH / ZK) * i] = V{in ? gb[i].x : T(0), in ? gb[i].y : T(0)};
        }
    };
    auto compute = [&](const ZStage &st) {
        const int r = lane & 15;
#pragma unroll
        for (int kk = 0; kk < ZK; kk += 4) {
            const int kr = kk + (lane >> 4);
            V fa[ZTA], fb[ZTB];
#pragma unroll
            for (int t = 0; t < ZTA; ++t) fa[t] = st.a[kr][wm + 16 * t + r];
#pragma unroll
            for (int t = 0; t < ZTB; ++t) fb[t] = st.b[kr][wn + 16 * t + r];
            // ZTA ZTB independent accumulators between two updates of one
#pragma unroll
            for (int a = 0; a < ZTA; ++a)
#pragma unroll
                for (int b = 0; b < ZTB; ++b)
                    cr[a][b] = ZMf<T>::op(fa[a].x, fb[b].x, cr[a][b]);
#pragma unroll
            for (int a = 0; a < ZTA; ++a)
#pragma unroll
                for (int b = 0; b < ZTB; ++b)
                    ci[a][b] = ZMf<T>::op(fa[a].x, fb[b].y, ci[a][b]);
#pragma unroll
            for (int a = 0; a < ZTA; ++a)
#pragma unroll
                for (int b = 0; b < ZTB; ++b)
                    cr[a][b] = ZMf<T>::op_nega(fa[a].y, fb[b].y, cr[a][b]);
#pragma unroll
            for (int a = 0; a < ZTA; ++a)
#pragma unroll
                for (int b = 0; b < ZTB; ++b)
                    ci[a][b] = ZMf<T>::op(fa[a].y, fb[b].x, ci[a][b]);
        }
    };

    const int nst = (k + ZK - 1) / ZK;
    if (nst > 0) {
        gload(0);
        sstore(S[0], 0);
        __syncthreads();
        for (int s = 0; s + 1 < nst; ++s) {
            gload((s + 1) * ZK);
            // the loads issue before the MFMAs and the stores wait after them,
            // so that the global latency hides behind the stage's products
            __builtin_amdgcn_sched_barrier(0);
            compute(S[s & 1]);
            __builtin_amdgcn_sched_barrier(0);
            // S[(s + 1) & 1] was last read before the previous barrier
            sstore(S[(s + 1) & 1], (s + 1) * ZK);
            __syncthreads();
        }
        compute(S[(nst - 1) & 1]);
    }

    // epilogue: lane holds rows ZMf<T>::row(lane, q) of column lane % 16 of each tile
    const bool bz = beta.r == T(0) && beta.i == T(0);
#pragma unroll
    for (int a = 0; a < ZTA; ++a)
#pragma unroll
        for (int b = 0; b < ZTB; ++b)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = i0 + wm + 16 * a + ZMf<T>::row(lane, q);
                const int col = j0 + wn + 16 * b + (lane & 15);
                if (row < m && col < n) {
                    const cplx<T> v = cmul(alpha, cplx<T>{cr[a][b][q], ci[a][b][q]});
                    V *pc = C + (long)row + (long)col * ldc;
                    cplx<T> res = v;
                    if (!bz) {
                        const V c0 = *pc;
                        res = cadd(v, cmul(beta, cplx<T>{c0.x, c0.y}));
                    }
                    *pc = V{res.r, res.i};
                }
            }
}

template <typename T, int TM, int TN, int WM, int WN, int OCC>
static void zgemm_launch(dim3 grd, hipStream_t st, int m, int n, int k, cplx<T> alpha, const void *A, long lda,
                         long sA, const void *B, long ldb, long sB, cplx<T> beta, void *C, long ldc, long sC)
{
    typedef typename ZMf<T>::V V;
    hipLaunchKernelGGL((k_zgemm_nn<T, TM, TN, WM, WN, OCC>), grd, dim3((TM / WM) * (TN / WN) * 64), 0, st, m, n, k,
                       alpha, (const V *)A, lda, sA, (const V *)B, ldb, sB, beta, (V *)C, ldc, sC);
}

template <typename T>
static int gemm_nn(int m, int n, int k, cplx<T> alpha, const void *A, long lda, long sA, const void *B, long ldb,
                   long sB, cplx<T> beta, void *C, long ldc, long sC, int batch, hipStream_t st, int shape)
{
    if (m <= 0 || n <= 0 || batch <= 0) return 0;
    static const int TMS[] = {128, 128, 64, 128, 64};
    static const int TNS[] = {128, 64, 64, 64, 32};
    MIDAS_CHECK(shape >= 0 && shape < 5, -1, "gemm_nn: tile shape %d unknown", shape);
    const int tm = TMS[shape], tn = TNS[shape];
    MIDAS_CHECK((n + tn - 1) / tn <= 65535 && batch <= 65535, -1, "gemm_nn: n %d or batch %d too large", n, batch);
    MIDAS_CHECK(k >= 0 && lda >= m && (k == 0 || ldb >= k) && ldc >= m, -1,
                "gemm_nn: leading dimensions %ld %ld %ld do not fit m %d n %d k %d", lda, ldb, ldc, m, n, k);
    const dim3 grd((m + tm - 1) / tm, (n + tn - 1) / tn, batch);
    switch (shape) {
    case 0: zgemm_launch<T, 128, 128, 64, 32, 1>(grd, st, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC); break;
    case 1: zgemm_launch<T, 128, 64, 64, 32, 2>(grd, st, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC); break;
    case 2: zgemm_launch<T, 64, 64, 32, 32, 4>(grd, st, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC); break;
    case 3: zgemm_launch<T, 128, 64, 32, 32, 2>(grd, st, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC); break;
    default: zgemm_launch<T, 64, 32, 32, 16, 4>(grd, st, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC); break;
    }
    MIDAS_HIP(hipGetLastError());
    return 0;
}

int launch_zgemm_nn(int m, int n, int k, z64 alpha, const void *A, long lda, long sA, const void *B, long ldb, long sB,
                    z64 beta, void *C, long ldc, long sC, int batch, hipStream_t st, int shape)
{
    return gemm_nn<double>(m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, st, shape);
}

int launch_cgemm_nn(int m, int n, int k, c32 alpha, const void *A, long lda, long sA, const void *B, long ldb, long sB,
                    c32 beta, void *C, long ldc, long sC, int batch, hipStream_t st, int shape)
{
    return gemm_nn<float>(m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, st, shape);
}

}  // namespace midas
