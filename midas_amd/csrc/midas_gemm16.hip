// Split-precision trailing update of the complex-FP32 LU on FP16 / BF16
// matrix cores (gfx950 v_mfma_f32_32x32x16_{f16,bf16}).
//
//   C(m x n) -= A(m x k) B(k x n)        C, A, B complex FP32 (interleaved)
//
// A = L21: partial pivoting on |Re|+|Im| bounds every multiplier's parts by
// sqrt(2), so A needs no scaling.  B = U12 is scaled per column by an exact
// power of two (entries <= 1 in magnitude); the epilogue undoes the scale.
// Each operand element x is split as hi = rn16(x), lo = rn16(x - hi):
//   terms = 1 :  A_hi B_hi
//   terms = 3 :  A_hi B_hi + A_hi B_lo + A_lo B_hi        (~FP32-level products)
// Accumulation is FP32 in the MFMA accumulators; the complex product uses 4
// real products (4M) per term pair.
//
// Layout: two small split kernels write the operands once per update as
// half-precision planes (plane = {hi, lo} x {re, im}) in a tile-major order,
// so each 128-row x 16-k stage tile is one contiguous 4 KB run.  The GEMM
// computes C^T = B^T A^T: the MFMA's row operand comes from B, its column
// operand from A, so a lane's output column is a row of C and the epilogue's
// read-modify-write of C is 256-byte contiguous per column.  Workgroup tile
// 128 x 128 (four waves of 64 x 64 = 2 x 2 MFMA tiles of 32 x 32), LDS double
// buffered, one stage = 16 k.  Tiles are ordered so that consecutive
// workgroups of one XCD share a column block of B (L2 reuse).
#include "midas_common.h"
#include "midas_lu.h"
#include "midas_tune.h"

#include <rocblas/rocblas.h>

#include <algorithm>

namespace midas {

typedef float f16v __attribute__((ext_vector_type(16)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef short s8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16;

constexpr int T_M = 128, T_N = 128, T_K = 16, T_ROW = T_K + 8;   // LDS row: 16 halves + 8 pad (48 B)

template <bool BF> struct half_t;
template <> struct half_t<false> {
    __device__ static inline u16 cvt(float x) { return __builtin_bit_cast(u16, (_Float16)x); }
    __device__ static inline float back(u16 x) { return (float)__builtin_bit_cast(_Float16, x); }
};
template <> struct half_t<true> {
    __device__ static inline u16 cvt(float x) { return __builtin_bit_cast(u16, (__bf16)x); }
    __device__ static inline float back(u16 x) { return (float)__builtin_bit_cast(__bf16, x); }
};

template <bool BF> __device__ inline f16v mma32(s8 a, s8 b, f16v c);
template <> __device__ inline f16v mma32<false>(s8 a, s8 b, f16v c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(h8, a), __builtin_bit_cast(h8, b), c, 0, 0, 0);
}
template <> __device__ inline f16v mma32<true>(s8 a, s8 b, f16v c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(b8, a), __builtin_bit_cast(b8, b), c, 0, 0, 0);
}


// plane element offset (in halves): plane pc, row, k with tile-major order
__host__ __device__ inline long plane_off(int pc, int row, int k, int rtiles, int kst)
{
    return ((((long)pc * rtiles + row / T_M) * kst + k / T_K) * T_M + row % T_M) * T_K + k % T_K;
}

// A (m x K, lda) -> planes (of I - A with eye_minus); thread per (row, group of 8 k)
template <bool BF, int P>
__global__ void k_split_a(const float2 *__restrict__ A, long lda, int m, int K, int rtiles, int kst, u16 *__restrict__ out,
                          long sA = 0, long sOut = 0, int eye_minus = 0)
{
    A += blockIdx.z * sA;      // matrix of a batch
    out += blockIdx.z * sOut;
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    const int k0 = blockIdx.y * 8;
    if (row >= rtiles * T_M) return;
    u16 h[2][2][8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        float2 v = (row < m && k0 + t < K) ? A[(long)row + (long)(k0 + t) * lda] : float2{0.f, 0.f};
        if (eye_minus) v = float2{(row == k0 + t && row < m && row < K ? 1.f : 0.f) - v.x, -v.y};
        u16 hr = half_t<BF>::cvt(v.x), hi = half_t<BF>::cvt(v.y);
        h[0][0][t] = hr;
        h[0][1][t] = hi;
        if (P == 2) {
            h[1][0][t] = half_t<BF>::cvt(v.x - half_t<BF>::back(hr));
            h[1][1][t] = half_t<BF>::cvt(v.y - half_t<BF>::back(hi));
        }
    }
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            s8 v;
#pragma unroll
            for (int t = 0; t < 8; ++t) v[t] = (short)h[p][c][t];
            *(s8 *)(out + plane_off(p * 2 + c, row, k0, rtiles, kst)) = v;
        }
}

// B (K x n, ldb) -> per-column power-of-two scale + planes; one wave per
// column, each lane 8 consecutive k of every 512 (K <= SPLIT_B_KMAX).
constexpr int SPLIT_B_KMAX = 2048;
template <bool BF, int P>
__global__ __launch_bounds__(64) void k_split_b(const float2 *__restrict__ B, long ldb, int K, int n, int rtiles,
                                                int kst, u16 *__restrict__ out, float *__restrict__ scale,
                                                long sB = 0, long sOut = 0, long sScale = 0, int eye_minus = 0)
{
    // (eye_minus: the planes of I - B, the identity of B's rows)
    B += blockIdx.y * sB;      // matrix of a batch
    out += blockIdx.y * sOut;
    scale += blockIdx.y * sScale;
    constexpr int G = SPLIT_B_KMAX / 512;
    const int col = blockIdx.x;
    const int t = threadIdx.x;
    const int Kp = kst * T_K;
    float2 v[G][8];
    float mx = 0.f;
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int kk = 512 * g + 8 * t + q;
            v[g][q] = (col < n && kk < K) ? B[(long)kk + (long)col * ldb] : float2{0.f, 0.f};
            if (eye_minus) v[g][q] = float2{(col < n && kk < K && kk == col ? 1.f : 0.f) - v[g][q].x, -v[g][q].y};
            mx = fmaxf(mx, fmaxf(fabsf(v[g][q].x), fabsf(v[g][q].y)));
        }
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    int e = 0;
    if (mx > 0.f && isfinite(mx)) frexpf(mx, &e);
    const float s = ldexpf(1.0f, -e);
    if (t == 0 && col < n) scale[col] = s;
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int k0 = 512 * g + 8 * t;
        if (k0 >= Kp) break;
        u16 h[2][2][8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            float xr = v[g][q].x * s, xi = v[g][q].y * s;
            u16 hr = half_t<BF>::cvt(xr), hi = half_t<BF>::cvt(xi);
            h[0][0][q] = hr;
            h[0][1][q] = hi;
            if (P == 2) {
                h[1][0][q] = half_t<BF>::cvt(xr - half_t<BF>::back(hr));
                h[1][1][q] = half_t<BF>::cvt(xi - half_t<BF>::back(hi));
            }
        }
#pragma unroll
        for (int p = 0; p < P; ++p)
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                s8 w;
#pragma unroll
                for (int q = 0; q < 8; ++q) w[q] = (short)h[p][c][q];
                *(s8 *)(out + plane_off(p * 2 + c, col, k0, rtiles, kst)) = w;
            }
    }
}

// C -= A B from the planes.  TERMS pairs: (hi,hi) [, (B hi, A lo), (B lo, A hi)].
// Stage tiles go straight from global memory to LDS (global_load_lds_dwordx4:
// each wave moves 1 KB of a 4 KB plane tile, no register staging) through a
// ring of NS LDS slots: the copies of the next NS-1 stages are in flight while
// a stage computes, which covers the memory latency with one wave per SIMD.
// LDS rows are 32 B (16 halves), unpadded: the fragment reads (lane -> row
// l%32, 16-byte chunk l/32) and the copies are bank-conflict free.
template <int SPS, int NPL> struct StageLds {
    u16 a[NPL][SPS][T_M][T_K];
    u16 b[NPL][SPS][T_N][T_K];
};

// Four LDS ring slots as distinct objects, so the compiler can tell a slot
// being refilled by the copy engine from the slot being read.
__shared__ __attribute__((aligned(16))) unsigned char g_slot0[32768];
__shared__ __attribute__((aligned(16))) unsigned char g_slot1[32768];
__shared__ __attribute__((aligned(16))) unsigned char g_slot2[32768];
__shared__ __attribute__((aligned(16))) unsigned char g_slot3[32768];
template <int I, typename S> __device__ inline S *ring_slot()
{
    if constexpr (I == 0) return (S *)g_slot0;
    else if constexpr (I == 1) return (S *)g_slot1;
    else if constexpr (I == 2) return (S *)g_slot2;
    else return (S *)g_slot3;
}

// s_waitcnt for vector-memory count <= N only (gfx9 encoding)
#define MIDAS_WAIT_VM(N) __builtin_amdgcn_s_waitcnt(((N) & 15) | (((N) >> 4) << 14) | (7 << 4) | (15 << 8))

template <bool BF, int P, int TERMS>
__global__ __launch_bounds__(256, 1) void k_trail16(int m, int n, int kst, int mt, int nt, int group,
                                                    const u16 *__restrict__ Ap,
                                                    const u16 *__restrict__ Bp, const float *__restrict__ bscale,
                                                    float2 *__restrict__ C, long ldc)
{
    constexpr int NPL = 2 * P;          // planes per operand
    constexpr int SPS = P == 1 ? 2 : 1; // 16-k granules per stage
    constexpr int NS = 4;               // LDS ring slots
    constexpr int IPS = NPL * SPS * 2;  // copy instructions per wave per stage
    typedef StageLds<SPS, NPL> Stage;
    static_assert(sizeof(Stage) <= 32768, "stage larger than an LDS slot");

    // Tile order.  Each XCD (workgroup ids are dealt round-robin to the 8
    // XCDs) walks a contiguous range of the sequence L; L runs over groups of
    // `group` row tiles, all column tiles of a group, row tiles fastest.  The
    // host sizes `group` so that a group's A tiles take about half of the
    // XCD's 4 MB L2 and stay there while its column tiles stream.
    const int total = mt * nt;
    const int id = blockIdx.x;
    const int full = (total / 8) * 8;
    const int L = id < full ? (id % 8) * (full / 8) + id / 8 : id;
    const int grp = L / (group * nt);
    const int rows_g = min(group, mt - grp * group);
    const int li = L - grp * group * nt;
    const int tm = grp * group + li % rows_g, tn = li / rows_g;
    const int m0 = tm * T_M, n0 = tn * T_N;

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // wave-uniform (scalar)
    const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
    const int nst = kst / SPS;

    // real part as two accumulators (Br Ar and Bi Ai, subtracted in the
    // epilogue), so no operand needs a sign flip in the inner loop
    f16v accr[2][2], acci[2][2], accn[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                accr[a][b][q] = 0.f;
                acci[a][b][q] = 0.f;
                accn[a][b][q] = 0.f;
            }

    // One wave copies 1 KB (64 lanes x 16 B) of a plane tile into LDS at
    // M0 + lane*16.  Written as one asm statement that also sets M0, so the
    // copy is invisible to the compiler's wait-count pass (which would
    // otherwise drain every copy in flight before each LDS read); completion
    // is awaited explicitly with MIDAS_WAIT_VM before the slot is read.  The
    // instruction has no register result, so nothing is left in flight in a
    // register.  Nothing else in this kernel uses M0.
    typedef __attribute__((address_space(3))) void lds_void;
    auto copy16 = [&](const u16 *gsrc, u16 *ldst) {
        const unsigned lds = (unsigned)(size_t)(lds_void *)ldst;
        asm volatile("s_mov_b32 m0, %1\n\tglobal_load_lds_dwordx4 %0, off" ::"v"(gsrc), "s"(lds) : "memory");
    };
    auto dma = [&](Stage *dst, int st) {
#pragma unroll
        for (int pc = 0; pc < NPL; ++pc)
#pragma unroll
            for (int g = 0; g < SPS; ++g) {
                const int ks = st * SPS + g;
                // scalar tile base + per-lane 16-byte offset
                const u16 *ga = Ap + ((((long)pc * mt + tm) * kst + ks) * T_M) * T_K + wave * 512;
                const u16 *gb = Bp + ((((long)pc * nt + tn) * kst + ks) * T_N) * T_K + wave * 512;
                copy16(ga + lane * 8, &dst->a[pc][g][0][0] + wave * 512);
                copy16(gb + lane * 8, &dst->b[pc][g][0][0] + wave * 512);
            }
    };
    auto compute = [&](const Stage *src) {
        const int r32 = lane & 31, kc = (lane >> 5) * 8;
#pragma unroll
        for (int g = 0; g < SPS; ++g) {
            // MFMA row operand from B (rows = columns of C), column operand from A;
            // B fragments for the whole granule, A fragments one row tile at a time
            s8 fb[NPL][2];
#pragma unroll
            for (int pc = 0; pc < NPL; ++pc)
#pragma unroll
                for (int t = 0; t < 2; ++t) fb[pc][t] = *(const s8 *)&src->b[pc][g][wn + 32 * t + r32][kc];
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                s8 fa[NPL];
#pragma unroll
                for (int pc = 0; pc < NPL; ++pc) fa[pc] = *(const s8 *)&src->a[pc][g][wm + 32 * a + r32][kc];
                // four independent accumulators per product part (the b tiles x re/im)
#pragma unroll
                for (int tp = 0; tp < TERMS; ++tp) {
                    const int pb = (tp == 2) ? 1 : 0;   // B plane set (hi / lo)
                    const int pa = (tp == 1) ? 1 : 0;   // A plane set
                    // C^T: Dr = Br^T Ar^T - Bi^T Ai^T (accr - accn), Di = Br^T Ai^T + Bi^T Ar^T
#pragma unroll
                    for (int b = 0; b < 2; ++b) {
                        accr[a][b] = mma32<BF>(fb[2 * pb][b], fa[2 * pa], accr[a][b]);
                        acci[a][b] = mma32<BF>(fb[2 * pb][b], fa[2 * pa + 1], acci[a][b]);
                    }
#pragma unroll
                    for (int b = 0; b < 2; ++b) {
                        accn[a][b] = mma32<BF>(fb[2 * pb + 1][b], fa[2 * pa + 1], accn[a][b]);
                        acci[a][b] = mma32<BF>(fb[2 * pb + 1][b], fa[2 * pa], acci[a][b]);
                    }
                }
            }
        }
    };
    // one ring step: wait for stage s, refill the slot freed by stage s-1, compute stage s
    auto step = [&](int s, const Stage *cur, Stage *refill) {
        if (s + NS - 2 < nst) MIDAS_WAIT_VM(IPS * (NS - 2));
        else MIDAS_WAIT_VM(0);
        // plain barrier: this wave's copies of stage s are complete (wait
        // above); a fenced __syncthreads would also drain the copies still
        // in flight for the later stages
        asm volatile("s_barrier" ::: "memory");
        if (s + NS - 1 < nst) dma(refill, s + NS - 1);
        compute(cur);
    };

    // nst is a multiple of NS (the planes are zero-padded)
    Stage *s0 = ring_slot<0, Stage>(), *s1 = ring_slot<1, Stage>(), *s2 = ring_slot<2, Stage>(),
          *s3 = ring_slot<3, Stage>();
    dma(s0, 0);
    dma(s1, 1);
    dma(s2, 2);
    for (int s = 0; s < nst - NS; s += NS) {
        step(s, s0, s3);
        step(s + 1, s1, s0);
        step(s + 2, s2, s1);
        step(s + 3, s3, s2);
    }
    // C of the first row tile is loaded ahead of the last ring pass, so its
    // latency overlaps the last MFMAs
    double pre[2][16];
    {
        const int row = m0 + wm + (lane & 31);
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int col = n0 + wn + 32 * b + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
                pre[b][q] = (row < m && col < n)
                                ? __builtin_nontemporal_load((const double *)(C + (long)row + (long)col * ldc))
                                : 0.0;
            }
        const int s = nst - NS;
        step(s, s0, s3);
        step(s + 1, s1, s0);
        step(s + 2, s2, s1);
        step(s + 3, s3, s2);
    }
    // epilogue: 32x32 output, lane holds column lane%32 of D = row of C;
    // register q holds D row 8*(q/4) + 4*(lane/32) + q%4 = column of C.
    // All loads of a row tile are issued before any use (one wave per SIMD:
    // nothing else hides their latency); C is streamed non-temporally so it
    // does not evict the operand planes other workgroups still read.
    float inv[2][16];
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int col = n0 + wn + 32 * b + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
            inv[b][q] = col < n ? bscale[col] : 1.0f;
        }
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int q = 0; q < 16; ++q) inv[b][q] = 1.0f / inv[b][q];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const int row = m0 + wm + 32 * a + (lane & 31);
        double cv[2][16];
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int col = n0 + wn + 32 * b + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
                if (a == 0) cv[b][q] = pre[b][q];
                else
                    cv[b][q] = (row < m && col < n)
                                   ? __builtin_nontemporal_load((const double *)(C + (long)row + (long)col * ldc))
                                   : 0.0;
            }
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int col = n0 + wn + 32 * b + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
                if (row < m && col < n) {
                    float2 c0 = __builtin_bit_cast(float2, cv[b][q]);
                    const float dr = accr[a][b][q] - accn[a][b][q];
                    float2 c1 = float2{c0.x - dr * inv[b][q], c0.y - acci[a][b][q] * inv[b][q]};
                    __builtin_nontemporal_store(__builtin_bit_cast(double, c1), (double *)(C + (long)row + (long)col * ldc));
                }
            }
    }
}

// ---------------------------------------------------------------------------
// Single-term variant with two workgroups per CU.  With one workgroup per CU
// the read-modify-write of C at the end of each tile and the LDS / barrier
// waits of the main loop leave the matrix cores idle; a second resident
// workgroup fills those gaps.  To fit two: the real part accumulates
// Br Ar + (-Bi) Ai in one set (B carries a negated imaginary plane), so a wave
// holds 2 x 4 accumulator tiles; the -Bi fragment is made in registers (a
// sign-bit flip of each half) rather than read as a fifth plane.
//   * Stages are 16 k deep in a ring of four 16 KB slots (64 KB per
//     workgroup).  At step s the wave holds stage s's fragments in registers
//     (read during step s-1), reads stage s+1's from LDS, and the slot that
//     held stage s is refilled with stage s+4: the LDS reads overlap the
//     matrix-core work instead of preceding it.
//   * LDS rows are 32 B (two 16-byte halves).  The half holding logical half c
//     of row r sits at c ^ x(r), x(r) = bit 2 ^ bit 3 of r, so the 16 lanes of
//     a fragment read that share a bank cycle land on distinct banks (the
//     unswizzled layout measured 2-way conflicts on half of the LDS cycles).
//     The copies to LDS keep their fixed destinations (M0 + 16 lane); each
//     lane fetches the global half that belongs at its destination.
//   * The epilogue of a tile that lies fully inside C uses buffer loads and
//     stores (a scalar descriptor per 32 x 32 block, a scalar offset per
//     column, one 32-bit lane offset), so all loads of a block are in flight
//     together and nothing spills; per-element bounds checks (tiles on the
//     edge of C) made every access a branch and serialised the waits.
// ---------------------------------------------------------------------------
struct Stage2 {
    u16 a[2][T_M][T_K];
    u16 b[2][T_N][T_K];
};
constexpr int TRAIL_NS = 4;   // ring slots; the stage count is padded to a multiple of 4

template <bool BF>
__global__ __launch_bounds__(256, 2) void k_trail16x(int m, int n, int kst, int mt, int nt, int group,
                                                     const u16 *__restrict__ Ap, const u16 *__restrict__ Bp,
                                                     const float *__restrict__ bscale, float2 *__restrict__ C, long ldc,
                                                     long sP = 0, long sScale = 0, long sC = 0)
{
    // matrix of a batch (grid y): its planes at sP halves from the previous
    // one's, its scales at sScale, its C at sC
    Ap += blockIdx.y * sP;
    Bp += blockIdx.y * sP;
    bscale += blockIdx.y * sScale;
    C += blockIdx.y * sC;
    constexpr int IPS = 4;   // copy instructions per wave per stage (one 1 KB part of each of 4 plane tiles)
    constexpr int NS = TRAIL_NS;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
    __shared__ float sinv[T_N];
    __shared__ __attribute__((aligned(16))) Stage2 ring[NS];

    // tile order as in k_trail16
    const int total = mt * nt;
    const int id = blockIdx.x;
    const int full = (total / 8) * 8;
    const int L = id < full ? (id % 8) * (full / 8) + id / 8 : id;
    const int grp = L / (group * nt);
    const int rows_g = min(group, mt - grp * group);
    const int li = L - grp * group * nt;
    const int tm = grp * group + li % rows_g, tn = li / rows_g;
    const int m0 = tm * T_M, n0 = tn * T_N;

    // 1 / column scale of the tile's columns, read by the epilogue from LDS
    // (a global load there would also wait for the C loads in flight); the
    // first ring barrier orders these stores before any read
    if (tid < T_N) sinv[tid] = 1.0f / (n0 + tid < n ? bscale[n0 + tid] : 1.0f);

    f16v accr[2][2], acci[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                accr[a][b][q] = 0.f;
                acci[a][b][q] = 0.f;
            }

    // copies straight to LDS, as in k_trail16 (one asm statement sets M0).
    // Lane l writes bytes [16 l, 16 l + 16) of the wave's 1 KB part: row
    // srow, physical half l & 1, which holds logical half (l & 1) ^ x(srow).
    typedef __attribute__((address_space(3))) void lds_void;
    auto copy16 = [&](const u16 *gsrc, u16 *ldst) {
        const unsigned lds = (unsigned)(size_t)(lds_void *)ldst;
        asm volatile("s_mov_b32 m0, %1\n\tglobal_load_lds_dwordx4 %0, off" ::"v"(gsrc), "s"(lds) : "memory");
    };
    const int srow = wave * 32 + (lane >> 1);
    const int soff = (lane >> 1) * T_K + (((lane & 1) ^ ((srow >> 2) ^ (srow >> 3))) & 1) * 8;
    auto dma = [&](Stage2 *dst, int ks) {
#pragma unroll
        for (int pc = 0; pc < 2; ++pc)
            copy16(Ap + ((((long)pc * mt + tm) * kst + ks) * T_M) * T_K + wave * 512 + soff,
                   &dst->a[pc][0][0] + wave * 512);
#pragma unroll
        for (int pc = 0; pc < 2; ++pc)
            copy16(Bp + ((((long)pc * nt + tn) * kst + ks) * T_N) * T_K + wave * 512 + soff,
                   &dst->b[pc][0][0] + wave * 512);
    };
    // a stage's fragments: lane l reads row l % 32 of each 32-row tile,
    // logical half l / 32 (physical half (l / 32) ^ x(row))
    struct Frag {
        s8 b[2][2];   // [plane][column tile]
        s8 a[2][2];   // [row tile][plane]
    };
    auto fetch = [&](const Stage2 *src, Frag &f) {
        const int r32 = lane & 31, kc = (((lane >> 5) ^ ((r32 >> 2) ^ (r32 >> 3))) & 1) * 8;
#pragma unroll
        for (int pc = 0; pc < 2; ++pc)
#pragma unroll
            for (int t = 0; t < 2; ++t) f.b[pc][t] = *(const s8 *)&src->b[pc][wn + 32 * t + r32][kc];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int pc = 0; pc < 2; ++pc) f.a[a][pc] = *(const s8 *)&src->a[pc][wm + 32 * a + r32][kc];
    };
    auto compute = [&](const Frag &f) {
        // -Bi: the sign bit of every half flipped
        typedef unsigned u4v __attribute__((ext_vector_type(4)));
        s8 nb[2];
#pragma unroll
        for (int t = 0; t < 2; ++t)
            nb[t] = __builtin_bit_cast(s8, __builtin_bit_cast(u4v, f.b[1][t]) ^
                                               (u4v){0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u});
        // C^T: Dr = Br^T Ar^T + (-Bi)^T Ai^T, Di = Br^T Ai^T + Bi^T Ar^T
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                accr[a][b] = mma32<BF>(f.b[0][b], f.a[a][0], accr[a][b]);
                acci[a][b] = mma32<BF>(f.b[0][b], f.a[a][1], acci[a][b]);
            }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                accr[a][b] = mma32<BF>(nb[b], f.a[a][1], accr[a][b]);
                acci[a][b] = mma32<BF>(f.b[1][b], f.a[a][0], acci[a][b]);
            }
    };

    // prologue: stages 0 .. NS-1 in flight; stage 0 landed, its fragments read
    Frag F[2];
#pragma unroll
    for (int u = 0; u < NS; ++u) dma(ring + u, u);
    MIDAS_WAIT_VM(IPS * (NS - 1));
    asm volatile("s_barrier" ::: "memory");
    fetch(ring, F[0]);
    // Step s: wait for this wave's copies of stage s+1; then, with every
    // wave's LDS reads of stage s complete (lgkmcnt(0) before the plain
    // barrier), refill stage s's slot with stage s+NS, read stage s+1's
    // fragments (the last step reads a stale slot, unused) and run stage s.
    auto step = [&](int s, int fc, int snext, int sfree) {
        const int later = min(kst - 1, s + NS - 1) - (s + 1);   // copies issued after stage s+1's
        if (later >= 2) MIDAS_WAIT_VM(IPS * 2);
        else if (later == 1) MIDAS_WAIT_VM(IPS);
        else MIDAS_WAIT_VM(0);
        asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
        if (s + NS < kst) dma(ring + sfree, s + NS);
        fetch(ring + snext, F[fc ^ 1]);
        // keep the fragment reads ahead of the matrix-core work
        __builtin_amdgcn_sched_barrier(0);
        compute(F[fc]);
    };
    for (int s = 0; s < kst; s += NS) {
#pragma unroll
        for (int u = 0; u < NS; ++u) step(s + u, u & 1, (u + 1) % NS, u);
    }

    // Epilogue: four 32 x 32 blocks e (a = e & 1, b = e >> 1).  Lane holds
    // column lane%32 of D (a row of C); register q holds D row
    // 8*(q/4) + 4*(lane/32) + q%4 (a column of C).  C is streamed; two blocks
    // are in flight: each block's loads are issued before the previous block
    // is written.
    if (m0 + T_M <= m && n0 + T_N <= n) {
        const unsigned lob = (unsigned)((lane & 31) * 8) + (unsigned)(lane >> 5) * (unsigned)(4 * ldc * 8);
        const int nrec = (int)(32 * ldc * 8 + 256);
        auto rsrc = [&](int e) {
            return __builtin_amdgcn_make_buffer_rsrc(
                (void *)(C + (long)(n0 + wn + 32 * (e >> 1)) * ldc + m0 + wm + 32 * (e & 1)), 0, nrec, 0x00020000);
        };
        typedef decltype(__builtin_amdgcn_raw_buffer_load_b64(rsrc(0), 0, 0, 0)) b64t;
        auto fload = [&](int e, double *cv) {
            const auto r = rsrc(e);
#pragma unroll
            for (int q = 0; q < 16; ++q)
                cv[q] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(
                                                       r, lob, (8 * (q >> 2) + (q & 3)) * (int)ldc * 8, 0));
        };
        auto fstore = [&](int e, const double *cv, const f16v &dr, const f16v &di) {
            const auto r = rsrc(e);
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const float inv = sinv[wn + 32 * (e >> 1) + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3)];
                const float2 c0 = __builtin_bit_cast(float2, cv[q]);
                const float2 c1 = float2{c0.x - dr[q] * inv, c0.y - di[q] * inv};
                __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(b64t, c1), r, lob,
                                                      (8 * (q >> 2) + (q & 3)) * (int)ldc * 8, 0);
            }
        };
        double cv0[16], cv1[16];
        fload(0, cv0);
        fload(1, cv1);
        __builtin_amdgcn_sched_barrier(0);
        fstore(0, cv0, accr[0][0], acci[0][0]);
        __builtin_amdgcn_sched_barrier(0);
        fload(2, cv0);
        __builtin_amdgcn_sched_barrier(0);
        fstore(1, cv1, accr[1][0], acci[1][0]);
        __builtin_amdgcn_sched_barrier(0);
        fload(3, cv1);
        __builtin_amdgcn_sched_barrier(0);
        fstore(2, cv0, accr[0][1], acci[0][1]);
        __builtin_amdgcn_sched_barrier(0);
        fstore(3, cv1, accr[1][1], acci[1][1]);
        return;
    }
    // a tile on the edge of C: bounds-checked accesses
    const int loff = (lane & 31) + 4 * (lane >> 5) * (int)ldc;
    auto cload = [&](int e, double *cv) {
        const int row = m0 + wm + 32 * (e & 1) + (lane & 31);
        const double *Cb = (const double *)(C + (long)(n0 + wn + 32 * (e >> 1)) * ldc + m0 + wm + 32 * (e & 1));
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int col = n0 + wn + 32 * (e >> 1) + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
            cv[q] = (row < m && col < n) ? __builtin_nontemporal_load(Cb + (long)(8 * (q >> 2) + (q & 3)) * ldc + loff)
                                         : 0.0;
        }
    };
    auto cstore = [&](int e, const double *cv, const f16v &dr, const f16v &di) {
        const int row = m0 + wm + 32 * (e & 1) + (lane & 31);
        double *Cb = (double *)(C + (long)(n0 + wn + 32 * (e >> 1)) * ldc + m0 + wm + 32 * (e & 1));
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int col = n0 + wn + 32 * (e >> 1) + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
            if (row < m && col < n) {
                const float inv = sinv[col - n0];
                const float2 c0 = __builtin_bit_cast(float2, cv[q]);
                const float2 c1 = float2{c0.x - dr[q] * inv, c0.y - di[q] * inv};
                __builtin_nontemporal_store(__builtin_bit_cast(double, c1),
                                            Cb + (long)(8 * (q >> 2) + (q & 3)) * ldc + loff);
            }
        }
    };
    double cv0[16], cv1[16];
    cload(0, cv0);
    cload(1, cv1);
    cstore(0, cv0, accr[0][0], acci[0][0]);
    cload(2, cv0);
    cstore(1, cv1, accr[1][0], acci[1][0]);
    cload(3, cv1);
    cstore(2, cv0, accr[0][1], acci[0][1]);
    cstore(3, cv1, accr[1][1], acci[1][1]);
}

// 16-k granules, rounded up to a multiple of 8 = ring slots x granules per stage
// for both variants (split planes are zero-padded)
static inline int kstages(int k) { return ((k + 8 * T_K - 1) / (8 * T_K)) * 8; }

size_t trail16_scratch_bytes(int mmax, int nmax, int kmax)
{
    const size_t rt_m = (mmax + T_M - 1) / T_M, rt_n = (nmax + T_N - 1) / T_N, kst = kstages(kmax);
    return 2 * 4 * (rt_m * T_M + rt_n * T_N) * kst * T_K + 4 * rt_n * T_N + 256;
}

// Split-operand planes of one matrix for an (m x n x k) update, in bytes
// (a multiple of 256): the matrices of a batch sit this far apart in scratch.
static size_t trail16_matrix_bytes(int m, int n, int k)
{
    return (trail16_scratch_bytes(m, n, k) + 255) & ~(size_t)255;
}

template <bool BF, int P, int TERMS>
static int run_trail(int m, int n, int k, const float2 *A, long lda, const float2 *B, long ldb, float2 *C, long ldc,
                     char *scratch, hipStream_t st, int batch, long stride, bool eye_minus_a, bool eye_minus_b,
                     bool a_ready)
{
    const int mt = (m + T_M - 1) / T_M, nt = (n + T_N - 1) / T_N, kst = kstages(k);
    u16 *Ap = (u16 *)scratch;
    u16 *Bp = Ap + (size_t)2 * P * mt * T_M * kst * T_K;
    float *sc = (float *)(Bp + (size_t)2 * P * nt * T_N * kst * T_K);
    MIDAS_CHECK(batch >= 1 && batch <= 65535 && (batch == 1 || P == 1), -1, "trail16: batch %d (P %d)", batch, P);
    if constexpr (P == 1) {
        // two workgroups per CU (MIDAS_TRAIL_1WG=1 selects the one-workgroup kernel)
        const char *e = getenv("MIDAS_TRAIL_1WG");
        if (!(e && atoi(e) != 0) || batch > 1) {
            MIDAS_CHECK(k <= SPLIT_B_KMAX, -1, "trail16: k = %d above %d", k, SPLIT_B_KMAX);
            // 16-k stages, padded to a multiple of the ring's slots (zero planes)
            const int ks = ((k + T_K - 1) / T_K + TRAIL_NS - 1) / TRAIL_NS * TRAIL_NS;
            u16 *Bq = Ap + (size_t)2 * mt * T_M * ks * T_K;
            float *sq = (float *)(Bq + (size_t)2 * nt * T_N * ks * T_K);
            // per matrix of a batch: planes (A's, then B's, then the scales)
            // of trail16_matrix_bytes each
            const long mb = (long)trail16_matrix_bytes(m, n, k);
            const long sP = mb / 2, sS = mb / 4;
            if (!a_ready)
                hipLaunchKernelGGL((k_split_a<BF, 1>), dim3(mt * T_M / 256 + 1, ks * 2, batch), dim3(256), 0, st, A,
                                   lda, m, k, mt, ks, Ap, stride, sP, (int)eye_minus_a);
            hipLaunchKernelGGL((k_split_b<BF, 1>), dim3(nt * T_N, batch), dim3(64), 0, st, B, ldb, k, n, nt, ks, Bq, sq,
                               stride, sP, sS, (int)eye_minus_b);
            // algorithmic bytes: C read and written, both operands once in FP16
            ProfScope ps(st, PROF_TRAIL, 8.0 * m * n * k * batch, (16.0 * m * n + 4.0 * ((double)m + n) * k) * batch);
            const long a_tile = (long)T_M * ks * T_K * 2 * 2;
            const int group = (int)std::max(2L, std::min(16L, ((long)tune(TUNE_TRAIL_GROUP_KIB) << 10) / a_tile));
            hipLaunchKernelGGL((k_trail16x<BF>), dim3(mt * nt, batch), dim3(256), 0, st, m, n, ks, mt, nt, group, Ap,
                               Bq, sq, C, ldc, sP, sS, stride);
            MIDAS_HIP(hipGetLastError());
            return 0;
        }
    }
    MIDAS_CHECK(kst * T_K <= 512, -1, "trail16: k = %d above 512", k);
    if (!a_ready)
        hipLaunchKernelGGL((k_split_a<BF, P>), dim3(mt * T_M / 256 + 1, kst * 2), dim3(256), 0, st, A, lda, m, k, mt,
                           kst, Ap, 0L, 0L, (int)eye_minus_a);
    hipLaunchKernelGGL((k_split_b<BF, P>), dim3(nt * T_N), dim3(64), 0, st, B, ldb, k, n, nt, kst, Bp, sc, 0L, 0L, 0L,
                       (int)eye_minus_b);
    // live accounting of the GEMM kernel alone: algorithmic flops 8 m n k
    // (complex multiply-add); bytes = C read and written (complex FP32) plus
    // one read of the operand planes (P planes of 2 B per real component)
    ProfScope ps(st, PROF_TRAIL, 8.0 * m * n * k, 16.0 * m * n + 4.0 * P * ((double)m + n) * k);
    // row tiles per group: A tiles of the group in ~2 MB of the XCD's L2
    const long a_tile = (long)T_M * kst * T_K * 2 * P * 2;
    const int group = (int)std::max(2L, std::min(16L, ((long)tune(TUNE_TRAIL_GROUP_KIB) << 10) / a_tile));
    hipLaunchKernelGGL((k_trail16<BF, P, TERMS>), dim3(mt * nt), dim3(256), 0, st, m, n, kst, mt, nt, group, Ap, Bp,
                       sc, C, ldc);
    MIDAS_HIP(hipGetLastError());
    return 0;
}

int launch_trail16(int mode, int m, int n, int k, const float2 *A, long lda, const float2 *B, long ldb, float2 *C,
                   long ldc, void *scratch, hipStream_t st, int batch, long stride, bool eye_minus_a, bool eye_minus_b,
                   bool a_ready)
{
    if (m <= 0 || n <= 0 || k <= 0) return 0;
    MIDAS_CHECK(scratch, -1, "trail16: no scratch");
    char *s = (char *)scratch;
    const bool ea = eye_minus_a, eb = eye_minus_b, ar = a_ready;
    switch (mode) {
    case TRAIL_F16_1: return run_trail<false, 1, 1>(m, n, k, A, lda, B, ldb, C, ldc, s, st, batch, stride, ea, eb, ar);
    case TRAIL_F16_3: return run_trail<false, 2, 3>(m, n, k, A, lda, B, ldb, C, ldc, s, st, batch, stride, ea, eb, false);
    case TRAIL_BF16_1: return run_trail<true, 1, 1>(m, n, k, A, lda, B, ldb, C, ldc, s, st, batch, stride, ea, eb, ar);
    case TRAIL_BF16_3: return run_trail<true, 2, 3>(m, n, k, A, lda, B, ldb, C, ldc, s, st, batch, stride, ea, eb, false);
    }
    MIDAS_CHECK(false, -1, "trail16: bad mode %d", mode);
}

size_t trail16_batch_bytes(int mmax, int nmax, int kmax, int batch)
{
    return trail16_matrix_bytes(mmax, nmax, kmax) * (size_t)batch;
}

// One rocBLAS handle per (host thread, device), created on first use and kept
// for the life of the process; its stream is set before every call.
static rocblas_handle blas_handle()
{
    constexpr int MAXDEV = 64;
    thread_local rocblas_handle h[MAXDEV] = {};
    int d = -1;
    if (hipGetDevice(&d) != hipSuccess || d < 0 || d >= MAXDEV) return nullptr;
    if (!h[d] && rocblas_create_handle(&h[d]) != rocblas_status_success) h[d] = nullptr;
    return h[d];
}

int blas_zgemm_nn(int m, int n, int k, z64 alpha, const void *A, long lda, long sA, const void *B, long ldb, long sB,
                  z64 beta, void *C, long ldc, long sC, int batch, hipStream_t st)
{
    if (m <= 0 || n <= 0 || k <= 0 || batch <= 0) return 0;
    rocblas_handle h = blas_handle();
    MIDAS_CHECK(h, -1, "zgemm: no rocBLAS handle for this device");
    rocblas_status s = rocblas_set_stream(h, st);
    MIDAS_CHECK(s == rocblas_status_success, -1, "zgemm: rocblas_set_stream status %d", (int)s);
    const rocblas_double_complex al(alpha.r, alpha.i), be(beta.r, beta.i);
    const rocblas_operation nt = rocblas_operation_none;
    const auto *a = (const rocblas_double_complex *)A, *b = (const rocblas_double_complex *)B;
    auto *c = (rocblas_double_complex *)C;
    s = batch == 1 ? rocblas_zgemm(h, nt, nt, m, n, k, &al, a, lda, b, ldb, &be, c, ldc)
                   : rocblas_zgemm_strided_batched(h, nt, nt, m, n, k, &al, a, lda, sA, b, ldb, sB, &be, c, ldc, sC,
                                                   batch);
    MIDAS_CHECK(s == rocblas_status_success, -1, "zgemm: rocblas status %d", (int)s);
    return 0;
}

// FP64 trailing update C -= A B: a plain (strided-batched) ZGEMM on the
// hand-written k_zgemm_nn, or on rocBLAS when fp64_gemm_library = 1.
template <>
int TrailingGemm<double>::operator()(int m, int n, int k, const void *A, long lda, const void *B, long ldb, void *C,
                                     long ldc, hipStream_t st) const
{
    if (m <= 0 || n <= 0 || k <= 0 || batch <= 0) return 0;
    ProfScope ps(st, PROF_TRAIL, 8.0 * m * n * k * batch, 16.0 * (2.0 * m * n + (double)m * k + (double)k * n) * batch);
    if (!library || !tune(TUNE_FP64_GEMM_LIBRARY))
        return launch_zgemm_nn(m, n, k, z64{-1.0, 0.0}, A, lda, stride, B, ldb, stride, z64{1.0, 0.0}, C, ldc, stride,
                               batch, st, tune(TUNE_ZGEMM_TILE));
    return blas_zgemm_nn(m, n, k, z64{-1.0, 0.0}, A, lda, stride, B, ldb, stride, z64{1.0, 0.0}, C, ldc, stride, batch,
                         st);
}

template <>
int TrailingGemm<float>::operator()(int m, int n, int k, const void *A, long lda, const void *B, long ldb, void *C,
                                    long ldc, hipStream_t st) const
{
    if (mode == TRAIL_NATIVE) {
        ProfScope ps(st, PROF_TRAIL, 8.0 * m * n * k * batch, 8.0 * (2.0 * m * n + (double)m * k + (double)k * n) * batch);
        return launch_cgemm<float>(m, n, k, c32{-1.f, 0.f}, A, lda, 0, B, ldb, 0, c32{1.f, 0.f}, C, ldc, st, batch,
                                   stride, stride, stride);
    }
    MIDAS_CHECK(batch == 1 || mode == TRAIL_F16_1 || mode == TRAIL_BF16_1, -1,
                "trail16: batched split-precision updates are single-term only");
    return launch_trail16(mode, m, n, k, (const float2 *)A, lda, (const float2 *)B, ldb, (float2 *)C, ldc, scratch,
                          st, batch, stride, eye_minus_a, eye_minus_b, a_ready && batch == 1);
}

}  // namespace midas
