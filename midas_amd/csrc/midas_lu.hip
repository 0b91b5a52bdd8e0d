// Dense complex LU with partial pivoting and triangular solves (gfx950).
//
// Replaces ScaLAPACK PZGESV (Compute_EFields_MoM.f90:173, Compute_EFields_CBFME.f90:912)
// on one GPU: right-looking blocked LU, column-major, interleaved complex.
//
//   outer panel  nb columns : factored in inner blocks of IB columns;
//   inner block             : one launch for its IB column steps (k_getf2_block):
//                             rows held in registers, one grid barrier per column
//                             for the pivot search (LAPACK getf2 order); then the
//                             block's interchanges on the other panel columns and
//                             the IBxIB unit-lower solve (k_swap_trsm);
//                             (for matrices too tall for one co-resident grid:
//                             one fused kernel per column, k_getf2_step)
//   inner block update      : GEMM on the rest of the panel;
//   end of a panel          : L11^-1 of its diagonal block (one wave per column);
//   outside the panel       : row swaps (laswp), U12 = L11^-1 A12 as one GEMM with
//                             that inverse, trailing GEMM A22 -= L21 U12.
//
// T = double is the FP64 baseline; T = float holds the complex-FP32 factors of
// the mixed-precision solvers, whose trailing GEMM runs either on FP32 MFMA or
// on the FP16/BF16 split kernels (midas_gemm16.hip).
#include "midas_common.h"
#include "midas_lu.h"
#include "midas_tune.h"

#include <algorithm>

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

// ---------------------------------------------------------------------------
// Inner block [jb, jb+ncols) (ncols <= IB) in one launch.  Thread t of
// workgroup b holds row jb + 256 b + t of the block's columns in registers for
// all IB column steps.  Per column j:
//   local argmax of |Re|+|Im| over rows >= j  ->  publish the candidate
//   row's IB values and, in workgroup 0, row j, and (max, row) with the step
//   tag.  FP32: the rows as 64-bit words (step tag | float bits), so no store
//   needs to complete before the record (tag | row | max, one 64-bit word)
//   goes out; a reader checks the tag of every word it takes.  FP64: the
//   rows, a wait for those stores, max and row, then a separate tag.
//   wait until every workgroup's record of this step is there
//   every workgroup reduces the candidates (first row on ties, as i?amax),
//   takes the pivot row from the winner's published copy, applies the
//   interchange to the rows it holds, and updates its rows below j (scale by
//   1/pivot, rank-1 update of the columns right of j).
// All exchanged data goes through device-scope loads and stores, so no cache
// write-back is needed; the tags grow across launches (host count `base`).
// Published data is double-buffered by column parity: a workgroup can be at
// most one step ahead of any reader.  The other panel columns are interchanged
// afterwards by k_swap_trsm.  The FP32 kernel stays within 64 VGPRs, so its
// waves fit next to the trailing GEMM's (448 registers per lane) instead of
// displacing them.  Workgroups of a matrix wait on each other and must become
// co-resident; a wait that runs out sets w.fault (reported by getrf) and every
// wave still exits.
// ---------------------------------------------------------------------------
constexpr int BLK_T = 256;          // threads = rows per workgroup
constexpr int BLK_MAXWG = 1024;     // workgroups per matrix that may wait on each other
constexpr int BLK_SPIN = 1 << 21;   // polls before a wait gives up (about two seconds)

// 4/8/16-byte loads and stores that are coherent across the XCDs (device-scope
// relaxed atomics)
template <typename X> __device__ inline X ld_coh(const X *p)
{
    if constexpr (sizeof(X) == 4) {
        return __builtin_bit_cast(X, __hip_atomic_load((unsigned *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    } else if constexpr (sizeof(X) == 8) {
        return __builtin_bit_cast(
            X, __hip_atomic_load((unsigned long long *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    } else {
        static_assert(sizeof(X) == 16, "ld_coh: 4, 8 or 16 bytes");
        struct P2 { unsigned long long a, b; } v;
        v.a = __hip_atomic_load((unsigned long long *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        v.b = __hip_atomic_load((unsigned long long *)p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return __builtin_bit_cast(X, v);
    }
}
template <typename X> __device__ inline void st_coh(X *p, X x)
{
    if constexpr (sizeof(X) == 4) {
        __hip_atomic_store((unsigned *)p, __builtin_bit_cast(unsigned, x), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else if constexpr (sizeof(X) == 8) {
        __hip_atomic_store((unsigned long long *)p, __builtin_bit_cast(unsigned long long, x), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    } else {
        static_assert(sizeof(X) == 16, "st_coh: 4, 8 or 16 bytes");
        struct P2 { unsigned long long a, b; } v = __builtin_bit_cast(P2, x);
        __hip_atomic_store((unsigned long long *)p, v.a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store((unsigned long long *)p + 1, v.b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}
// wait until this wave's stores have completed (device-scope stores are then
// visible to every workgroup)
__device__ inline void stores_done() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// tid mod 32, recomputed where it is used: an address formed from it is not
// hoisted out of the column loop and kept live (and spilled) across it
__device__ inline int lane32()
{
    int x = threadIdx.x & 31;
    asm volatile("" : "+v"(x));
    return x;
}

template <typename T> __device__ inline bool pivot_better(T v, int i, T bv, int bi)
{
    return v > bv || (v == bv && i < bi);
}

// Argmax over the wave of (v, i), v >= 0 or v < 0 for no candidate, the
// smallest i on ties when i grows with the lane (a wave's rows do).  FP32: the
// key "float bits + 1" (0 for none) is reduced by DPP row shifts and row
// broadcasts -- VALU only, where lane shuffles (ds_bpermute) queue behind the
// trailing update's LDS fragment reads -- and the first lane holding the
// maximum gives (v, i).  FP64: lane shuffles.
template <typename T> __device__ inline void wave_argmax(T &v, int &i)
{
    if constexpr (sizeof(T) == 4) {
        const unsigned key = v < (T)0 ? 0u : __float_as_uint((float)v) + 1u;
        unsigned m = key;
        auto up = [](unsigned a, unsigned b) { return a > b ? a : b; };
        m = up(m, (unsigned)__builtin_amdgcn_update_dpp(0, (int)m, 0x111, 0xf, 0xf, true));   // row_shr:1
        m = up(m, (unsigned)__builtin_amdgcn_update_dpp(0, (int)m, 0x112, 0xf, 0xf, true));   // row_shr:2
        m = up(m, (unsigned)__builtin_amdgcn_update_dpp(0, (int)m, 0x114, 0xf, 0xf, true));   // row_shr:4
        m = up(m, (unsigned)__builtin_amdgcn_update_dpp(0, (int)m, 0x118, 0xf, 0xf, true));   // row_shr:8
        m = up(m, (unsigned)__builtin_amdgcn_update_dpp((int)m, (int)m, 0x142, 0xa, 0xf, false));   // row_bcast:15
        m = up(m, (unsigned)__builtin_amdgcn_update_dpp((int)m, (int)m, 0x143, 0xc, 0xf, false));   // row_bcast:31
        const unsigned w = __builtin_amdgcn_readlane(m, 63);
        const int lane = __ffsll((long long)__ballot(key == w)) - 1;
        v = __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
        i = __builtin_amdgcn_readlane(i, lane);
    } else {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const T ov = __shfl_xor(v, off, 64);
            const int oi = __shfl_xor(i, off, 64);
            if (pivot_better(ov, oi, v, i)) {
                v = ov;
                i = oi;
            }
        }
    }
}

// argmax over the workgroup of NW waves; every thread gets the result.
template <typename T, int NW = BLK_T / 64> __device__ inline void argmax_block(T &v, int &i, T *rv, int *ri)
{
    wave_argmax(v, i);
    if ((threadIdx.x & 63) == 0) { rv[threadIdx.x >> 6] = v; ri[threadIdx.x >> 6] = i; }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NW; ++q)
        if (pivot_better(rv[q], ri[q], v, i)) { v = rv[q]; i = ri[q]; }
    __syncthreads();
}

template <typename T, int WPE>
__global__ __launch_bounds__(BLK_T, WPE) void k_getf2_block(typename cvec<T>::t *A, long lda, int n, int jb,
                                                            int ncols, int *ipiv, LuWork<T> w, unsigned base)
{
    typedef typename cvec<T>::t V;
    __shared__ V urow[LU_IB];
    __shared__ V urowj[LU_IB];   // FP32: row j's values for the thread holding row p
    __shared__ T rv[BLK_T / 64];
    __shared__ int ri[BLK_T / 64];
    const int tid = threadIdx.x, b = blockIdx.x, nwg = gridDim.x, zb = blockIdx.y;
    A += zb * w.sA;
    ipiv += (long)zb * n;
    V *cand = (V *)w.cand + (long)zb * 2 * w.maxblk * LU_IB;
    V *rowj = (V *)w.rowj + (long)zb * 2 * LU_IB;
    // FP32: the same buffers as tagged 64-bit words, two per complex value
    unsigned long long *candw = (unsigned long long *)w.cand + (long)zb * 2 * w.maxblk * LU_IB * 2;
    unsigned long long *rowjw = (unsigned long long *)w.rowj + (long)zb * 2 * LU_IB * 2;
    T *bmax = w.pmax + (long)zb * 2 * w.maxblk;
    int *bidx = w.pidx + (long)zb * 2 * w.maxblk;
    unsigned *tag = w.tag + (long)zb * w.maxblk;
    const int r0 = jb + b * BLK_T, i = r0 + tid;
    const bool mine = i < n;
    bool waited_out = false;
    // the panel is the critical path: its waves take issue slots first on
    // SIMDs shared with the trailing update
    __builtin_amdgcn_s_setprio(3);

    // a[t]: this thread's row.  The block's columns rotate one place left after
    // every step, so the current column is always t = 0 and every register
    // index is a constant; after IB steps they are back in place.
    V *Ab = A + r0 + (long)jb * lda;   // uniform base: per-thread offsets stay 32-bit
    V a[LU_IB];
#pragma unroll
    for (int t = 0; t < LU_IB; ++t) a[t] = (mine && t < ncols) ? Ab[(long)t * lda + tid] : V{0, 0};

    for (int c = 0; c < LU_IB; ++c) {
        if (c < ncols) {
            const int j = jb + c, par = c & 1;
            const unsigned want = base + (unsigned)c + 1u;
            // ---- local candidate, published -----------------------------------
            T best = (T)-1;
            int bi = 0x7fffffff;
            if (mine && i >= j) {
                best = fabs(a[0].x) + fabs(a[0].y);
                bi = i;
            }
            argmax_block(best, bi, rv, ri);
            bi = __builtin_amdgcn_readfirstlane(bi);   // uniform: keeps address math scalar
            T gv = (T)-1;
            int gi = 0x7fffffff;
            if constexpr (sizeof(T) == 4) {
                // the candidate row and row j as tagged words (step tag | float
                // bits): a reader checks every word's tag, so these stores need
                // no completion wait before the record is published
                // (staged through LDS so that one wave stores each row, a word
                // per lane; row j lies in workgroup 0, thread c)
                if (bi == i) {
#pragma unroll
                    for (int t = 0; t < LU_IB; ++t) urow[t] = a[t];
                }
                if (b == 0 && tid == c) {
#pragma unroll
                    for (int t = 0; t < LU_IB; ++t) urowj[t] = a[t];
                }
                __syncthreads();
                const unsigned long long tg = (unsigned long long)want << 32;
                // wave 0: the candidate's 32 words; wave 1 of workgroup 0: row j's
                if (tid < 2 * LU_IB) {
                    unsigned long long *dst = candw + ((long)par * w.maxblk + b) * LU_IB * 2;
                    const int l = lane32();
                    st_coh(dst + l, tg | __float_as_uint(((float *)urow)[l]));
                } else if (b == 0 && tid >= 64 && tid < 64 + 2 * LU_IB) {
                    unsigned long long *dst = rowjw + par * LU_IB * 2;
                    const int l = lane32();
                    st_coh(dst + l, tg | __float_as_uint(((float *)urowj)[l]));
                }
                __syncthreads();   // urow, urowj are reused below
                // one 64-bit record: step tag (12 bits) | row (20 bits) | max (float bits)
                unsigned long long *rec = (unsigned long long *)w.pmax + (long)zb * 2 * w.maxblk;
                if (tid == 0)
                    st_coh(rec + par * w.maxblk + b,
                           ((unsigned long long)(want & 0xfffu) << 52) |
                               ((unsigned long long)(bi == 0x7fffffff ? 0xfffffu : (unsigned)bi) << 32) |
                               __builtin_bit_cast(unsigned, (float)best));
                // ---- wait for every workgroup's record of this step -----------
                for (int q = tid; q < nwg; q += BLK_T) {
                    int spins = 0;
                    unsigned long long r = ld_coh(rec + par * w.maxblk + q);
                    while (!waited_out && (unsigned)(r >> 52) != (want & 0xfffu)) {
                        if (++spins > BLK_SPIN) {
                            st_coh(w.fault, 1);
                            waited_out = true;
                        }
                        __builtin_amdgcn_s_sleep(1);
                        r = ld_coh(rec + par * w.maxblk + q);
                    }
                    const unsigned ri32 = (unsigned)(r >> 32) & 0xfffffu;
                    const int vi = ri32 == 0xfffffu ? 0x7fffffff : (int)ri32;
                    const T v = (T)__builtin_bit_cast(float, (unsigned)r);
                    if (pivot_better(v, vi, gv, gi)) { gv = v; gi = vi; }
                }
            } else {
                if (bi == i) {
#pragma unroll
                    for (int t = 0; t < LU_IB; ++t) st_coh(cand + ((long)par * w.maxblk + b) * LU_IB + t, a[t]);
                }
                // row j lies in workgroup 0, thread c
                if (b == 0 && tid == c) {
#pragma unroll
                    for (int t = 0; t < LU_IB; ++t) st_coh(rowj + par * LU_IB + t, a[t]);
                }
                stores_done();
                __syncthreads();
                if (tid == 0) {
                    st_coh(bmax + par * w.maxblk + b, best);
                    st_coh(bidx + par * w.maxblk + b, bi);
                    stores_done();
                    st_coh(tag + b, want);
                }
                // ---- wait for every workgroup's tag of this step --------------
                for (int q = tid; q < nwg; q += BLK_T) {
                    int spins = 0;
                    while (!waited_out && (int)(ld_coh(tag + q) - want) < 0) {
                        if (++spins > BLK_SPIN) {
                            st_coh(w.fault, 1);
                            waited_out = true;
                        }
                        __builtin_amdgcn_s_sleep(1);
                    }
                    const T v = ld_coh(bmax + par * w.maxblk + q);
                    const int vi = ld_coh(bidx + par * w.maxblk + q);
                    if (pivot_better(v, vi, gv, gi)) { gv = v; gi = vi; }
                }
            }
            // ---- global pivot -------------------------------------------------
            argmax_block(gv, gi, rv, ri);
            const int p = __builtin_amdgcn_readfirstlane(gi == 0x7fffffff ? j : gi);
            if constexpr (sizeof(T) == 4) {
                // wave 0: the pivot row's 32 words; wave 1 of the workgroup
                // holding row p: row j's (row p takes them)
                const bool holds_p = p != j && (p - jb) / BLK_T == b;
                auto fetch = [&](const unsigned long long *src, float *dst) {
                    int spins = 0;
                    unsigned long long v = ld_coh(src);
                    while (!waited_out && (unsigned)(v >> 32) != want) {
                        if (++spins > BLK_SPIN) {
                            st_coh(w.fault, 1);
                            waited_out = true;
                        }
                        __builtin_amdgcn_s_sleep(1);
                        v = ld_coh(src);
                    }
                    *dst = __uint_as_float((unsigned)v);
                };
                if (tid < 2 * LU_IB) {
                    const unsigned long long *src =
                        p == j ? rowjw + par * LU_IB * 2 : candw + ((long)par * w.maxblk + (p - jb) / BLK_T) * LU_IB * 2;
                    const int l = lane32();
                    fetch(src + l, (float *)urow + l);
                } else if (holds_p && tid >= 64 && tid < 64 + 2 * LU_IB) {
                    const int l = lane32();
                    fetch(rowjw + par * LU_IB * 2 + l, (float *)urowj + l);
                }
                __syncthreads();
                if (holds_p && i == p) {
#pragma unroll
                    for (int t = 0; t < LU_IB; ++t) a[t] = urowj[t];
                }
            } else {
                if (tid < LU_IB)
                    urow[tid] = p == j ? ld_coh(rowj + par * LU_IB + tid)
                                       : ld_coh(cand + ((long)par * w.maxblk + (p - jb) / BLK_T) * LU_IB + tid);
                // row p takes row j's old values (loaded alongside the pivot row)
                if (p != j && i == p) {
#pragma unroll
                    for (int t = 0; t < LU_IB; ++t) a[t] = ld_coh(rowj + par * LU_IB + t);
                }
                __syncthreads();
            }
            if (b == 0 && tid == 0) ipiv[j] = p;
            // row j takes the pivot row
            if (p != j && i == j) {
#pragma unroll
                for (int t = 0; t < LU_IB; ++t) a[t] = urow[t];
            }
            // ---- scale and rank-1 update of the rows below j ------------------
            if (mine && i > j) {
                const V d = urow[0];
                cplx<T> pinv = {(T)0, (T)0};
                if (d.x != (T)0 || d.y != (T)0) pinv = cinv(cplx<T>{d.x, d.y});
                const cplx<T> l = cmul(cplx<T>{a[0].x, a[0].y}, pinv);
                a[0] = V{l.r, l.i};
                const int nright = ncols - c;   // positions 1 .. nright-1: the columns right of j
#pragma unroll
                for (int t = 1; t < LU_IB; ++t)
                    if (t < nright) {
                        const cplx<T> r = csub(cplx<T>{a[t].x, a[t].y}, cmul(l, cplx<T>{urow[t].x, urow[t].y}));
                        a[t] = V{r.r, r.i};
                    }
            }
            // exactly singular: record the first zero pivot (LAPACK INFO), keep going
            if (b == 0 && tid == 0 && urow[0].x == (T)0 && urow[0].y == (T)0) atomicCAS(w.info + zb, 0, j + 1);
            __syncthreads();
        }
        // rotate left by one column
        const V a0 = a[0];
#pragma unroll
        for (int t = 0; t < LU_IB - 1; ++t) a[t] = a[t + 1];
        a[LU_IB - 1] = a0;
    }

    if (mine) {
#pragma unroll
        for (int t = 0; t < LU_IB; ++t)
            if (t < ncols) Ab[(long)t * lda + tid] = a[t];
    }
}

// Inner block whose rows [jb, rend) fit one workgroup (rend - jb <= LOC_T),
// one workgroup per matrix of a batch (grid y): the same column steps as
// k_getf2_block (pivot = largest |re| + |im|, first row on ties, as izamax;
// the same scaling and rank-1 update), with one barrier per column: each
// wave reduces its candidates, and its winning lane leaves (value, row, the
// row's entries) in an LDS slot, next to row j's entries; after the barrier
// every thread picks the winner among the waves' slots and reads the pivot
// row from it.  Slots alternate by column parity, so a slot is rewritten only
// after every thread has passed the next column's barrier.  Then, as
// k_swap_trsm would (the same workgroup, no launch): the block's interchanges,
// in order, on the panel's columns [k, jb) and [jend, k1), and the unit-lower
// solve with L11 (from LDS) on [jend, k1).  It takes the inner blocks of the last rows of
// a factorisation and the diagonal blocks of panels whose pivot search is
// restricted to them.
// NT threads: 512, or 256 for at most 256 rows (four waves of at most 96
// registers fit a CU beside the trailing update's two workgroups; eight
// waves wait for a CU it leaves).
constexpr int LOC_T = 512;
template <typename T, int NT>
__global__ __launch_bounds__(NT, NT == 256 ? 5 : 1) void k_getf2_local(typename cvec<T>::t *A, long lda, long sA,
                                                                       int rend, int jb, int ncols, int *ipiv, long sP,
                                                                       int *info, int k, int k1)
{
    typedef typename cvec<T>::t V;
    constexpr int NW = NT / 64;
    __shared__ V crow[2][NW][LU_IB];   // each wave's candidate row
    __shared__ V jrow[2][LU_IB];       // row j
    __shared__ T cval[2][NW];
    __shared__ int cidx[2][NW];
    __shared__ V l11[LU_IB][LU_IB];    // the block's diagonal 16 x 16, [row][column]
    __shared__ int piv[LU_IB];
    const int tid = threadIdx.x, zb = blockIdx.y, wv = tid >> 6, lane = tid & 63;
    A += zb * sA;
    ipiv += zb * sP;
    const int i = jb + tid;
    const bool mine = i < rend;
    __builtin_amdgcn_s_setprio(3);
    V *Ab = A + jb + (long)jb * lda;
    V a[LU_IB];
#pragma unroll
    for (int t = 0; t < LU_IB; ++t) a[t] = (mine && t < ncols) ? Ab[(long)t * lda + tid] : V{0, 0};
    for (int c = 0; c < LU_IB; ++c) {
        if (c < ncols) {
            const int j = jb + c, par = c & 1;
            T best = (T)-1;
            int bi = 0x7fffffff;
            if (mine && i >= j) {
                best = fabs(a[0].x) + fabs(a[0].y);
                bi = i;
            }
            wave_argmax(best, bi);
            if (i == bi) {
                cval[par][wv] = best;
                cidx[par][wv] = bi;
#pragma unroll
                for (int t = 0; t < LU_IB; ++t) crow[par][wv][t] = a[t];
            } else if (lane == 0 && bi == 0x7fffffff) {
                cval[par][wv] = best;
                cidx[par][wv] = bi;
            }
            if (i == j) {
#pragma unroll
                for (int t = 0; t < LU_IB; ++t) jrow[par][t] = a[t];
            }
            __syncthreads();
            T gv = cval[par][0];
            int gi = cidx[par][0], gw = 0;
#pragma unroll
            for (int q = 1; q < NW; ++q)
                if (pivot_better(cval[par][q], cidx[par][q], gv, gi)) {
                    gv = cval[par][q];
                    gi = cidx[par][q];
                    gw = q;
                }
            const int p = __builtin_amdgcn_readfirstlane(gi == 0x7fffffff ? j : gi);
            const V *urow = gi == 0x7fffffff ? jrow[par] : crow[par][__builtin_amdgcn_readfirstlane(gw)];
            if (tid == 0) {
                ipiv[j] = p;
                piv[c] = p;
            }
            if (p != j) {
                if (i == j) {
#pragma unroll
                    for (int t = 0; t < LU_IB; ++t) a[t] = urow[t];
                } else if (i == p) {
#pragma unroll
                    for (int t = 0; t < LU_IB; ++t) a[t] = jrow[par][t];
                }
            }
            const V d = urow[0];
            if (mine && i > j) {
                cplx<T> pinv = {(T)0, (T)0};
                if (d.x != (T)0 || d.y != (T)0) pinv = cinv(cplx<T>{d.x, d.y});
                const cplx<T> l = cmul(cplx<T>{a[0].x, a[0].y}, pinv);
                a[0] = V{l.r, l.i};
                const int nright = ncols - c;
#pragma unroll
                for (int t = 1; t < LU_IB; ++t)
                    if (t < nright) {
                        const cplx<T> r = csub(cplx<T>{a[t].x, a[t].y}, cmul(l, cplx<T>{urow[t].x, urow[t].y}));
                        a[t] = V{r.r, r.i};
                    }
            }
            if (tid == 0 && d.x == (T)0 && d.y == (T)0) atomicCAS(info + zb, 0, j + 1);
        }
        const V a0 = a[0];
#pragma unroll
        for (int t = 0; t < LU_IB - 1; ++t) a[t] = a[t + 1];
        a[LU_IB - 1] = a0;
    }
    if (mine) {
#pragma unroll
        for (int t = 0; t < LU_IB; ++t)
            if (t < ncols) Ab[(long)t * lda + tid] = a[t];
    }
    const int jend = jb + ncols, nl = jb - k, nr = k1 - jend;
    if (nl + nr <= 0) return;   // workgroup-uniform
    if (tid < ncols) {
#pragma unroll
        for (int t = 0; t < LU_IB; ++t) l11[tid][t] = a[t];
    }
    stores_done();
    __syncthreads();
    // the interchanges, in order (the pivots are in LDS; a row pair at a time
    // keeps the registers to those of the column steps)
    bool moved = false;
#pragma unroll
    for (int r = 0; r < LU_IB; ++r) moved |= r < ncols && piv[r] != jb + r;
    for (int t = tid; t < nl + nr; t += NT) {
        const int c = t < nl ? k + t : jend + (t - nl);
        V *col = A + (long)c * lda;
        if (moved) {
            for (int r = 0; r < ncols; ++r) {
                const int p = piv[r];
                if (p != jb + r) {
                    const V u = col[jb + r], v = col[p];
                    col[jb + r] = v;
                    col[p] = u;
                }
            }
            stores_done();
        }
        if (c < jend) continue;
        cplx<T> x[LU_IB];
#pragma unroll
        for (int r = 0; r < LU_IB; ++r)
            if (r < ncols) x[r] = cplx<T>{col[jb + r].x, col[jb + r].y};
#pragma unroll
        for (int r = 1; r < LU_IB; ++r) {
            if (r >= ncols) continue;
#pragma unroll
            for (int q = 0; q < r; ++q) x[r] = csub(x[r], cmul(cplx<T>{l11[r][q].x, l11[r][q].y}, x[q]));
        }
#pragma unroll
        for (int r = 1; r < LU_IB; ++r)
            if (r < ncols) col[jb + r] = V{x[r].r, x[r].i};
    }
}

// L21 from the staging buffer S (rows x cols, lds) back into the matrix (D;
// null: L21 is in place already), and the largest modulus of its entries into
// *amax (float bits, atomic max: the order of non-negative floats is that of
// their bits).
template <typename T>
__global__ __launch_bounds__(256) void k_copy_amax(const typename cvec<T>::t *S, long lds, long sS,
                                                   typename cvec<T>::t *D, long ldd, long sD, int rows, int cols,
                                                   unsigned *amax)
{
    S += blockIdx.z * sS;   // matrix of a batch
    D += blockIdx.z * sD;
    // a few thousand workgroups, each over a share of rows and columns, so
    // that few atomics meet on *amax
    __shared__ float wm[4];
    float m = 0.0f;
    for (int c = blockIdx.y; c < cols; c += gridDim.y)
        for (int r = blockIdx.x * 256 + threadIdx.x; r < rows; r += gridDim.x * 256) {
            const typename cvec<T>::t v = S[(long)r + (long)c * lds];
            if (D) D[(long)r + (long)c * ldd] = v;
            m = fmaxf(m, (float)sqrt((double)v.x * v.x + (double)v.y * v.y));
        }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = fmaxf(fmaxf(wm[0], wm[1]), fmaxf(wm[2], wm[3]));
        if (m > 0.0f) atomicMax(amax, __float_as_uint(m));
    }
}

// Interchanges ipiv[jb, jend) on panel columns [k, jb) and [jend, k1); then,
// on [jend, k1), forward substitution with the unit-lower block
// L(jb:jend, jb:jend) on rows [jb, jend).  One thread per column.  The
// interchanges are composed first into at most 2 IB (position, source row)
// pairs, so each column issues all its loads before its stores.
template <typename T>
__global__ __launch_bounds__(64) void k_swap_trsm(typename cvec<T>::t *A, long lda, long sA, const int *ipiv, int n,
                                                  int jb, int jend, int k, int k1)
{
    typedef typename cvec<T>::t V;
    __shared__ int pos[2 * LU_IB], src[2 * LU_IB], piv[LU_IB];
    __shared__ int npair;
    A += blockIdx.y * sA;
    ipiv += (long)blockIdx.y * n;
    // the block's pivots in one parallel load (a serial loop would wait on
    // each global load in turn)
    if (threadIdx.x < jend - jb) piv[threadIdx.x] = ipiv[jb + threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0) {
        int m = 0;
        for (int j = jb; j < jend; ++j) {
            const int p = piv[j - jb];
            if (p == j) continue;
            int sj = -1, sp = -1;
            for (int s = 0; s < m; ++s) {
                if (pos[s] == j) sj = s;
                if (pos[s] == p) sp = s;
            }
            if (sj < 0) { pos[m] = j; src[m] = j; sj = m++; }
            if (sp < 0) { pos[m] = p; src[m] = p; sp = m++; }
            const int t = src[sj];
            src[sj] = src[sp];
            src[sp] = t;
        }
        npair = m;
    }
    __syncthreads();
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nl = jb - k;
    if (t >= nl + (k1 - jend)) return;
    const int c = t < nl ? k + t : jend + (t - nl);
    V *col = A + (long)c * lda;
    const int m = npair;
    cplx<T> v[2 * LU_IB];   // (a plain struct: an array of double2 is not kept in registers)
#pragma unroll
    for (int s = 0; s < 2 * LU_IB; ++s)
        if (s < m) {
            const V x = col[src[s]];
            v[s] = cplx<T>{x.x, x.y};
        }
#pragma unroll
    for (int s = 0; s < 2 * LU_IB; ++s)
        if (s < m) col[pos[s]] = V{v[s].r, v[s].i};
    if (c < jend) return;
    const int ib = jend - jb;
    const V *L = A + (long)jb + (long)jb * lda;
    cplx<T> x[LU_IB];
#pragma unroll
    for (int r = 0; r < LU_IB; ++r)
        if (r < ib) x[r] = cplx<T>{col[jb + r].x, col[jb + r].y};
#pragma unroll
    for (int r = 1; r < LU_IB; ++r) {
        if (r >= ib) continue;
#pragma unroll
        for (int q = 0; q < r; ++q) {
            const V l = L[(long)r + (long)q * lda];
            x[r] = csub(x[r], cmul(cplx<T>{l.x, l.y}, x[q]));
        }
    }
#pragma unroll
    for (int r = 1; r < LU_IB; ++r)
        if (r < ib) col[jb + r] = V{x[r].r, x[r].i};
}

// Rank-IB update of the rest of the outer panel after an inner block:
//   A(jend:n, jend:c1) -= L(jend:n, jb:jend) U(jb:jend, jend:c1).
// A stream over the panel's remaining columns (each element read and written
// once): thread = row, its IB multipliers in registers, the workgroup's
// UPD_COLS columns of U in LDS (read as broadcasts).  FP32: kept under 104
// registers (5 waves per SIMD) so its waves fit beside the trailing GEMM's.
// UPD_COLS = 64 amortises each row's multipliers over many columns; for few
// rows (the last panels) 16 columns per workgroup give four times the
// workgroups and a quarter of the serial column batches.
template <typename T, int WPE, int UPD_COLS>
__global__ __launch_bounds__(256, WPE) void k_panel_update(typename cvec<T>::t *A, long lda, long sA, int n, int jb,
                                                           int ib, int jend, int c1)
{
    typedef typename cvec<T>::t V;
    __shared__ cplx<T> u[UPD_COLS][LU_IB];
    __builtin_amdgcn_s_setprio(3);   // panel critical path (see k_getf2_block)
    A += blockIdx.z * sA;
    const int c0 = jend + blockIdx.y * UPD_COLS;
    const int nc = min(UPD_COLS, c1 - c0);
    for (int e = threadIdx.x; e < UPD_COLS * LU_IB; e += 256) {
        const int c = e / LU_IB, t = e % LU_IB;
        V x = V{0, 0};
        if (c < nc && t < ib) x = A[(long)(jb + t) + (long)(c0 + c) * lda];
        u[c][t] = cplx<T>{x.x, x.y};
    }
    __syncthreads();
    const int r = jend + blockIdx.x * 256 + threadIdx.x;
    if (r >= n) return;
    cplx<T> l[LU_IB];
#pragma unroll
    for (int t = 0; t < LU_IB; ++t) {
        V x = V{0, 0};
        if (t < ib) x = A[(long)r + (long)(jb + t) * lda];
        l[t] = cplx<T>{x.x, x.y};
    }
    // columns in batches: all loads of a batch are issued before its stores
    // (a store could alias a later load, so the compiler would not hoist them)
    constexpr int CB = sizeof(T) == 4 ? 16 : 4;
    V *col = A + (long)r + (long)c0 * lda;
    for (int cb = 0; cb < nc; cb += CB) {
        cplx<T> x[CB];
#pragma unroll
        for (int q = 0; q < CB; ++q) {
            V v = V{0, 0};
            if (cb + q < nc) v = col[(long)(cb + q) * lda];
            x[q] = cplx<T>{v.x, v.y};
        }
#pragma unroll
        for (int q = 0; q < CB; ++q)
            if (cb + q < nc) {
                cplx<T> acc = x[q];
#pragma unroll
                for (int t = 0; t < LU_IB; ++t) acc = csub(acc, cmul(l[t], u[cb + q][t]));
                col[(long)(cb + q) * lda] = V{acc.r, acc.i};
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

// Apply the row interchanges ipiv[k0..k1) to columns [c0, c1) (and [c2, c3)),
// one thread per column (256 per workgroup).  The workgroup stages the
// interchanges in LDS first and returns at once when all are identities (the
// rule with a pivot search that keeps the diagonal).
constexpr int LASWP_LDS = 1024;
template <typename T>
__global__ __launch_bounds__(256) void k_laswp(typename cvec<T>::t *A, long lda, long sA, const int *ipiv, long sP,
                                               int k0, int k1, int c0, int c1, int c2, int c3)
{
    typedef typename cvec<T>::t V;
    __shared__ int pv[LASWP_LDS];
    A += blockIdx.y * sA;
    ipiv += blockIdx.y * sP;
    int moved = 0;
    for (int j = threadIdx.x; j < k1 - k0; j += blockDim.x) {
        const int p = ipiv[k0 + j];
        if (j < LASWP_LDS) pv[j] = p;
        moved |= p != k0 + j;
    }
    if (!__syncthreads_or(moved)) return;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int w0 = c1 - c0;
    int c = t < w0 ? c0 + t : c2 + (t - w0);
    if (t >= w0 + (c3 - c2)) return;
    V *col = A + (long)c * lda;
    for (int j = k0; j < k1; ++j) {
        const int p = j - k0 < LASWP_LDS ? pv[j - k0] : ipiv[j];
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
// per QC steps; QC shrinks as R grows so that the lower (panel-stream) FP32
// variant stays within 96 registers and fits beside the trailing update.
// LOWER: unit lower L, x = L^-1 e_c; otherwise upper U with its diagonal,
// x = U^-1 e_c.  Inv is ldi x bs.
template <typename T, bool LOWER, int R>
__device__ __forceinline__ void trinv_wave(const typename cvec<T>::t *D, long lda, int bs, typename cvec<T>::t *Inv,
                                           long ldi)
{
    typedef typename cvec<T>::t V;
    constexpr int TRINV_QC = R >= 8 ? (sizeof(T) == 4 ? 3 : 2) : (R >= 4 ? 4 : 8);
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
template <typename T, bool LOWER, int R>
__global__ __launch_bounds__(256) void k_trinv_wave(const typename cvec<T>::t *D, long lda, long sD, int bs,
                                                    typename cvec<T>::t *Inv, long ldi, long sI, long sD2 = 0,
                                                    long sI2 = 0)
{
    trinv_wave<T, LOWER, R>(D + blockIdx.y * sD + blockIdx.z * sD2, lda, bs, Inv + blockIdx.y * sI + blockIdx.z * sI2,
                            ldi);
}
// Both inverses of one LU-factored diagonal block in one launch: grid z = 0
// the unit-lower L^-1 into Li, z = 1 the upper U^-1 into Ui.
template <typename T, int R>
__global__ __launch_bounds__(256) void k_trinv_lu(const typename cvec<T>::t *D, long lda, long sD, int bs,
                                                  typename cvec<T>::t *Li, typename cvec<T>::t *Ui, long ldi, long sI)
{
    D += blockIdx.y * sD;
    if (blockIdx.z == 0) trinv_wave<T, true, R>(D, lda, bs, Li + blockIdx.y * sI, ldi);
    else trinv_wave<T, false, R>(D, lda, bs, Ui + blockIdx.y * sI, ldi);
}

// Inverses of bs x bs triangular blocks: grid y over the matrices of a batch
// (strides sD, sI), grid z over nblk blocks of each (strides sD2, sI2).
template <typename T, bool LOWER>
static int trinv(const typename cvec<T>::t *D, long lda, int bs, typename cvec<T>::t *Inv, long ldi, hipStream_t st,
                 int batch = 1, long sD = 0, long sI = 0, int nblk = 1, long sD2 = 0, long sI2 = 0)
{
    MIDAS_CHECK(bs <= 512, -1, "trinv: block %d above 512", bs);
    MIDAS_CHECK(nblk >= 1 && nblk <= 65535 && batch <= 65535, -1, "trinv: %d x %d blocks", batch, nblk);
    const dim3 g((bs + 3) / 4, batch, nblk), b(256);
    if (bs <= 128)
        hipLaunchKernelGGL((k_trinv_wave<T, LOWER, 2>), g, b, 0, st, D, lda, sD, bs, Inv, ldi, sI, sD2, sI2);
    else if (bs <= 256)
        hipLaunchKernelGGL((k_trinv_wave<T, LOWER, 4>), g, b, 0, st, D, lda, sD, bs, Inv, ldi, sI, sD2, sI2);
    else
        hipLaunchKernelGGL((k_trinv_wave<T, LOWER, 8>), g, b, 0, st, D, lda, sD, bs, Inv, ldi, sI, sD2, sI2);
    MIDAS_HIP(hipGetLastError());
    return 0;
}

// L^-1 and U^-1 of an LU-factored bs x bs block, one launch for a strided
// batch of matrices (sD, sI).
template <typename T>
static int trinv_lu(const typename cvec<T>::t *D, long lda, int bs, typename cvec<T>::t *Li, typename cvec<T>::t *Ui,
                    long ldi, hipStream_t st, int batch = 1, long sD = 0, long sI = 0)
{
    MIDAS_CHECK(bs <= 512 && batch <= 65535, -1, "trinv: block %d above 512 or batch %d", bs, batch);
    const dim3 g((bs + 3) / 4, batch, 2), b(256);
    if (bs <= 128)
        hipLaunchKernelGGL((k_trinv_lu<T, 2>), g, b, 0, st, D, lda, sD, bs, Li, Ui, ldi, sI);
    else if (bs <= 256)
        hipLaunchKernelGGL((k_trinv_lu<T, 4>), g, b, 0, st, D, lda, sD, bs, Li, Ui, ldi, sI);
    else
        hipLaunchKernelGGL((k_trinv_lu<T, 8>), g, b, 0, st, D, lda, sD, bs, Li, Ui, ldi, sI);
    MIDAS_HIP(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// Inverses of the triangular factors of a bs x bs diagonal block by recursive
// doubling.  32 x 32 leaves are inverted one per wave (lane c carries column c
// of the leaf's inverse in registers; the leaf's columns are read with
// wave-uniform addresses).  Each level then merges pairs of b x b inverses
// into 2b x 2b ones,
//   L = [A 0; C B]:  L^-1 = [A^-1 0; -B^-1 C A^-1  B^-1]
//   U = [A E; 0 G]:  U^-1 = [A^-1  -A^-1 E G^-1; 0 G^-1]
// as two small products (T = C A^-1 then -B^-1 T; T = E G^-1 then -A^-1 T),
// every merge of the level, both factors and every matrix of a batch in one
// launch per product: 1 + 2 log2(bs / 32) launches in all, where one wave per
// column of the inverse walking the whole triangle took bs dependent steps and
// read the triangle once per column.  The blocks opposite the inverses'
// triangles are written as zeros.
// ---------------------------------------------------------------------------
constexpr int TI_LEAF = 32;

// (FP32 within 96 registers: it runs beside the trailing update)
template <typename T>
__global__ __launch_bounds__(64, sizeof(T) == 4 ? 5 : 1) void k_trinv_leaf(const typename cvec<T>::t *__restrict__ D,
                                                                           long lda, long sD, int bs,
                                                                           typename cvec<T>::t *Li,
                                                                           typename cvec<T>::t *Ui, long ldi, long sI)
{
    typedef typename cvec<T>::t V;
    V *Inv = blockIdx.z ? Ui : Li;
    if (!Inv) return;
    const int p = blockIdx.x * TI_LEAF, b = min(TI_LEAF, bs - p);
    D += blockIdx.y * sD + p + (long)p * lda;
    Inv += blockIdx.y * sI + p + (long)p * ldi;
    // lane: column c of the leaf's inverse, rows [16 h, 16 h + 16) of it
    // (lanes of columns >= b compute along and store nothing)
    const int lane = threadIdx.x, c = lane & 31, h = lane >> 5;
    // the factor's triangle staged in LDS without its diagonal (zero
    // elsewhere and outside the b x b block), U's diagonal apart (one past b)
    __shared__ cplx<T> sl[TI_LEAF][TI_LEAF];   // [column][row]
    __shared__ cplx<T> sd[TI_LEAF];
    const bool lower = blockIdx.z == 0;
    for (int e = lane; e < TI_LEAF * TI_LEAF; e += 64) {
        const int r = e % TI_LEAF, q = e / TI_LEAF;
        const bool in = r < b && q < b && (lower ? r > q : r < q);
        const V v = in ? D[r + (long)q * lda] : V{0, 0};
        sl[q][r] = cplx<T>{v.x, v.y};
    }
    if (lane < TI_LEAF) {
        const V d = lane < b ? D[lane + (long)lane * lda] : V{1, 0};
        sd[lane] = cplx<T>{d.x, d.y};
    }
    __syncthreads();
    cplx<T> x[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) x[j] = cplx<T>{(T)(16 * h + j == c ? 1 : 0), (T)0};
    // x_q (final once step q begins) from the lane holding row q of column c
    auto xq_of = [&](int q) {
        const cplx<T> v = x[q & 15];
        const int src = c + 32 * (q >> 4);
        return cplx<T>{__shfl(v.r, src, 64), __shfl(v.i, src, 64)};
    };
    if (lower) {
        // unit lower: x_r -= L(r, q) x_q (r > q), q ascending
#pragma unroll
        for (int q = 0; q < TI_LEAF - 1; ++q) {
            const cplx<T> xq = xq_of(q);
#pragma unroll
            for (int j = 0; j < 16; ++j) x[j] = csub(x[j], cmul(sl[q][16 * h + j], xq));
        }
    } else {
        // upper: x_q /= U(q, q), then x_r -= U(r, q) x_q (r < q), q descending
#pragma unroll
        for (int q = TI_LEAF - 1; q >= 0; --q) {
            const cplx<T> xq = cmul(xq_of(q), cinv(sd[q]));
            if (h == (q >> 4)) x[q & 15] = xq;
#pragma unroll
            for (int j = 0; j < 16; ++j) x[j] = csub(x[j], cmul(sl[q][16 * h + j], xq));
        }
    }
    if (c < b) {
#pragma unroll
        for (int j = 0; j < 16; ++j)
            if (16 * h + j < b) Inv[16 * h + j + (long)c * ldi] = V{x[j].r, x[j].i};
    }
}

// Merges at level b (pairs of b x b inverses at p = 2 m b): stage 0 forms T
// for every merge m into Tw (per merge at m b^2, ld = rows of T), stage 1 the
// off-diagonal block of the merged inverse and the zero block opposite it.
// Grid x: merges x (ceil(b / 32))^2 tiles of 32 x 32, y: matrices, z: L / U.
template <typename T>
__global__ __launch_bounds__(256) void k_trinv_merge(const typename cvec<T>::t *__restrict__ D, long lda, long sD,
                                                     int bs, int b, int stage, typename cvec<T>::t *Li,
                                                     typename cvec<T>::t *Ui, long ldi, long sI,
                                                     typename cvec<T>::t *Tw, long sT, long sTz)
{
    typedef typename cvec<T>::t V;
    V *Inv = blockIdx.z ? Ui : Li;
    if (!Inv) return;
    const int tpm = (b + 31) / 32;
    const int m = blockIdx.x / (tpm * tpm), t = blockIdx.x % (tpm * tpm);
    const int p = 2 * m * b, b1 = b, b2 = min(b, bs - p - b);
    if (b2 <= 0) return;
    D += blockIdx.y * sD;
    Inv += blockIdx.y * sI;
    V *Tm = Tw + blockIdx.y * sT + blockIdx.z * sTz + (long)m * b * b;
    // Out (M x N) = alpha P (M x K) Q (K x N)
    int M, N, K;
    const V *P, *Q;
    long ldp, ldq, ldo;
    V *O, *Z;
    T alpha;
    if (blockIdx.z == 0) {
        M = b2;
        N = b1;
        Z = Inv + p + (long)(p + b1) * ldi;   // b1 x b2 zeros above the merged block's diagonal
        if (stage == 0) {
            K = b1, P = D + (p + b1) + (long)p * lda, ldp = lda, Q = Inv + p + (long)p * ldi, ldq = ldi;
            O = Tm, ldo = b2, alpha = 1;
        } else {
            K = b2, P = Inv + (p + b1) + (long)(p + b1) * ldi, ldp = ldi, Q = Tm, ldq = b2;
            O = Inv + (p + b1) + (long)p * ldi, ldo = ldi, alpha = -1;
        }
    } else {
        M = b1;
        N = b2;
        Z = Inv + (p + b1) + (long)p * ldi;   // b2 x b1 zeros below it
        if (stage == 0) {
            K = b2, P = D + p + (long)(p + b1) * lda, ldp = lda, Q = Inv + (p + b1) + (long)(p + b1) * ldi, ldq = ldi;
            O = Tm, ldo = b1, alpha = 1;
        } else {
            K = b1, P = Inv + p + (long)p * ldi, ldp = ldi, Q = Tm, ldq = b1;
            O = Inv + p + (long)(p + b1) * ldi, ldo = ldi, alpha = -1;
        }
    }
    const int i0 = (t % tpm) * 32, j0 = (t / tpm) * 32;
    const int tid = threadIdx.x, row = tid & 31, jc = tid >> 5;
    if (stage == 1) {
        // zero block: N x M, tile rows [j0, j0 + 32), columns [i0, i0 + 32)
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int zr = j0 + row, zc = i0 + jc + 8 * u;
            if (zr < N && zc < M) Z[zr + (long)zc * ldi] = V{0, 0};
        }
    }
    if (i0 >= M || j0 >= N) return;   // block-uniform
    __shared__ cplx<T> sp[32][33], sq[32][33];
    cplx<T> acc[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) acc[u] = cplx<T>{0, 0};
    // the next 32-deep slices of P and Q are loaded while the current one is
    // multiplied (one load latency per product, not one per slice)
    V na[4], nq[4];
    auto load = [&](int k0) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int r = row, c = jc + 8 * u;
            na[u] = (i0 + r < M && k0 + c < K) ? P[(i0 + r) + (long)(k0 + c) * ldp] : V{0, 0};
            nq[u] = (k0 + r < K && j0 + c < N) ? Q[(k0 + r) + (long)(j0 + c) * ldq] : V{0, 0};
        }
    };
    load(0);
    for (int k0 = 0; k0 < K; k0 += 32) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            sp[jc + 8 * u][row] = cplx<T>{na[u].x, na[u].y};
            sq[jc + 8 * u][row] = cplx<T>{nq[u].x, nq[u].y};
        }
        __syncthreads();
        if (k0 + 32 < K) load(k0 + 32);
#pragma unroll 8
        for (int kk = 0; kk < 32; ++kk) {
            const cplx<T> a = sp[kk][row];
#pragma unroll
            for (int u = 0; u < 4; ++u) acc[u] = cadd(acc[u], cmul(a, sq[jc + 8 * u][kk]));
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int r = i0 + row, c = j0 + jc + 8 * u;
        if (r < M && c < N) O[r + (long)c * ldo] = V{alpha * acc[u].r, alpha * acc[u].i};
    }
}

// Recursive doubling pays from 512 columns on: below, its 1 + 2 log2(bs / 32)
// launches (each waiting for room beside the trailing update) cost more than
// the one launch of the column-per-wave inverse.
static bool use_trinv_rec(int bs) { return bs > 256; }

// Elements of the product scratch per factor and matrix for a bs x bs block.
static long trinv_rec_scratch(int bs)
{
    long need = 0;
    for (int b = TI_LEAF; b < bs; b *= 2) {
        const int nm = (bs - b + 2 * b - 1) / (2 * b);
        need = std::max(need, (long)(nm - 1) * b * b + (long)b * std::min(b, bs - 2 * (nm - 1) * b - b));
    }
    return need;
}

// L^-1 (Li, unless null) and U^-1 (Ui, unless null) of the LU-factored
// bs x bs block D by recursive doubling; scr: trinv_rec_scratch(bs) elements
// per factor, the L part first, at stride sT per matrix.
template <typename T>
static int trinv_rec(const typename cvec<T>::t *D, long lda, int bs, typename cvec<T>::t *Li,
                     typename cvec<T>::t *Ui, long ldi, typename cvec<T>::t *scr, hipStream_t st, int batch, long sD,
                     long sI, long sT)
{
    MIDAS_CHECK(bs >= 1 && batch >= 1 && batch <= 65535, -1, "trinv: block %d, batch %d", bs, batch);
    const int nz = Ui ? 2 : 1;
    hipLaunchKernelGGL(k_trinv_leaf<T>, dim3((bs + TI_LEAF - 1) / TI_LEAF, batch, nz), dim3(64), 0, st, D, lda, sD, bs,
                       Li, Ui, ldi, sI);
    MIDAS_HIP(hipGetLastError());
    const long sTz = trinv_rec_scratch(bs);
    for (int b = TI_LEAF; b < bs; b *= 2) {
        const int nm = (bs - b + 2 * b - 1) / (2 * b), tpm = (b + 31) / 32;
        for (int stage = 0; stage < 2; ++stage) {
            hipLaunchKernelGGL(k_trinv_merge<T>, dim3(nm * tpm * tpm, batch, nz), dim3(256), 0, st, D, lda, sD, bs, b,
                               stage, Li, Ui, ldi, sI, scr, sT, sTz);
            MIDAS_HIP(hipGetLastError());
        }
    }
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

// bytes (a multiple of 4) of device memory set to zero by a kernel
__global__ void k_zero_words(unsigned *p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0;
}
static int zero_words(void *p, size_t bytes, hipStream_t st)
{
    const size_t n = bytes / 4;
    if (!n) return 0;
    const unsigned g = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_zero_words, dim3(g), dim3(256), 0, st, (unsigned *)p, n);
    MIDAS_HIP(hipGetLastError());
    return 0;
}
int zero_device(void *p, size_t bytes, hipStream_t st) { return zero_words(p, bytes, st); }
#define TRY_LU(x)               \
    do {                        \
        int r_ = (x);           \
        if (r_) return r_;      \
    } while (0)

template <typename T> static cplx<T> cone() { return {(T)1, (T)0}; }
template <typename T> static cplx<T> cmone() { return {(T)-1, (T)0}; }
template <typename T> static cplx<T> czero() { return {(T)0, (T)0}; }

// Columns [k, k+nb) of an n x n matrix factored right-looking in inner
// blocks of LU_IB columns: each block's pivots and rank-IB update reach only
// these columns.
template <typename T>
static int panel_leaf(typename cvec<T>::t *A, long lda, int n, int k, int nb, int *ipiv, LuWork<T> &w, hipStream_t st)
{
    // one launch per inner block while the workgroups of a matrix fit on the
    // device together, else one launch per column (MIDAS_LU_COLUMN_STEPS=1
    // forces the latter, so that both stay tested)
    const char *cs_env = getenv("MIDAS_LU_COLUMN_STEPS");
    const bool column_steps = cs_env && atoi(cs_env) != 0;
    const bool fused = !column_steps && (n - k + BLK_T - 1) / BLK_T <= BLK_MAXWG &&
                       (sizeof(T) == 8 || n < 0xfffff);   // FP32 records hold a 20-bit row
    for (int jb = k; jb < k + nb; jb += LU_IB) {
        int jend = min(jb + LU_IB, k + nb);
        int rc = k + nb - jend;
        if (!column_steps && n - jb <= LOC_T) {
            // (its interchanges and U12 solve on the other panel columns included)
            if (n - jb <= 256)
                hipLaunchKernelGGL((k_getf2_local<T, 256>), dim3(1, w.batch), dim3(256), 0, st, A, lda, w.sA, n, jb,
                                   jend - jb, ipiv, (long)w.n, w.info, k, k + nb);
            else
                hipLaunchKernelGGL((k_getf2_local<T, LOC_T>), dim3(1, w.batch), dim3(LOC_T), 0, st, A, lda, w.sA, n, jb,
                                   jend - jb, ipiv, (long)w.n, w.info, k, k + nb);
            MIDAS_HIP(hipGetLastError());
        } else if (fused) {
            const int nblk = (n - jb + BLK_T - 1) / BLK_T;
            MIDAS_CHECK(nblk <= w.maxblk, -2, "lu: step workspace too small");
            // FP32: at most 64 VGPRs (8 waves per SIMD), see k_getf2_block
            constexpr int wpe = sizeof(T) == 4 ? 8 : 4;
            hipLaunchKernelGGL((k_getf2_block<T, wpe>), dim3(nblk, w.batch), dim3(BLK_T), 0, st, A, lda, n, jb,
                               jend - jb, ipiv, w, w.tag_base);
            w.tag_base += (unsigned)(jend - jb);
            const int nc = (jb - k) + rc;
            if (nc > 0)
                hipLaunchKernelGGL(k_swap_trsm<T>, dim3((nc + 63) / 64, w.batch), dim3(64), 0, st, A, lda, w.sA, ipiv,
                                   n, jb, jend, k, k + nb);
            MIDAS_HIP(hipGetLastError());
        } else {
            w.jb = jb;
            for (int j = jb; j <= jend; ++j) {
                int rows = n - j;
                int nblk = rows > 0 ? (rows + STEP_ROWS - 1) / STEP_ROWS : 1;
                MIDAS_CHECK(nblk <= w.maxblk, -2, "lu: step workspace too small");
                hipLaunchKernelGGL(k_getf2_step<T>, dim3(nblk, w.batch), dim3(STEP_THREADS), 0, st, A, lda, n, j,
                                   j < jend ? jend : j, k, k + nb, ipiv, w);
            }
            MIDAS_HIP(hipGetLastError());
            if (rc > 0) {
                hipLaunchKernelGGL(k_trsm_lunit_small<T>, dim3((rc + 63) / 64, w.batch), dim3(64), 0, st, A, lda, w.sA,
                                   jb, jend - jb, jend, k + nb);
                MIDAS_HIP(hipGetLastError());
            }
        }
        // rest of the outer panel: columns [jend, k+nb)
        if (rc > 0 && n - jend > 0) {
            constexpr int wpe = sizeof(T) == 4 ? 5 : 4;
            const int rb = (n - jend + 255) / 256;
            if (rb * w.batch >= 64)
                hipLaunchKernelGGL((k_panel_update<T, wpe, 64>), dim3(rb, (rc + 63) / 64, w.batch), dim3(256), 0, st, A,
                                   lda, w.sA, n, jb, jend - jb, jend, k + nb);
            else
                hipLaunchKernelGGL((k_panel_update<T, wpe, 16>), dim3(rb, (rc + 15) / 16, w.batch), dim3(256), 0, st, A,
                                   lda, w.sA, n, jb, jend - jb, jend, k + nb);
            MIDAS_HIP(hipGetLastError());
        }
    }
    return 0;
}

// Columns [k, k+nb) factored recursively: the left half, its interchanges and
// L11^-1 carried over the right half, the right half's rows below updated by
// one GEMM (complex FP32 / FP64 on the matrix cores), the right half, and its
// interchanges carried back over the left half; halves of at most
// tune(TUNE_PANEL_LEAF) columns go to panel_leaf.  Against one right-looking pass over the
// whole panel, the columns right of an inner block are read and written once
// per recursion level instead of once per inner block.  Scratch: w.u12p
// (L11^-1 of the left half, then L11^-1 A12).
template <typename T>
static int panel_rec(typename cvec<T>::t *A, long lda, int n, int k, int nb, int *ipiv, LuWork<T> &w, hipStream_t st)
{
    typedef typename cvec<T>::t V;
    if (nb <= tune(TUNE_PANEL_LEAF)) return panel_leaf<T>(A, lda, n, k, nb, ipiv, w, st);
    const int h = ((nb / 2 + LU_IB - 1) / LU_IB) * LU_IB, h2 = nb - h;
    int r = panel_rec<T>(A, lda, n, k, h, ipiv, w, st);
    if (r) return r;
    hipLaunchKernelGGL(k_laswp<T>, dim3((h2 + 255) / 256, w.batch), dim3(256), 0, st, A, lda, w.sA, ipiv, (long)w.n, k, k + h,
                       k + h, k + nb, 0, 0);
    MIDAS_HIP(hipGetLastError());
    // A12 <- L11^-1 A12 through the staging buffer (nb x nb per matrix)
    V *Li = (V *)w.u12p, *P = Li + (long)h * h;
    const long sU = (long)w.nb * w.nb;
    // L11^-1 by recursive doubling, its product scratch after Li and P
    if (use_trinv_rec(h) && trinv_rec_scratch(h) <= sU - (long)h * nb)
        r = trinv_rec<T>(A + k + (long)k * lda, lda, h, Li, nullptr, h, Li + (long)h * nb, st, w.batch, w.sA, sU, sU);
    else
        r = trinv<T, true>(A + k + (long)k * lda, lda, h, Li, h, st, w.batch, w.sA, sU);
    if (r) return r;
    r = launch_cgemm<T>(h, h2, h, cone<T>(), Li, h, 0, A + k + (long)(k + h) * lda, lda, 0, czero<T>(), P, h, st,
                        w.batch, sU, w.sA, sU);
    if (r) return r;
    hipLaunchKernelGGL(k_copy2d<T>, dim3((h + 255) / 256, h2, w.batch), dim3(256), 0, st, P, (long)h,
                       A + k + (long)(k + h) * lda, lda, h, h2, sU, w.sA);
    MIDAS_HIP(hipGetLastError());
    // A22 -= A21 A12 for the rows below
    if (n - k - h > 0) {
        r = launch_cgemm<T>(n - k - h, h2, h, cmone<T>(), A + (k + h) + (long)k * lda, lda, 0,
                            A + k + (long)(k + h) * lda, lda, 0, cone<T>(), A + (k + h) + (long)(k + h) * lda, lda,
                            st, w.batch, w.sA, w.sA, w.sA);
        if (r) return r;
    }
    r = panel_rec<T>(A, lda, n, k + h, h2, ipiv, w, st);
    if (r) return r;
    hipLaunchKernelGGL(k_laswp<T>, dim3((h + 255) / 256, w.batch), dim3(256), 0, st, A, lda, w.sA, ipiv, (long)w.n, k + h,
                       k + nb, k, k + h, 0, 0);
    MIDAS_HIP(hipGetLastError());
    return 0;
}

// Panel [k, k+nb) of an n x n matrix, then the inverse of its unit-lower
// diagonal block L11: U12 = L11^-1 A12 becomes one GEMM on the main stream,
// and getrs reuses it.
template <typename T>
static int panel(typename cvec<T>::t *A, long lda, int n, int k, int nb, int *ipiv, LuWork<T> &w, hipStream_t st,
                 const TrailingGemm<T> *ts = nullptr)
{
    typedef typename cvec<T>::t V;
    if (!w.diag_pivot) {
        int r = panel_rec<T>(A, lda, n, k, nb, ipiv, w, st);
        if (r) return r;
        V *Li = (V *)w.linv + (long)(k / w.nb) * w.nb * w.nb;
        const long sU = (long)w.nb * w.nb;
        if (use_trinv_rec(nb) && trinv_rec_scratch(nb) <= sU)
            return trinv_rec<T>(A + k + (long)k * lda, lda, nb, Li, nullptr, w.nb, (V *)w.u12p, st, w.batch, w.sA,
                                w.inv_stride(), sU);
        return trinv<T, true>(A + k + (long)k * lda, lda, nb, Li, w.nb, st, w.batch, w.sA, w.inv_stride());
    }
    // Pivot search restricted to the diagonal block: rows [k, k + nb) factored
    // with partial pivoting among themselves (P11 A11 = L11 U11: the recursive
    // panel of a matrix of order k + nb), the inverses of L11 (for U12) and of
    // U11 formed, and L21 = A21 U11^-1 as one GEMM.  The multiplier of a row
    // below the block is its candidate value over the pivot, so where every
    // |l| <= 1/sqrt(2) (|re| + |im| is within sqrt(2) of the modulus) no row
    // below the block would have beaten a pivot the block chose: partial
    // pivoting over the whole column takes the same pivots.  The largest |l|
    // goes to w.lmax for the caller to report.
    int r = panel_rec<T>(A, lda, k + nb, k, nb, ipiv, w, st);
    if (r) return r;
    V *D = A + k + (long)k * lda;
    const long sI = w.inv_stride(), off = (long)(k / w.nb) * w.nb * w.nb;
    V *Ui = (V *)w.uinv + off;
    // both inverses by recursive doubling (product scratch: the panel's staging buffer)
    const long sU = (long)w.nb * w.nb;
    if (use_trinv_rec(nb) && 2 * trinv_rec_scratch(nb) <= sU)
        r = trinv_rec<T>(D, lda, nb, (V *)w.linv + off, Ui, w.nb, (V *)w.u12p, st, w.batch, w.sA, sI, sU);
    else
        r = trinv_lu<T>(D, lda, nb, (V *)w.linv + off, Ui, w.nb, st, w.batch, w.sA, sI);
    if (r) return r;
    const int mr = n - k - nb;
    if (mr <= 0) return 0;
    V *A21 = A + (k + nb) + (long)k * lda, *S = (V *)w.l21;
    const long sS = (long)mr * nb;
    if (ts && ts->scratch && sizeof(T) == 4 && w.batch == 1 && (ts->mode == TRAIL_F16_1 || ts->mode == TRAIL_BF16_1) &&
        tune(TUNE_L21_SPLIT)) {
        // on the split-precision matrix cores, in place: A21 -= A21 (I - U11^-1)
        TrailingGemm<T> t = *ts;
        t.eye_minus_b = true;
        if ((r = t(mr, nb, nb, A21, lda, Ui, w.nb, A21, lda, st))) return r;
        hipLaunchKernelGGL(k_copy_amax<T>, dim3(min((mr + 255) / 256, 64), min(nb, 32), 1), dim3(256), 0, st, A21, lda,
                           0L, (V *)nullptr, lda, 0L, mr, nb, w.lmax);
        MIDAS_HIP(hipGetLastError());
        return 0;
    }
    if ((r = launch_cgemm<T>(mr, nb, nb, cone<T>(), A21, lda, 0, Ui, w.nb, 0, czero<T>(), S, mr, st, w.batch, w.sA, sI,
                             sS)))
        return r;
    hipLaunchKernelGGL(k_copy_amax<T>, dim3(min((mr + 255) / 256, 64), min(nb, 32), w.batch), dim3(256), 0, st, S,
                       (long)mr, sS, A21, lda, w.sA, mr, nb, w.lmax);
    MIDAS_HIP(hipGetLastError());
    return 0;
}

// U12(:, c0:c1) = L11^-1 A12(:, c0:c1) through a staging buffer (w.u12, or
// stage: nb x (c1 - c0) per matrix at stride sst), for the panel at k.  With
// a single-term split trailing GEMM (ts) and u12_split set, in place on its
// matrix cores instead: A12 -= (I - L11^-1) A12, the operand I - L11^-1
// formed while it is split, A12 split before the update overwrites it.
template <typename T>
static int u12_solve(typename cvec<T>::t *A, long lda, int k, int nb, int c0, int c1, LuWork<T> &w, hipStream_t st,
                     cplx<T> *stage = nullptr, long sst = 0, const TrailingGemm<T> *ts = nullptr)
{
    typedef typename cvec<T>::t V;
    const int nc = c1 - c0;
    if (nc <= 0) return 0;
    const V *Li = (const V *)w.linv + (long)(k / w.nb) * w.nb * w.nb;
    if (ts && sizeof(T) == 4 && w.batch == 1 && (ts->mode == TRAIL_F16_1 || ts->mode == TRAIL_BF16_1) &&
        tune(TUNE_U12_SPLIT)) {
        TrailingGemm<T> t = *ts;
        t.eye_minus_a = true;
        V *A12 = A + k + (long)c0 * lda;
        return t(nb, nc, nb, Li, w.nb, A12, lda, A12, lda, st);
    }
    V *tmp = (V *)(stage ? stage : w.u12);
    const long st12 = stage ? sst : (long)w.nb * w.n;
    int r = launch_cgemm<T>(nb, nc, nb, cone<T>(), Li, w.nb, 0, A + k + (long)c0 * lda, lda, 0, czero<T>(), tmp, nb,
                            st, w.batch, w.inv_stride(), w.sA, st12);
    if (r) return r;
    hipLaunchKernelGGL(k_copy2d<T>, dim3((nb + 255) / 256, nc, w.batch), dim3(256), 0, st, tmp, (long)nb,
                       A + k + (long)c0 * lda, lda, nb, nc, st12, w.sA);
    MIDAS_HIP(hipGetLastError());
    return 0;
}

// The panel of an outer step at column k, W columns, on stream ps.  One
// inner panel when W <= nb; else two: factor the first nb columns, carry their
// interchanges and U12 over the second, update the second's rows below
// (tgp), factor it, and carry its interchanges back over the first.
template <typename T>
static int outer_panel(typename cvec<T>::t *A, long lda, int n, int k, int W, int *ipiv, LuWork<T> &w, hipStream_t ps,
                       const TrailingGemm<T> &tgp)
{
    const int h1 = min(w.nb, W), h2 = W - h1;
    int r = panel<T>(A, lda, n, k, h1, ipiv, w, ps, &tgp);
    if (r || h2 <= 0) return r;
    hipLaunchKernelGGL(k_laswp<T>, dim3((h2 + 255) / 256, w.batch), dim3(256), 0, ps, A, lda, w.sA, ipiv, (long)w.n, k,
                       k + h1, k + h1, k + W, 0, 0);
    MIDAS_HIP(hipGetLastError());
    r = u12_solve<T>(A, lda, k, h1, k + h1, k + W, w, ps, w.u12p, (long)w.nb * w.nb, &tgp);
    if (r) return r;
    r = tgp(n - k - h1, h2, h1, A + (k + h1) + (long)k * lda, lda, A + k + (long)(k + h1) * lda, lda,
            A + (k + h1) + (long)(k + h1) * lda, lda, ps);
    if (r) return r;
    r = panel<T>(A, lda, n, k + h1, h2, ipiv, w, ps, &tgp);
    if (r) return r;
    hipLaunchKernelGGL(k_laswp<T>, dim3((h1 + 255) / 256, w.batch), dim3(256), 0, ps, A, lda, w.sA, ipiv, (long)w.n, k + h1,
                       k + W, k, k + h1, 0, 0);
    MIDAS_HIP(hipGetLastError());
    return 0;
}

// U12 of an outer step (rows [k, k + W)) over columns [c0, c1) and the trailing
// update of those columns: with two inner panels, the second's rows get the
// first's update before their own triangular solve.
template <typename T>
static int outer_update(typename cvec<T>::t *A, long lda, int n, int k, int W, int c0, int c1, LuWork<T> &w,
                        hipStream_t st, const TrailingGemm<T> &tg, bool reuse_a = false)
{
    if (c1 <= c0) return 0;
    const int h1 = min(w.nb, W), h2 = W - h1;
    // the smaller products on a scratch of their own (when there is one), so
    // that the main update's A planes, split by the step's first call, serve
    // its second (reuse_a) as well
    TrailingGemm<T> ts = tg;
    if (tg.scratch2) ts.scratch = tg.scratch2;
    int r = u12_solve<T>(A, lda, k, h1, c0, c1, w, st, nullptr, 0, &ts);
    if (r) return r;
    if (h2 > 0) {
        r = ts(h2, c1 - c0, h1, A + (k + h1) + (long)k * lda, lda, A + k + (long)c0 * lda, lda,
               A + (k + h1) + (long)c0 * lda, lda, st);
        if (r) return r;
        r = u12_solve<T>(A, lda, k + h1, h2, c0, c1, w, st, nullptr, 0, &ts);
        if (r) return r;
    }
    const int nr = n - k - W;
    if (nr <= 0) return 0;
    TrailingGemm<T> tm = tg;
    tm.a_ready = reuse_a && tg.scratch2 != nullptr;
    return tm(nr, c1 - c0, W, A + (k + W) + (long)k * lda, lda, A + k + (long)c0 * lda, lda,
              A + (k + W) + (long)c0 * lda, lda, st);
}

// Look-ahead: the panel of step k+1 is factored on a second stream while the
// main stream applies the rest of step k's trailing update.  Per step, on the
// main stream: wait for panel k, row swaps outside it, U12 = L11^-1 A12,
// update of the next panel's columns (then signal the panel stream), update
// of the remaining columns.  The panel stream only touches the next panel's
// columns, which the remaining update never reads or writes.
template <typename T>
int getrf(typename cvec<T>::t *A, long lda, int n, int *ipiv, LuWork<T> &w, hipStream_t st, TrailingGemm<T> tg,
          TrailingGemm<T> tgp)
{
    typedef typename cvec<T>::t V;
    // lu_lookahead = 0 keeps the panels on the main stream (one stream, no
    // cross-stream waits: for profilers that serialise dispatches)
    hipStream_t ps = tune(TUNE_LU_LOOKAHEAD) ? w.pstream : st;
    MIDAS_CHECK(w.batch == 1 || tg.batch == w.batch, -1, "getrf: trailing GEMM not batched like the factorisation");
    MIDAS_CHECK(w.outer == 1 || (w.outer == 2 && tgp.scratch != tg.scratch) || sizeof(T) == 8 || tg.mode == TRAIL_NATIVE,
                -1, "getrf: two-panel steps need a panel-stream GEMM scratch of its own");
    MIDAS_CHECK(!w.diag_pivot || (w.l21 && w.lmax), -1, "getrf: the diagonal-block pivot search needs its staging buffers");
    const int nbo = w.nb * (w.outer == 2 ? 2 : 1);
    // last-workgroup counters, info, the fault flag and the step tags; the
    // tagged records and words (a fill kernel: safe under stream capture)
    TRY_LU(zero_words(w.counter, sizeof(int) * (2 * w.batch + 1 + w.maxblk * w.batch), st));
    TRY_LU(zero_words(w.pmax, 8 * 2 * w.maxblk * w.batch, st));
    TRY_LU(zero_words(w.cand, 16 * 2 * w.maxblk * LU_IB * w.batch, st));
    TRY_LU(zero_words(w.rowj, 16 * 2 * LU_IB * w.batch, st));
    if (w.diag_pivot) TRY_LU(zero_words(w.lmax, sizeof(unsigned), st));
    w.tag_base = 0;
    MIDAS_HIP(hipEventRecord(w.ev_next, st));
    MIDAS_HIP(hipStreamWaitEvent(ps, w.ev_next, 0));
    int r = outer_panel<T>(A, lda, n, 0, min(nbo, n), ipiv, w, ps, tgp);
    if (r) return r;
    MIDAS_HIP(hipEventRecord(w.ev_panel, ps));
    for (int k = 0; k < n; k += nbo) {
        const int W = min(nbo, n - k);
        MIDAS_HIP(hipStreamWaitEvent(st, w.ev_panel, 0));
        // swaps outside the outer panel: columns [0, k) and [k+W, n)
        int ncols = k + (n - k - W);
        if (ncols > 0) {
            hipLaunchKernelGGL(k_laswp<T>, dim3((ncols + 255) / 256, w.batch), dim3(256), 0, st, A, lda, w.sA, ipiv, (long)w.n,
                               k, k + W, 0, k, k + W, n);
            MIDAS_HIP(hipGetLastError());
        }
        int nr = n - k - W;
        if (nr <= 0) continue;
        // The next panel's columns first (U12, then their trailing update), so
        // that the panel stream can start on them; then the rest.
        const int W1 = min(nbo, nr);
        r = outer_update<T>(A, lda, n, k, W, k + W, k + W + W1, w, st, tg);
        if (r) return r;
        MIDAS_HIP(hipEventRecord(w.ev_next, st));
        // The rest of the update is enqueued before the panel: a panel is
        // hundreds of small launches, and the host takes milliseconds to issue
        // them; issued first, the big update runs while they are issued.
        r = outer_update<T>(A, lda, n, k, W, k + W + W1, n, w, st, tg, true);
        if (r) return r;
        MIDAS_HIP(hipStreamWaitEvent(ps, w.ev_next, 0));
        r = outer_panel<T>(A, lda, n, k + W, W1, ipiv, w, ps, tgp);
        if (r) return r;
        MIDAS_HIP(hipEventRecord(w.ev_panel, ps));
    }
    MIDAS_HIP(hipStreamWaitEvent(st, w.ev_panel, 0));
    // inverses of the diagonal blocks of U, used by getrs (those of L come from
    // the panels, and with the diagonal-block pivot search those of U too):
    // the full blocks in one launch (they are independent), then the last,
    // partial one
    const int nfull = w.diag_pivot ? 0 : n / w.nb;
    if (nfull > 0) {
        r = trinv<T, false>(A, lda, w.nb, (V *)w.uinv, w.nb, st, w.batch, w.sA, w.inv_stride(), nfull,
                            (long)w.nb * (lda + 1), (long)w.nb * w.nb);
        if (r) return r;
    }
    if (n % w.nb && !w.diag_pivot) {
        const int k = nfull * w.nb;
        r = trinv<T, false>(A + k + (long)k * lda, lda, n - k, (V *)w.uinv + (long)nfull * w.nb * w.nb, w.nb, st,
                            w.batch, w.sA, w.inv_stride());
        if (r) return r;
    }
    return w.defer_fault_check ? 0 : lu_fault_check<T>(w, st);
}

template <typename T> int lu_fault_check(const LuWork<T> &w, hipStream_t st)
{
    int fault = 0;
    MIDAS_HIP(hipMemcpyAsync(&fault, w.fault, sizeof(int), hipMemcpyDeviceToHost, st));
    MIDAS_HIP(hipStreamSynchronize(st));
    MIDAS_CHECK(fault == 0, -3, "lu: a panel wait timed out (workgroups not co-resident)");
    return 0;
}
template int lu_fault_check<double>(const LuWork<double> &, hipStream_t);
template int lu_fault_check<float>(const LuWork<float> &, hipStream_t);

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

// Few right-hand sides: C(0:m, :) (+)= A(0:m, 0:kd) X, a row strip streamed
// once.
//   sub == 0:  C = A X      (a diagonal-block inverse times B_k)
//   sub == 1:  C -= A X     (the update of the rows below/above)
// X (kd x nrhs, nrhs <= R) is staged in LDS; workgroup 0 also writes it to Y
// when Y is given (the solved block B_k).  64 rows per workgroup, one row per
// lane; the NW waves take an equal share of the kd columns each, and their
// partial sums are added in wave order (deterministic).  NW = 4 for the tall
// updates; NW = 8 or 16 for the square diagonal blocks, where only kd / 64
// workgroups exist and each wave must cover few columns.  The bound is HBM:
// each element of A is read once, 8 (c64) or 16 (c128) bytes.
constexpr int STRIP_ROWS = 64;
template <typename T, int R, int NW>
__global__ __launch_bounds__(64 * NW) void k_strip_mv(int m, int kd, const typename cvec<T>::t *A, long lda, long sA,
                                                      const typename cvec<T>::t *X, long ldx, long sX,
                                                      typename cvec<T>::t *C, long ldc, long sC, int nrhs, int sub,
                                                      typename cvec<T>::t *Y, long ldy, long sY)
{
    typedef typename cvec<T>::t V;
    extern __shared__ double strip_smem[];
    cplx<T> *xs = (cplx<T> *)strip_smem;      // kd x R, row-major
    cplx<T> *part = xs + (long)kd * R;        // waves 1..NW-1: 64 x R each
    const int zb = blockIdx.z;
    A += zb * sA;
    X += zb * sX;
    C += zb * sC;
    for (int e = threadIdx.x; e < kd * R; e += 64 * NW) {
        const int c = e / R, q = e % R;
        V v = V{0, 0};
        if (q < nrhs) v = X[c + (long)q * ldx];
        xs[e] = cplx<T>{v.x, v.y};
    }
    __syncthreads();
    if (Y != nullptr && blockIdx.x == 0) {
        Y += zb * sY;
        for (int e = threadIdx.x; e < kd * R; e += 64 * NW) {
            const int c = e / R, q = e % R;
            if (q < nrhs) Y[c + (long)q * ldy] = V{xs[e].r, xs[e].i};
        }
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int r = blockIdx.x * STRIP_ROWS + lane;
    const int per = (kd + NW - 1) / NW;
    const int c0 = min(kd, wv * per), c1 = min(kd, c0 + per);
    cplx<T> acc[R];
#pragma unroll
    for (int q = 0; q < R; ++q) acc[q] = cplx<T>{(T)0, (T)0};
    if (r < m) {
        constexpr int U = sizeof(T) == 4 ? 16 : 8;   // loads in flight per lane
        const V *col = A + (long)c0 * lda + r;
        for (int c = c0; c < c1; c += U) {
            V a[U];
#pragma unroll
            for (int u = 0; u < U; ++u) a[u] = c + u < c1 ? col[(long)(c - c0 + u) * lda] : V{0, 0};
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (c + u >= c1) continue;
                const cplx<T> av = {a[u].x, a[u].y};
#pragma unroll
                for (int q = 0; q < R; ++q) acc[q] = cadd(acc[q], cmul(av, xs[(c + u) * R + q]));
            }
        }
    }
    if (wv > 0) {
#pragma unroll
        for (int q = 0; q < R; ++q) part[((wv - 1) * STRIP_ROWS + lane) * R + q] = acc[q];
    }
    __syncthreads();
    if (wv == 0 && r < m) {
        for (int w2 = 0; w2 < NW - 1; ++w2)
#pragma unroll
            for (int q = 0; q < R; ++q) acc[q] = cadd(acc[q], part[(w2 * STRIP_ROWS + lane) * R + q]);
#pragma unroll
        for (int q = 0; q < R; ++q)
            if (q < nrhs) {
                V *o = C + r + (long)q * ldc;
                if (sub) {
                    const V v = *o;
                    *o = V{v.x - acc[q].r, v.y - acc[q].i};
                } else {
                    *o = V{acc[q].r, acc[q].i};
                }
            }
    }
}

template <typename T, int R, int NW>
static int strip_mv(int m, int kd, const typename cvec<T>::t *A, long lda, long sA, const typename cvec<T>::t *X,
                    long ldx, long sX, typename cvec<T>::t *C, long ldc, long sC, int nrhs, int sub,
                    typename cvec<T>::t *Y, long ldy, long sY, int batch, hipStream_t st)
{
    const size_t lds = sizeof(cplx<T>) * ((size_t)kd * R + (size_t)(NW - 1) * STRIP_ROWS * R);
    const int gx = max(1, (m + STRIP_ROWS - 1) / STRIP_ROWS);
    hipLaunchKernelGGL((k_strip_mv<T, R, NW>), dim3(gx, 1, batch), dim3(64 * NW), lds, st, m, kd, A, lda, sA, X, ldx,
                       sX, C, ldc, sC, nrhs, sub, Y, ldy, sY);
    MIDAS_HIP(hipGetLastError());
    return 0;
}

// Blocked substitution for at most R right-hand sides: per diagonal block,
// Y_k = inv(D_k) B_k into tmp, then one streaming pass over the block column
// that updates the remaining rows and writes Y_k back to B_k.  Every factor
// element is read once per solve.
template <typename T, int R>
static int getrs_few(const typename cvec<T>::t *A, long lda, int n, const LuWork<T> &w, typename cvec<T>::t *B,
                     long ldb, int nrhs, typename cvec<T>::t *tmp, long sT, hipStream_t st, long sB)
{
    typedef typename cvec<T>::t V;
    const int nb = w.nb, bt = w.batch;
    const long sA = w.sA, sI = w.inv_stride();
    constexpr int DW = R <= 2 ? 16 : 8;   // waves per diagonal-block workgroup (no spills at 8 RHS)
    int r;
    for (int k = 0; k < n; k += nb) {
        const int b = min(nb, n - k);
        const V *Li = (const V *)w.linv + (long)(k / nb) * nb * nb;
        if ((r = strip_mv<T, R, DW>(b, b, Li, nb, sI, B + k, ldb, sB, tmp, nb, sT, nrhs, 0, nullptr, 0, 0, bt, st)))
            return r;
        if ((r = strip_mv<T, R, 4>(n - k - b, b, A + (k + b) + (long)k * lda, lda, sA, tmp, nb, sT, B + k + b, ldb,
                                   sB, nrhs, 1, B + k, ldb, sB, bt, st)))
            return r;
    }
    for (int k = ((n - 1) / nb) * nb; k >= 0; k -= nb) {
        const int b = min(nb, n - k);
        const V *Ui = (const V *)w.uinv + (long)(k / nb) * nb * nb;
        if ((r = strip_mv<T, R, DW>(b, b, Ui, nb, sI, B + k, ldb, sB, tmp, nb, sT, nrhs, 0, nullptr, 0, 0, bt, st)))
            return r;
        if ((r = strip_mv<T, R, 4>(k, b, A + (long)k * lda, lda, sA, tmp, nb, sT, B, ldb, sB, nrhs, 1, B + k, ldb, sB,
                                   bt, st)))
            return r;
    }
    return 0;
}

long getrs_tmp_elems(int n, int nb, int nrhs) { return (long)max(n, nb) * nrhs; }

// Solve with the factors of getrf for every matrix of the batch: B (n x nrhs,
// ldb, stride sB per matrix); tmp needs getrs_tmp_elems(n, nb, nrhs) per matrix.
template <typename T>
int getrs(const typename cvec<T>::t *A, long lda, int n, const int *perm, const LuWork<T> &w,
          typename cvec<T>::t *B, long ldb, int nrhs, typename cvec<T>::t *tmp, hipStream_t st, long sB)
{
    typedef typename cvec<T>::t V;
    const int nb = w.nb, bt = w.batch;
    const long sA = w.sA, sI = w.inv_stride(), sT = getrs_tmp_elems(n, nb, nrhs);
    if (n <= 0 || nrhs <= 0) return 0;
    MIDAS_CHECK(nrhs <= 65535, -1, "getrs: nrhs too large for one launch");
    hipLaunchKernelGGL(k_permute_rows<T>, dim3((n + 255) / 256, nrhs, bt), dim3(256), 0, st, B, ldb, sB, tmp, (long)n,
                       sT, perm, n);
    hipLaunchKernelGGL(k_copy2d<T>, dim3((n + 255) / 256, nrhs, bt), dim3(256), 0, st, tmp, (long)n, B, ldb, n, nrhs,
                       sT, sB);
    MIDAS_HIP(hipGetLastError());
    // few right-hand sides: streaming passes instead of thin GEMMs
    if (nb <= 512 && !getenv("MIDAS_GETRS_GEMM")) {
        if (nrhs <= 2) return getrs_few<T, 2>(A, lda, n, w, B, ldb, nrhs, tmp, sT, st, sB);
        if constexpr (sizeof(T) == 4)
            if (nrhs <= 8) return getrs_few<T, 8>(A, lda, n, w, B, ldb, nrhs, tmp, sT, st, sB);
    }
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

// sizes of the ten work buffers (pmax, pidx, counter, cand, rowj, pinv, linv, uinv, u12, u12p)
template <typename T> static void lu_sizes(int n, int nb, int batch, size_t need[10])
{
    const size_t maxblk = (n + BLK_T - 1) / BLK_T + 1;   // workgroups of either panel kernel
    const size_t inv = (size_t)((n + nb - 1) / nb) * nb * nb;
    // 16 bytes per candidate element: complex FP64, or complex FP32 as two tagged words
    need[0] = (size_t)8 * 2 * maxblk * batch;                        // pmax (FP32: 64-bit records)
    need[1] = sizeof(int) * 2 * maxblk * batch;                      // pidx
    need[2] = sizeof(int) * (2 * (size_t)batch + 1 + maxblk * batch);   // counter, info, fault, tags
    need[3] = (size_t)16 * 2 * maxblk * LU_IB * batch;               // cand
    need[4] = (size_t)16 * 2 * LU_IB * batch;                        // rowj
    need[5] = sizeof(cplx<T>) * batch;                               // pinv
    need[6] = sizeof(cplx<T>) * inv * batch;                         // linv
    need[7] = sizeof(cplx<T>) * inv * batch;                         // uinv
    need[8] = sizeof(cplx<T>) * (size_t)nb * n * batch;              // u12
    need[9] = sizeof(cplx<T>) * (size_t)nb * nb * batch;             // u12p
}
static inline size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

template <typename T>
size_t LuWork<T>::bytes_needed(int n, int nb, int batch)
{
    size_t need[10], t = 0;
    lu_sizes<T>(n, nb, batch, need);
    for (size_t b : need) t += round256(b);
    return t;
}

template <typename T>
int LuWork<T>::alloc(int n_, int nb_, int batch_, long sA_)
{
    n = n_;
    nb = nb_;
    batch = batch_;
    sA = sA_;
    maxblk = (n + BLK_T - 1) / BLK_T + 1;
    size_t need[10];
    lu_sizes<T>(n, nb, batch, need);
    void **buf[10] = {(void **)&pmax, (void **)&pidx, (void **)&counter, (void **)&cand, (void **)&rowj,
                      (void **)&pinv, (void **)&linv, (void **)&uinv, (void **)&u12, (void **)&u12p};
    for (int i = 0; i < 10; ++i) {
        if (*buf[i] && need[i] <= cap[i]) continue;
        MIDAS_CHECK(!borrowed, -1, "lu: lent workspace too small (buffer %d: %zu of %zu bytes)", i, cap[i], need[i]);
        if (*buf[i]) MIDAS_HIP(hipFree(*buf[i]));
        *buf[i] = nullptr;
        cap[i] = 0;
        MIDAS_HIP(hipMalloc(buf[i], need[i]));
        cap[i] = need[i];
    }
    info = counter + batch;
    fault = counter + 2 * batch;
    tag = (unsigned *)(counter + 2 * batch + 1);
    if (!pstream) MIDAS_HIP(hipStreamCreate(&pstream));
    if (!ev_panel) MIDAS_HIP(hipEventCreateWithFlags(&ev_panel, hipEventDisableTiming));
    if (!ev_next) MIDAS_HIP(hipEventCreateWithFlags(&ev_next, hipEventDisableTiming));
    return 0;
}

template <typename T>
int LuWork<T>::lend(void *mem, size_t bytes, int n_, int nb_, int batch_, hipStream_t ps, hipEvent_t e1, hipEvent_t e2)
{
    MIDAS_CHECK(mem && ps && e1 && e2, -1, "lu: lend needs memory, a stream and two events");
    size_t need[10];
    lu_sizes<T>(n_, nb_, batch_, need);
    MIDAS_CHECK(bytes >= bytes_needed(n_, nb_, batch_), -1, "lu: lent %zu bytes, need %zu", bytes,
                bytes_needed(n_, nb_, batch_));
    release();
    void **buf[10] = {(void **)&pmax, (void **)&pidx, (void **)&counter, (void **)&cand, (void **)&rowj,
                      (void **)&pinv, (void **)&linv, (void **)&uinv, (void **)&u12, (void **)&u12p};
    char *q = (char *)mem;
    for (int i = 0; i < 10; ++i) {
        *buf[i] = q;
        cap[i] = need[i];
        q += round256(need[i]);
    }
    pstream = ps;
    ev_panel = e1;
    ev_next = e2;
    borrowed = true;
    return alloc(n_, nb_, batch_, 0);
}

template <typename T>
void LuWork<T>::release()
{
    if (!borrowed) {
        if (pstream) {
            hipStreamSynchronize(pstream);
            if (own_stream) hipStreamDestroy(pstream);
        }
        hipFree(pmax);
        hipFree(pidx);
        hipFree(counter);
        hipFree(pinv);
        hipFree(linv);
        hipFree(uinv);
        hipFree(u12);
        hipFree(u12p);
        hipFree(cand);
        hipFree(rowj);
        if (own_stream) {
            if (ev_panel) hipEventDestroy(ev_panel);
            if (ev_next) hipEventDestroy(ev_next);
        }
    }
    borrowed = false;
    own_stream = true;
    pstream = nullptr;
    ev_panel = ev_next = nullptr;
    pmax = nullptr;
    pidx = counter = info = fault = nullptr;
    tag = nullptr;
    pinv = nullptr;
    linv = uinv = u12 = u12p = cand = rowj = nullptr;
    for (size_t &c : cap) c = 0;
}

template <typename T>
void LuWork<T>::share_stream(hipStream_t ps, hipEvent_t e1, hipEvent_t e2)
{
    if (!ps || pstream == ps) return;
    if (own_stream && !borrowed) {
        if (pstream) {
            hipStreamSynchronize(pstream);
            hipStreamDestroy(pstream);
        }
        if (ev_panel) hipEventDestroy(ev_panel);
        if (ev_next) hipEventDestroy(ev_next);
    }
    pstream = ps;
    ev_panel = e1;
    ev_next = e2;
    own_stream = false;
}

template struct LuWork<double>;
template struct LuWork<float>;
template int getrf<double>(double2 *, long, int, int *, LuWork<double> &, hipStream_t, TrailingGemm<double>,
                            TrailingGemm<double>);
template int getrf<float>(float2 *, long, int, int *, LuWork<float> &, hipStream_t, TrailingGemm<float>,
                           TrailingGemm<float>);
template int getrs<double>(const double2 *, long, int, const int *, const LuWork<double> &, double2 *, long, int,
                           double2 *, hipStream_t, long);
template int getrs<float>(const float2 *, long, int, const int *, const LuWork<float> &, float2 *, long, int,
                          float2 *, hipStream_t, long);

}  // namespace midas
