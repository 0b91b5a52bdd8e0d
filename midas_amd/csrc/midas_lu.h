// Dense LU workspace and entry points (see midas_lu.hip).
#pragma once
#include "midas_common.h"

namespace midas {

constexpr int LU_IB = 16;   // inner panel width (column-step kernels)

template <typename T> struct LuWork {
    int n = 0, nb = 128, maxblk = 0;
    int batch = 1;                   // matrices factored together (same n), A stride sA
    long sA = 0;
    int jb = 0;                      // first column of the current inner block
    T *pmax = nullptr;               // per-workgroup pivot candidates (two buffers)
    int *pidx = nullptr;
    int *counter = nullptr;          // last-workgroup counter
    int *info = nullptr;             // first exactly-zero pivot (1-based), 0 if none
    unsigned *tag = nullptr;         // step tag per workgroup of the inner-block kernel, per matrix
    int *fault = nullptr;            // set when an inner-block wait gave up
    unsigned tag_base = 0;           // column steps issued so far (host count)
    cplx<T> *cand = nullptr;         // candidate pivot rows, one per workgroup (two buffers)
    cplx<T> *rowj = nullptr;         // row j before its interchange (two buffers)
    cplx<T> *pinv = nullptr;         // 1 / current pivot
    cplx<T> *linv = nullptr;         // inverses of the diagonal blocks of L (nb x nb each)
    cplx<T> *uinv = nullptr;         // inverses of the diagonal blocks of U
    cplx<T> *u12 = nullptr;          // nb x n staging for U12 = L11^-1 A12
    cplx<T> *u12p = nullptr;         // nb x nb staging of the panel stream (two-panel steps)
    int outer = 1;                   // inner panels per outer step: 1, or 2 (trailing updates of 2 nb columns)
    hipStream_t pstream = nullptr;   // panel stream (look-ahead), default priority
    hipEvent_t ev_panel = nullptr, ev_next = nullptr;
    bool own_stream = true;          // pstream and the events made (and destroyed) here
    size_t cap[10] = {};              // bytes held by each buffer above (grow-only)
    bool borrowed = false;           // buffers, stream and events lent by the caller (lend())
    bool defer_fault_check = false;  // getrf leaves the wait-fault check to lu_fault_check (stream capture)
    // pivot search restricted to each panel's diagonal block: L21 = A21 U11^-1
    // staged in l21 (n x nb per matrix of the batch), the largest |L21| of the
    // batch in *lmax (float bits)
    bool diag_pivot = false;
    cplx<T> *l21 = nullptr;
    unsigned *lmax = nullptr;
    // Sizes the workspace for (n, nb, batch, sA): buffers grow when too small
    // and are kept otherwise, so a LuWork reused for one size allocates once;
    // the panel stream and its events are created on first use.
    int alloc(int n, int nb, int batch = 1, long sA = 0);
    void release();
    // Bytes of device memory alloc(n, nb, batch) needs, and a workspace lent
    // from `mem` (that many bytes, 256-byte aligned) with the given panel
    // stream (may be the main stream: the factorisation then runs without
    // look-ahead overlap) and events: alloc() then only checks the sizes.
    static size_t bytes_needed(int n, int nb, int batch = 1);
    int lend(void *mem, size_t bytes, int n, int nb, int batch, hipStream_t ps, hipEvent_t e1, hipEvent_t e2);
    // Use the caller's panel stream and events (kept for the life of its
    // context) instead of making them: streams made and destroyed call after
    // call are placed on the process's few hardware queues anew each time and
    // can share the main stream's, which serialises the look-ahead.
    void share_stream(hipStream_t ps, hipEvent_t e1, hipEvent_t e2);
    long inv_stride() const { return (long)((n + nb - 1) / nb) * nb * nb; }   // diagonal-block inverses per matrix
};

// Trailing update C(m x n) -= A(m x k) B(k x n), complex, column-major.
enum TrailingMode { TRAIL_NATIVE = 0, TRAIL_F16_1 = 1, TRAIL_F16_3 = 3, TRAIL_BF16_1 = 11, TRAIL_BF16_3 = 13 };

template <typename T> struct TrailingGemm {
    int mode = TRAIL_NATIVE;
    void *scratch = nullptr;        // split operand planes + column scales (trail16_scratch_bytes)
    int batch = 1;                  // batched factorisations: matrices at stride `stride`
    long stride = 0;
    bool library = true;            // FP64: rocBLAS ZGEMM allowed (where fp64_gemm_library asks for it)
    bool eye_minus_a = false;       // split modes: C -= (I - A) B instead (the identity of A's order)
    bool eye_minus_b = false;       // split modes: C -= A (I - B) instead
    bool a_ready = false;           // split modes: A's planes are in scratch already (a call with the same A)
    void *scratch2 = nullptr;       // two-panel steps: scratch of the U12 and second-panel products, so that
                                    // the main update's A planes outlive them
    int operator()(int m, int n, int k, const void *A, long lda, const void *B, long ldb, void *C, long ldc,
                   hipStream_t st) const;
};

size_t trail16_scratch_bytes(int mmax, int nmax, int kmax);
// scratch of a batched single-term split update (matrices of up to mmax x nmax, k <= kmax)
size_t trail16_batch_bytes(int mmax, int nmax, int kmax, int batch);

template <typename T> struct cvec;
template <> struct cvec<double> { typedef double2 t; };
template <> struct cvec<float> { typedef float2 t; };

// With w.outer == 2 the panel stream factors two nb-wide panels per outer step
// (updating the second with the first through tg_panel, which must own a
// scratch of its own) and the main stream's trailing updates take 2 nb columns
// at once.
template <typename T>
int getrf(typename cvec<T>::t *A, long lda, int n, int *ipiv, LuWork<T> &w, hipStream_t st, TrailingGemm<T> tg,
          TrailingGemm<T> tg_panel = TrailingGemm<T>());
// Solve with the factors of getrf (same LuWork: it holds the diagonal-block
// inverses); tmp needs getrs_tmp_elems(n, nb, nrhs) elements per matrix.
// With w.batch > 1 every matrix (stride w.sA) solves its own B (stride sB).
template <typename T>
int getrs(const typename cvec<T>::t *A, long lda, int n, const int *perm, const LuWork<T> &w, typename cvec<T>::t *B,
          long ldb, int nrhs, typename cvec<T>::t *tmp, hipStream_t st, long sB = 0);
long getrs_tmp_elems(int n, int nb, int nrhs);
// The check getrf makes at its end (a panel wait that timed out), for
// factorisations enqueued with w.defer_fault_check (synchronises st).
template <typename T> int lu_fault_check(const LuWork<T> &w, hipStream_t st);
// bytes (a multiple of 4) of device memory set to zero by a kernel (safe under
// stream capture, unlike hipMemsetAsync)
int zero_device(void *p, size_t bytes, hipStream_t st);
void ipiv_to_perm(const int *ipiv, int n, int *perm);
int ipiv_to_perm_dev(const int *ipiv, int *perm, int n, int batch, hipStream_t st);

// C = alpha A B + beta C through rocBLAS ZGEMM (strided batch when batch > 1):
// the FP64 products when the fp64_gemm_library parameter asks for the vendor
// library (a comparison point; the default is k_zgemm_nn).
int blas_zgemm_nn(int m, int n, int k, z64 alpha, const void *A, long lda, long sA, const void *B, long ldb, long sB,
                  z64 beta, void *C, long ldc, long sC, int batch, hipStream_t st);

// C = alpha A B + beta C, complex FP64 on FP64 MFMA (midas_zgemm.hip), strided
// batch when batch > 1; shape picks the tile (0: 128 x 128, 1: 128 x 64 of
// 64 x 32 waves, 2: 64 x 64, 3: 128 x 64 of 32 x 32 waves, 4: 64 x 32).
int launch_zgemm_nn(int m, int n, int k, z64 alpha, const void *A, long lda, long sA, const void *B, long ldb, long sB,
                    z64 beta, void *C, long ldc, long sC, int batch, hipStream_t st, int shape);
// The same kernel in complex FP32 (v_mfma_f32_16x16x4_f32).
int launch_cgemm_nn(int m, int n, int k, c32 alpha, const void *A, long lda, long sA, const void *B, long ldb, long sB,
                    c32 beta, void *C, long ldc, long sC, int batch, hipStream_t st, int shape);

// Eigen-decomposition of a strided batch of Hermitian n x n matrices G
// (midas_heev.hip): eigenvalues ascending into D (stride dstride), the
// eigenvectors over G (column c for D[c]); info[b] = 1 where the QL iteration
// did not converge.  work: heev_work_bytes(n, batch) of device memory.
// path 0 picks by order (1: the one-stage kernel, all pairs; 2: the
// large-order kernel, which computes only the eigenpairs with
// lambda >= cut lambda_max when cut > 0 and leaves the others 0).
size_t heev_work_bytes(int n, int batch);
// Pivoted Cholesky G ~= L L^H of a strided batch of Hermitian positive
// semi-definite n x n matrices (n <= 1024), stopped where the largest remaining
// diagonal is <= tolrel max diag(G); L is n x n per matrix with columns past
// the rank zeroed; rank[b] receives the number of columns (midas_heev.hip).
int launch_pchol_batched(int n, const double2 *G, long ld, long stride, double2 *L, long ldl, long lstride, int *rank,
                         int batch, double tolrel, hipStream_t st);
int launch_heev_batched(int n, double2 *G, long ld, long stride, double *D, long dstride, int *info, int batch,
                        void *work, hipStream_t st, double cut = 0.0, int path = 0);

// C = alpha op(A) op(B) + beta C; op: 0 none, 1 transpose, 2 conjugate transpose.
template <typename T>
int launch_cgemm(int m, int n, int k, cplx<T> alpha, const void *A, long lda, int tA, const void *B, long ldb,
                 int tB, cplx<T> beta, void *C, long ldc, hipStream_t st, int batch = 1, long sA = 0, long sB = 0,
                 long sC = 0);

}  // namespace midas
