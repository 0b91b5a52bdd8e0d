// C-ABI of the MIDAS gfx950 library (declared in include/midas_gpu.h).
#include "midas_common.h"
#include "midas_lu.h"
#include "midas_host.h"
#include "midas_tune.h"
#include "../../include/midas_gpu.h"
#include "../../include/midas_gpu_test.h"

#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

// Device scratch of the allocation-free entry point midas_dev_mom_solve_noalloc:
// every device the library runs on holds one copy until the process ends.  In
// host code `scratch` only names the symbol; its device address comes from
// hipGetSymbolAddress (midas_init, per device) and is what host code passes on.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)76 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace midas {

static thread_local char g_err[1024] = "";
void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}
const char *last_error() { return g_err; }

KernelProfile &profile()
{
    static KernelProfile p;
    return p;
}

}  // namespace midas

using namespace midas;

// ---------------------------------------------------------------------------
// device helpers for the solver
// ---------------------------------------------------------------------------
__global__ void k_z2c(const double2 *Z, long ldz, float2 *W, long ldw, int m, int n)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    int c = blockIdx.y;
    if (r < m && c < n) {
        double2 v = Z[(long)r + (long)c * ldz];
        W[(long)r + (long)c * ldw] = make_float2((float)v.x, (float)v.y);
    }
}
__global__ void k_c2z_add(const float2 *D, long ldd, double2 *X, long ldx, int m, int n, int accumulate)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    int c = blockIdx.y;
    if (r < m && c < n) {
        float2 d = D[(long)r + (long)c * ldd];
        double2 *x = X + (long)r + (long)c * ldx;
        if (accumulate) *x = make_double2(x->x + (double)d.x, x->y + (double)d.y);
        else *x = make_double2((double)d.x, (double)d.y);
    }
}
__global__ void k_copyz(const double2 *S, long lds, double2 *D, long ldd, int m, int n)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    int c = blockIdx.y;
    if (r < m && c < n) D[(long)r + (long)c * ldd] = S[(long)r + (long)c * lds];
}
// per-column max modulus of a (m x n) block -> out[c]
__global__ void k_colmax(const double2 *X, long ldx, int m, double *out)
{
    int c = blockIdx.x;
    double mx = 0.0;
    for (int r = threadIdx.x; r < m; r += blockDim.x) {
        double2 v = X[(long)r + (long)c * ldx];
        mx = fmax(mx, sqrt(v.x * v.x + v.y * v.y));
    }
    __shared__ double red[256];
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = red[0];
}
__global__ void k_max_reduce(const double *v, int n, double *out)
{
    double mx = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) mx = fmax(mx, v[i]);
    __shared__ double red[256];
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

static int trail_mode(int mode)
{
    switch (mode) {
    case MIDAS_SOLVE_C32_IR: return TRAIL_NATIVE;
    case MIDAS_SOLVE_F16_IR: return TRAIL_F16_1;
    case MIDAS_SOLVE_F16X3_IR: return TRAIL_F16_3;
    case MIDAS_SOLVE_BF16_IR: return TRAIL_BF16_1;
    case MIDAS_SOLVE_BF16X3_IR: return TRAIL_BF16_3;
    }
    return -1;
}

// Outer step width of the LU (tuned defaults, midas_tune_set): FP64; the
// mixed modes; and for the single-term split modes on large n a wider step
// (above 512: two panels of half its width per step, so the trailing updates
// carry twice the arithmetic per byte of C).
static int default_nb(int n, int mode)
{
    if (mode == MIDAS_SOLVE_FP64) return tune(TUNE_LU_NB_FP64);
    if ((mode == MIDAS_SOLVE_F16_IR || mode == MIDAS_SOLVE_BF16_IR) && n >= tune(TUNE_LU_TWO_PANEL_MIN_N))
        return tune(TUNE_LU_NB_SPLIT);
    return tune(TUNE_LU_NB_MIXED);
}

__global__ void k_zeroz(double2 *X, long ldx, int m, int n)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    int c = blockIdx.y;
    if (r < m && c < n) X[(long)r + (long)c * ldx] = make_double2(0.0, 0.0);
}

// perm from LAPACK interchanges for one matrix, on the device: one workgroup,
// the permutation held in LDS (n <= PERM_LDS_MAX), one thread walks the swaps.
constexpr int PERM_LDS_MAX = 16384;
__global__ __launch_bounds__(256) void k_ipiv_to_perm_lds(const int *ipiv, int *perm, int n)
{
    __shared__ int p[PERM_LDS_MAX];
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = i;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int i = 0; i < n; ++i) {
            const int q = ipiv[i];
            if (q != i) {
                const int t = p[i];
                p[i] = p[q];
                p[q] = t;
            }
        }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) perm[i] = p[i];
}

// ipiv (device) -> perm (device) for one matrix of order n
static int make_perm(const int *ipiv_d, int *perm_d, int n, hipStream_t st)
{
    if (n <= PERM_LDS_MAX) {
        hipLaunchKernelGGL(k_ipiv_to_perm_lds, dim3(1), dim3(256), 0, st, ipiv_d, perm_d, n);
        MIDAS_HIP(hipGetLastError());
        return 0;
    }
    std::vector<int> ipiv(n), perm(n);
    TRY(copy_sync(ipiv.data(), ipiv_d, sizeof(int) * n, hipMemcpyDeviceToHost, st));
    ipiv_to_perm(ipiv.data(), n, perm.data());
    TRY(copy_sync(perm_d, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
    return 0;
}

// FP64 LU of A (in place) and X = A^-1 B with it (PZGESV semantics).
static int fp64_lu_solve(midas_ctx *ctx, SolveWs &ws, int n, int nrhs, double2 *A, long lda, const double2 *B,
                         long ldb, double2 *X, long ldx, int nb, midas_solve_info_t &inf)
{
    hipStream_t st = ctx->st;
    LuWork<double> &w = ws.lud;
    if (!ws.lent) w.share_stream(ctx->ps, ctx->pev[0], ctx->pev[1]);
    TRY(w.alloc(n, nb));
    TrailingGemm<double> tg;
    tg.library = !ws.lent;
    int r = getrf<double>(A, lda, n, ws.ipiv.as<int>(), w, st, tg);
    if (r) return r;
    MIDAS_HIP(hipMemcpyAsync(&inf.info, w.info, sizeof(int), hipMemcpyDeviceToHost, st));
    TRY(make_perm(ws.ipiv.as<int>(), ws.perm.as<int>(), n, st));
    hipLaunchKernelGGL(k_copyz, dim3((n + 255) / 256, nrhs), dim3(256), 0, st, B, ldb, X, ldx, n, nrhs);
    TRY(ws.t2.reserve(sizeof(double2) * (size_t)getrs_tmp_elems(n, nb, nrhs)));
    r = getrs<double>(A, lda, n, ws.perm.as<int>(), w, X, ldx, nrhs, ws.t2.as<double2>(), st);
    if (r) return r;
    MIDAS_HIP(hipStreamSynchronize(st));
    return 0;
}

// R = B - Z X and the column maxima of X and R (device work only; a stream
// capture may record it).
static int residual_launch(midas_ctx *ctx, SolveWs &ws, int n, int nrhs, const double2 *Z, long ldz,
                           const double2 *B, long ldb, const double2 *X, long ldx)
{
    hipStream_t st = ctx->st;
    double *colx = ws.norms.as<double>() + n, *colr = colx + nrhs;
    double2 *R = ws.R.as<double2>();
    const dim3 g2((n + 255) / 256, nrhs);
    int r;
    if (ws.mf_active && nrhs <= 4) {
        r = launch_residual_free(*ws.mf_cells, n / 3, ws.mf_k0, ws.mf_pi, X, ldx, nrhs, B, ldb, R, n, ws.mfs.p, st);
    } else if (nrhs <= 4) {
        r = launch_residual_few(Z, ldz, n, X, ldx, nrhs, B, ldb, R, n, ws.zx.p, st);
    } else {
        hipLaunchKernelGGL(k_copyz, g2, dim3(256), 0, st, B, ldb, R, (long)n, n, nrhs);
        // many right-hand sides: k_zgemm_nn (through launch_cgemm), or rocBLAS
        // when fp64_gemm_library = 1
        r = nrhs >= 32 && !ws.lent && tune(TUNE_FP64_GEMM_LIBRARY) ? blas_zgemm_nn(n, nrhs, n, z64{-1.0, 0.0}, Z, ldz, 0, X, ldx, 0, z64{1.0, 0.0},
                                                   R, n, 0, 1, st)
                                   : launch_cgemm<double>(n, nrhs, n, z64{-1.0, 0.0}, Z, ldz, 0, X, ldx, 0,
                                                          z64{1.0, 0.0}, R, n, st);
    }
    if (r) return r;
    hipLaunchKernelGGL(k_colmax, dim3(nrhs), dim3(256), 0, st, X, ldx, n, colx);
    hipLaunchKernelGGL(k_colmax, dim3(nrhs), dim3(256), 0, st, R, (long)n, n, colr);
    MIDAS_HIP(hipGetLastError());
    return 0;
}

// The column maxima to the host and the normwise backward error
// max_j ||r_j|| / (||Z|| ||x_j|| + ||b_j||) (infinity norms).
static int berr_read(midas_ctx *ctx, SolveWs &ws, int n, int nrhs, double anorm, const std::vector<double> &hb,
                     double *berr)
{
    hipStream_t st = ctx->st;
    const double *colx = ws.norms.as<double>() + n, *colr = colx + nrhs;
    std::vector<double> hx(nrhs), hr(nrhs);
    TRY(copy_sync(hx.data(), colx, sizeof(double) * nrhs, hipMemcpyDeviceToHost, st));
    TRY(copy_sync(hr.data(), colr, sizeof(double) * nrhs, hipMemcpyDeviceToHost, st));
    double b = 0.0;
    for (int c = 0; c < nrhs; ++c) {
        const double den = anorm * hx[c] + hb[c];
        const double e = den > 0 ? hr[c] / den : (hr[c] > 0 ? INFINITY : 0.0);
        b = (e > b || std::isnan(e)) ? e : b;
    }
    *berr = b;
    return 0;
}

static int residual_berr(midas_ctx *ctx, SolveWs &ws, int n, int nrhs, const double2 *Z, long ldz, const double2 *B,
                         long ldb, const double2 *X, long ldx, double anorm, const std::vector<double> &hb,
                         double *berr)
{
    TRY(residual_launch(ctx, ws, n, nrhs, Z, ldz, B, ldb, X, ldx));
    return berr_read(ctx, ws, n, nrhs, anorm, hb, berr);
}

// One refinement sweep's device work: D = LU^-1 R (FP32), X += D, R = B - Z X,
// column maxima.  The same launches every sweep, so they can be captured once.
struct SweepArgs {
    int n, nrhs;
    const float2 *A;
    long lda;
    const LuWork<float> *w;
    double2 *Z;
    long ldz;
    const double2 *B;
    long ldb;
    double2 *X;
    long ldx;
};
static int sweep_launch(midas_ctx *ctx, SolveWs &ws, const SweepArgs &a)
{
    hipStream_t st = ctx->st;
    const dim3 g2((a.n + 255) / 256, a.nrhs);
    float2 *D = ws.D.as<float2>();
    hipLaunchKernelGGL(k_z2c, g2, dim3(256), 0, st, ws.R.as<double2>(), (long)a.n, D, (long)a.n, a.n, a.nrhs);
    TRY(getrs<float>(a.A, a.lda, a.n, ws.perm.as<int>(), *a.w, D, (long)a.n, a.nrhs, ws.t2.as<float2>(), st));
    hipLaunchKernelGGL(k_c2z_add, g2, dim3(256), 0, st, D, (long)a.n, a.X, a.ldx, a.n, a.nrhs, 1);
    return residual_launch(ctx, ws, a.n, a.nrhs, a.Z, a.ldz, a.B, a.ldb, a.X, a.ldx);
}

// A captured sweep (hipGraph), destroyed with its owner.
struct SweepGraph {
    hipGraph_t g = nullptr;
    hipGraphExec_t e = nullptr;
    ~SweepGraph()
    {
        if (e) hipGraphExecDestroy(e);
        if (g) hipGraphDestroy(g);
    }
};
static int sweep_capture(midas_ctx *ctx, SolveWs &ws, const SweepArgs &a, SweepGraph &sg)
{
    hipStream_t st = ctx->st;
    MIDAS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int r = sweep_launch(ctx, ws, a);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (r) {
        if (g) hipGraphDestroy(g);
        return r;
    }
    MIDAS_HIP(e);
    sg.g = g;
    MIDAS_HIP(hipGraphInstantiate(&sg.e, sg.g, nullptr, nullptr, 0));
    return 0;
}

// Solve Z X = B on the device.  Z (n x n, ldz) FP64; if W32 is non-NULL it
// already holds the FP32 copy of Z (mixed modes).  For MIDAS_SOLVE_FP64 Z is
// factored in place when z_inplace, else copied; mixed modes factor Z in place
// only for the FP64 fallback, which z_inplace allows.  X (n x nrhs, ldx)
// receives the solution; B (n x nrhs, ldb) is read only.  Work buffers live in
// ctx->ws and only grow.
// Clears the matrix-free flag when a solve ends, however it ends.
struct MfOff {
    SolveWs &ws;
    ~MfOff() { ws.mf_active = false; }
};

int solve_dev(midas_ctx *ctx, int n, int nrhs, double2 *Z, long ldz, float2 *W32, long ldw, const double2 *B,
              long ldb, double2 *X, long ldx, int mode, const midas_solve_opts_t *opts, midas_solve_info_t *info,
              bool z_inplace, SolveWs *wsp)
{
    hipStream_t st = ctx->st;
    SolveWs &ws = wsp ? *wsp : ctx->ws;
    midas_solve_info_t inf{};
    const int max_iter = (opts && opts->max_iter > 0) ? opts->max_iter : 30;
    const double tol = (opts && opts->tol > 0) ? opts->tol : std::max(1e-14, std::sqrt((double)n) * DBL_EPSILON);
    const bool fallback = !(opts && opts->no_fallback) && z_inplace;
    int nb = (opts && opts->nb > 0) ? opts->nb : default_nb(n, mode);
    nb = nb > n ? n : nb;
    MIDAS_CHECK(n > 0 && nrhs > 0, -1, "solve: n=%d nrhs=%d", n, nrhs);
    MIDAS_CHECK(mode >= 0 && mode <= 5, -1, "solve: unknown mode %d", mode);
    TRY(ws.ipiv.reserve(sizeof(int) * n));
    TRY(ws.perm.reserve(sizeof(int) * n));
    TRY(ws.norms.reserve(sizeof(double) * (n + nrhs * 2 + 4)));
    double *rowsum = ws.norms.as<double>();
    double *scal = rowsum + n + 2 * nrhs;

    // ||Z||_inf before factoring (PZLANGE 'I'); the scratch also serves the
    // few-RHS residuals of refinement.  Matrix-free (MoM mixed solves): both
    // from the cells, Z untouched until a fallback fills it.
    ws.mf_active = ws.mf_cells && mode != MIDAS_SOLVE_FP64 && nrhs <= 4 && !ws.lent && 3 * ws.mf_cells->n == n;
    MfOff mf_off{ws};
    if (ws.mf_active) {
        TRY(ws.mfs.reserve(mf_scratch_bytes(n / 3, nrhs)));
        TRY(launch_rowabs_free(*ws.mf_cells, n / 3, ws.mf_k0, ws.mf_pi, rowsum, ws.mfs.p, st));
    } else {
        MIDAS_CHECK(Z, -1, "solve: no FP64 matrix");
        TRY(ws.zx.reserve(zx_scratch_bytes(n, nrhs <= 4 ? nrhs : 1)));
        TRY(launch_rowabs(Z, ldz, n, rowsum, ws.zx.p, st));
    }
    hipLaunchKernelGGL(k_max_reduce, dim3(1), dim3(256), 0, st, rowsum, n, scal);
    MIDAS_HIP(hipGetLastError());
    TRY(copy_sync(&inf.anorm, scal, sizeof(double), hipMemcpyDeviceToHost, st));

    Timer tm(st, !ws.lent);
    if (mode == MIDAS_SOLVE_FP64) {
        double2 *A = Z;
        long lda = ldz;
        if (!z_inplace) {
            TRY(ws.zcopy.reserve(sizeof(double2) * (size_t)n * n));
            A = ws.zcopy.as<double2>();
            lda = n;
            hipLaunchKernelGGL(k_copyz, dim3((n + 255) / 256, n), dim3(256), 0, st, Z, ldz, A, lda, n, n);
        }
        int r = fp64_lu_solve(ctx, ws, n, nrhs, A, lda, B, ldb, X, ldx, nb, inf);
        if (r) return r;
        inf.t_factor_ms = tm.ms();
        inf.status = MIDAS_STATUS_CONVERGED;
        if (info) *info = inf;
        return inf.info;
    }

    float2 *A = W32;
    long lda = ldw;
    if (!A) {
        TRY(ws.W.reserve(sizeof(float2) * (size_t)n * n));
        A = ws.W.as<float2>();
        lda = n;
        hipLaunchKernelGGL(k_z2c, dim3((n + 255) / 256, n), dim3(256), 0, st, Z, ldz, A, lda, n, n);
        MIDAS_HIP(hipGetLastError());
    }
    // an outer step wider than 512 is two inner panels of half its width
    // (single-term split modes only: their GEMM takes k up to 2048)
    const bool two = nb > 512 && (mode == MIDAS_SOLVE_F16_IR || mode == MIDAS_SOLVE_BF16_IR);
    MIDAS_CHECK(nb <= 512 || (two && nb <= 1024), -1,
                "solve: panel width %d: above 512 needs a single-term split mode, at most 1024", nb);
    const int nbi = two ? (nb + 1) / 2 : nb;
    LuWork<float> &w = ws.luf;
    if (!ws.lent) w.share_stream(ctx->ps, ctx->pev[0], ctx->pev[1]);
    TRY(w.alloc(n, nbi));
    w.outer = two ? 2 : 1;
    TRY(ws.trail.reserve(trail16_scratch_bytes(n, n, nb)));
    TrailingGemm<float> tg, tgp;
    tg.mode = tgp.mode = trail_mode(mode);
    tg.scratch = ws.trail.p;
    if (two) {
        TRY(ws.trailp.reserve(trail16_scratch_bytes(n, nbi, nbi)));
        tgp.scratch = ws.trailp.p;
        // (lent workspaces have no room for it: their steps split A twice)
        if (!ws.lent && tg.mode != TRAIL_NATIVE) {
            TRY(ws.trail2.reserve(trail16_scratch_bytes(nbi, n, nbi)));
            tg.scratch2 = ws.trail2.p;
        }
    }
    // pivot search in the diagonal blocks (MoM solves; not with lent buffers)
    w.diag_pivot = ws.pivot_search && !ws.lent;
    if (w.diag_pivot) {
        const size_t sl = sizeof(float2) * (size_t)n * nbi;
        TRY(ws.l21.reserve(sl + 256));
        w.l21 = ws.l21.as<c32>();
        w.lmax = (unsigned *)((char *)ws.l21.p + sl);
    }
    int r = getrf<float>(A, lda, n, ws.ipiv.as<int>(), w, st, tg, tgp);
    nb = nbi;   // the diagonal blocks getrs works with
    inf.pivot_search = w.diag_pivot ? 1 : 0;
    w.diag_pivot = false;
    if (r) return r;
    MIDAS_HIP(hipMemcpyAsync(&inf.info, w.info, sizeof(int), hipMemcpyDeviceToHost, st));
    if (inf.pivot_search) {
        unsigned lm = 0;
        MIDAS_HIP(hipMemcpyAsync(&lm, w.lmax, sizeof(unsigned), hipMemcpyDeviceToHost, st));
        MIDAS_HIP(hipStreamSynchronize(st));
        inf.max_l21 = (double)__builtin_bit_cast(float, lm);
    }
    TRY(make_perm(ws.ipiv.as<int>(), ws.perm.as<int>(), n, st));
    inf.t_factor_ms = tm.ms();
    // refinement: X = 0, R = B; repeat D = LU^-1 R (FP32), X += D, R = B - Z X (FP64)
    TRY(ws.R.reserve(sizeof(double2) * (size_t)n * nrhs));
    TRY(ws.D.reserve(sizeof(float2) * (size_t)n * nrhs));
    TRY(ws.t2.reserve(sizeof(float2) * (size_t)getrs_tmp_elems(n, nb, nrhs)));
    TRY(ws.xbest.reserve(sizeof(double2) * (size_t)n * nrhs));
    double2 *R = ws.R.as<double2>(), *Xb = ws.xbest.as<double2>();
    const dim3 g2((n + 255) / 256, nrhs);
    hipLaunchKernelGGL(k_copyz, g2, dim3(256), 0, st, B, ldb, R, (long)n, n, nrhs);
    std::vector<double> hb(nrhs);
    double *colr = rowsum + n + nrhs;
    hipLaunchKernelGGL(k_colmax, dim3(nrhs), dim3(256), 0, st, B, ldb, n, colr);
    TRY(copy_sync(hb.data(), colr, sizeof(double) * nrhs, hipMemcpyDeviceToHost, st));
    double berr = INFINITY, best = INFINITY, prev = INFINITY;
    int it = 0;
    bool best_saved = false;   // Xb holds an iterate better than the current X
    // X = 0, then every sweep adds its correction.  The sweeps repeat the same
    // hundreds of small launches (the few-RHS substitution is one per block
    // column): captured once into a hipGraph and replayed, the host issues one
    // launch per sweep (not under no-allocation use, and not around the
    // library GEMM of the many-RHS residual).
    hipLaunchKernelGGL(k_zeroz, g2, dim3(256), 0, st, X, ldx, n, nrhs);
    MIDAS_HIP(hipGetLastError());
    const SweepArgs sa{n, nrhs, A, lda, &w, Z, ldz, B, ldb, X, ldx};
    const bool use_graph = tune(TUNE_REFINE_GRAPH) && !ws.lent && nrhs < 32 && !profile().on;
    SweepGraph sg;
    if (use_graph) TRY(sweep_capture(ctx, ws, sa, sg));
    for (; it < max_iter; ++it) {
        if (use_graph) MIDAS_HIP(hipGraphLaunch(sg.e, st));
        else TRY(sweep_launch(ctx, ws, sa));
        if (it == 0) inf.t_solve_ms = tm.ms();
        TRY(berr_read(ctx, ws, n, nrhs, inf.anorm, hb, &berr));
        if (std::isnan(berr)) {
            best_saved = best < INFINITY;
            ++it;
            break;
        }
        if (berr <= tol) {
            best = berr;
            best_saved = false;
            ++it;
            break;
        }
        if (berr < best) {
            // keep this iterate: a later sweep may be worse
            best = berr;
            hipLaunchKernelGGL(k_copyz, g2, dim3(256), 0, st, X, ldx, Xb, (long)n, n, nrhs);
            best_saved = false;
        } else {
            best_saved = true;
        }
        // no progress: the sweeps stop gaining at least a factor of two
        const bool stalled = it >= 2 && berr > 0.5 * prev;
        prev = berr;
        if (stalled) {
            ++it;
            break;
        }
    }
    inf.iterations = it;
    if (best_saved) {
        // the last sweep made X worse than an earlier iterate: return that one
        hipLaunchKernelGGL(k_copyz, g2, dim3(256), 0, st, Xb, (long)n, X, ldx, n, nrhs);
        MIDAS_HIP(hipGetLastError());
    }
    inf.berr = best;
    inf.t_refine_ms = tm.ms();
    if (best <= tol) {
        inf.status = MIDAS_STATUS_CONVERGED;
    } else if (fallback && inf.info == 0) {
        // refinement cannot reach tol with this factorisation (ZCGESV's rule):
        // the FP64 LU of Z, in place (matrix-free: Z filled first, into the
        // context's buffer when the caller gave none)
        if (ws.mf_active) {
            if (!Z) {
                TRY(ws.zcopy.reserve(sizeof(double2) * (size_t)n * n));
                Z = ws.zcopy.as<double2>();
                ldz = n;
            }
            const CellSet all{n / 3, 0, nullptr};
            TRY(launch_fill(*ws.mf_cells, all, all, ws.mf_k0, ws.mf_quirks, false, Z, ldz, nullptr, 0, st));
        }
        r = fp64_lu_solve(ctx, ws, n, nrhs, Z, ldz, B, ldb, X, ldx, default_nb(n, MIDAS_SOLVE_FP64), inf);
        if (r) return r;
        inf.status = MIDAS_STATUS_FP64_FALLBACK;
        inf.t_refine_ms += tm.ms();
    } else {
        inf.status = MIDAS_STATUS_NOT_CONVERGED;
    }
    if (info) *info = inf;
    if (inf.info) return inf.info;
    MIDAS_CHECK(inf.status != MIDAS_STATUS_NOT_CONVERGED, MIDAS_ERR_NOT_CONVERGED,
                "solve: refinement stopped at backward error %.3g above tol %.3g after %d sweeps (mode %d); X is the "
                "best iterate", inf.berr, tol, it, mode);
    return 0;
}

namespace midas {
int mfma_probe(int which, double *dout, hipStream_t st);
int launch_trail16(int mode, int m, int n, int k, const float2 *A, long lda, const float2 *B, long ldb, float2 *C,
                   long ldc, void *scratch, hipStream_t st, int batch = 1, long stride = 0, bool eye_minus_a = false,
                   bool eye_minus_b = false, bool a_ready = false);
}

// ===========================================================================
// extern "C"
// ===========================================================================
extern "C" {

const char *midas_last_error(void) { return midas::last_error(); }

int midas_init(int device, midas_ctx **ctx)
{
    MIDAS_CHECK(ctx, -1, "midas_init: ctx is NULL");
    int ndev = 0;
    MIDAS_HIP(hipGetDeviceCount(&ndev));
    MIDAS_CHECK(device >= 0 && device < ndev, -1, "midas_init: device %d of %d", device, ndev);
    MIDAS_HIP(hipSetDevice(device));
    midas_ctx *c = new midas_ctx;
    c->device = device;
    hipError_t e = hipStreamCreate(&c->st);
    // the scratch array's address on this device, and the two events of the
    // allocation-free factorisation, made once here
    void *base = nullptr;
    if (e == hipSuccess) e = hipGetSymbolAddress(&base, HIP_SYMBOL(scratch));
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->nev[0], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->nev[1], hipEventDisableTiming);
    // the panel stream and the second CBFM stream, right after the main one
    if (e == hipSuccess) e = hipStreamCreate(&c->ps);
    if (e == hipSuccess) e = hipStreamCreate(&c->s2);
    for (hipEvent_t &ev : c->pev)
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    for (hipEvent_t &ev : c->aev)
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e != hipSuccess) {
        for (hipStream_t s : {c->st, c->ps, c->s2})
            if (s) hipStreamDestroy(s);
        for (hipEvent_t ev : c->nev)
            if (ev) hipEventDestroy(ev);
        for (hipEvent_t ev : c->pev)
            if (ev) hipEventDestroy(ev);
        for (hipEvent_t ev : c->aev)
            if (ev) hipEventDestroy(ev);
        delete c;
        set_error("midas_init: %s", hipGetErrorString(e));
        return -100 - (int)e;
    }
    c->noalloc_base = base;
    *ctx = c;
    return 0;
}

int midas_finalize(midas_ctx *ctx)
{
    if (!ctx) return 0;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->st);
    ctx->ws.release();
    ctx->nws.release();
    ctx->release_calls();
    for (hipGraphExec_t g : ctx->cbfm_graphs)
        if (g) hipGraphExecDestroy(g);
    hipStreamSynchronize(ctx->ps);
    hipStreamSynchronize(ctx->s2);
    for (hipStream_t s : {ctx->st, ctx->ps, ctx->s2})
        hipStreamDestroy(s);
    for (hipEvent_t ev : ctx->nev)
        if (ev) hipEventDestroy(ev);
    for (hipEvent_t ev : ctx->pev)
        if (ev) hipEventDestroy(ev);
    for (hipEvent_t ev : ctx->aev)
        if (ev) hipEventDestroy(ev);
    delete ctx;
    return 0;
}

int midas_release_workspace(midas_ctx *ctx)
{
    MIDAS_CHECK(ctx, -1, "release_workspace: ctx is NULL");
    MIDAS_HIP(hipSetDevice(ctx->device));
    MIDAS_HIP(hipStreamSynchronize(ctx->st));
    ctx->ws.release();
    ctx->release_calls();
    DBuf *b[] = {&ctx->cells.x, &ctx->cells.y, &ctx->cells.z, &ctx->cells.sc, &ctx->cells.eps, &ctx->cells.kappa,
                 &ctx->cells.che, &ctx->cells.znnpp, &ctx->cells.kchi, &ctx->cells.id, &ctx->cells.a};
    for (DBuf *x : b) x->release();
    return 0;
}

int midas_set_cbfm_budgets(midas_ctx *ctx, double lu_bytes, double strip_bytes, double dense_bytes, double aca_bytes)
{
    MIDAS_CHECK(ctx, -1, "set_cbfm_budgets: ctx is NULL");
    const double v[4] = {lu_bytes, strip_bytes, dense_bytes, aca_bytes};
    for (int i = 0; i < 4; ++i) {
        MIDAS_CHECK(v[i] >= 0 && std::isfinite(v[i]), -1, "set_cbfm_budgets: budget %d = %g", i, v[i]);
        ctx->cbfm_budget[i] = v[i];
    }
    return 0;
}

void *midas_stream(midas_ctx *ctx) { return ctx ? (void *)ctx->st : nullptr; }

int midas_profile(midas_ctx *ctx, int enable)
{
    MIDAS_CHECK(ctx, -1, "profile: ctx is NULL");
    KernelProfile &p = profile();
    MIDAS_HIP(hipStreamSynchronize(ctx->st));
    for (auto &r : p.recs) {
        hipEventDestroy(r.a);
        hipEventDestroy(r.b);
    }
    p.recs.clear();
    p.on = enable != 0;
    return 0;
}

int midas_profile_query(midas_ctx *ctx, int kind, int *launches, double *total_ms, double *flops, double *bytes)
{
    MIDAS_CHECK(ctx, -1, "profile_query: ctx is NULL");
    MIDAS_HIP(hipStreamSynchronize(ctx->st));
    int cnt = 0;
    double ms = 0, f = 0, b = 0;
    for (auto &r : profile().recs) {
        if (r.kind != kind) continue;
        float t = 0;
        MIDAS_HIP(hipEventElapsedTime(&t, r.a, r.b));
        ++cnt;
        ms += t;
        f += r.flops;
        b += r.bytes;
    }
    if (launches) *launches = cnt;
    if (total_ms) *total_ms = ms;
    if (flops) *flops = f;
    if (bytes) *bytes = b;
    return 0;
}

int midas_cell_params(midas_ctx *ctx, int n, const double *sc, const double *eps, double k0, int quirks, double *a_n,
                      double *che, double *znnpp, double *kappa)
{
    MIDAS_CHECK(ctx && n > 0 && sc && eps, -1, "cell_params: bad arguments");
    MIDAS_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->st;
    DBuf dsc, deps, da, dche, dzn, dk, dkc;
    TRY(upload(dsc, sc, sizeof(double) * n, st));
    TRY(upload(deps, eps, sizeof(double) * 2 * n, st));
    TRY(da.alloc(sizeof(double) * n));
    TRY(dche.alloc(sizeof(double) * 2 * n));
    TRY(dzn.alloc(sizeof(double) * 2 * n));
    TRY(dk.alloc(sizeof(double) * n));
    TRY(dkc.alloc(sizeof(double) * 2 * n));
    TRY(launch_cell_params(n, dsc.as<double>(), deps.as<double>(), k0, quirks, da.as<double>(), dche.as<double>(),
                           dzn.as<double>(), dk.as<double>(), dkc.as<double>(), st));
    if (a_n) MIDAS_HIP(hipMemcpyAsync(a_n, da.p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    if (che) MIDAS_HIP(hipMemcpyAsync(che, dche.p, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, st));
    if (znnpp) MIDAS_HIP(hipMemcpyAsync(znnpp, dzn.p, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, st));
    if (kappa) MIDAS_HIP(hipMemcpyAsync(kappa, dk.p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    MIDAS_HIP(hipStreamSynchronize(st));
    return 0;
}

int midas_mom_fill(midas_ctx *ctx, const midas_cells_t *cells, double k0, int quirks, double *Z, long ldz)
{
    MIDAS_CHECK(ctx && cells && Z, -1, "mom_fill: bad arguments");
    MIDAS_HIP(hipSetDevice(ctx->device));
    const int n = cells->n, n3 = 3 * cells->n;
    MIDAS_CHECK(ldz >= n3, -1, "mom_fill: ldz %ld < %d", ldz, n3);
    DevCellsOwned dc;
    TRY(dc.from_host(cells, k0, quirks, ctx->st));
    DBuf dz;
    TRY(dz.alloc(sizeof(double2) * (size_t)n3 * n3));
    TRY(launch_fill(dc.v, CellSet{n, 0, nullptr}, CellSet{n, 0, nullptr}, k0, quirks, false, dz.as<double2>(), n3,
                    nullptr, 0, ctx->st));
    MIDAS_HIP(hipMemcpy2DAsync(Z, sizeof(double2) * ldz, dz.p, sizeof(double2) * n3, sizeof(double2) * n3, n3,
                               hipMemcpyDeviceToHost, ctx->st));
    MIDAS_HIP(hipStreamSynchronize(ctx->st));
    return 0;
}

int midas_block_fill(midas_ctx *ctx, const midas_cells_t *b1, const midas_cells_t *b2, double k0, int quirks,
                     double *Z, long ldz)
{
    MIDAS_CHECK(ctx && b1 && b2 && Z, -1, "block_fill: bad arguments");
    MIDAS_HIP(hipSetDevice(ctx->device));
    const int n1 = b1->n, n2 = b2->n;
    MIDAS_CHECK(ldz >= 3L * n1, -1, "block_fill: ldz too small");
    // one device table: receivers first, then sources
    std::vector<double> x(n1 + n2), y(n1 + n2), z(n1 + n2), sc(n1 + n2), kap(n1 + n2), che(2 * (n1 + n2)),
        zn(2 * (n1 + n2));
    std::vector<int> id(n1 + n2);
    // constants: compute each set's own (receivers' constants are not used by the element)
    std::vector<double> a(n1 + n2);
    auto fillset = [&](const midas_cells_t *c, int off) -> int {
        for (int i = 0; i < c->n; ++i) {
            x[off + i] = c->x[i];
            y[off + i] = c->y[i];
            z[off + i] = c->z[i];
            sc[off + i] = c->sc ? c->sc[i] : 0.0;
            id[off + i] = c->id ? c->id[i] : i + 1;
        }
        if (c->kappa && c->che && c->znnpp) {
            for (int i = 0; i < c->n; ++i) {
                kap[off + i] = c->kappa[i];
                che[2 * (off + i)] = c->che[2 * i];
                che[2 * (off + i) + 1] = c->che[2 * i + 1];
                zn[2 * (off + i)] = c->znnpp[2 * i];
                zn[2 * (off + i) + 1] = c->znnpp[2 * i + 1];
            }
            return 0;
        }
        MIDAS_CHECK(c->sc && c->eps, -1, "block_fill: need sc/eps or constants");
        return midas_cell_params(ctx, c->n, c->sc, c->eps, k0, quirks, a.data() + off, che.data() + 2 * off,
                                 zn.data() + 2 * off, kap.data() + off);
    };
    TRY(fillset(b1, 0));
    TRY(fillset(b2, n1));
    midas_cells_t all{n1 + n2, x.data(), y.data(), z.data(), sc.data(), nullptr, kap.data(), che.data(), zn.data(),
                      id.data()};
    DevCellsOwned dc;
    TRY(dc.from_host(&all, k0, quirks, ctx->st));
    DBuf dz;
    TRY(dz.alloc(sizeof(double2) * (size_t)3 * n1 * 3 * n2));
    TRY(launch_fill(dc.v, CellSet{n1, 0, nullptr}, CellSet{n2, n1, nullptr}, k0, quirks,
                    (quirks & MIDAS_Q_FP32_BLOCKS) != 0, dz.as<double2>(), 3L * n1, nullptr, 0, ctx->st));
    MIDAS_HIP(hipMemcpy2DAsync(Z, sizeof(double2) * ldz, dz.p, sizeof(double2) * 3 * n1, sizeof(double2) * 3 * n1,
                               3 * n2, hipMemcpyDeviceToHost, ctx->st));
    MIDAS_HIP(hipStreamSynchronize(ctx->st));
    return 0;
}

int midas_plane_wave(midas_ctx *ctx, const midas_cells_t *cells, double k0, int quirks, int ntr, const double *th,
                     const double *ph, int ref32, double *V, long ldv)
{
    MIDAS_CHECK(ctx && cells && th && ph && V && ntr > 0, -1, "plane_wave: bad arguments");
    MIDAS_HIP(hipSetDevice(ctx->device));
    const int n = cells->n;
    MIDAS_CHECK(ldv >= 3L * n, -1, "plane_wave: ldv too small");
    DevCellsOwned dc;
    midas_cells_t geo = *cells;
    std::vector<double> zr;
    if (!(geo.kappa && geo.che && geo.znnpp) && !(geo.sc && geo.eps)) {
        // only the geometry enters the excitation: neutral constants
        zr.assign(2 * (size_t)n, 0.0);
        geo.kappa = zr.data();
        geo.che = zr.data();
        geo.znnpp = zr.data();
    }
    TRY(dc.from_host(&geo, k0, quirks, ctx->st));
    DBuf dth, dph, dv;
    TRY(upload(dth, th, sizeof(double) * ntr, ctx->st));
    TRY(upload(dph, ph, sizeof(double) * ntr, ctx->st));
    TRY(dv.alloc(sizeof(double2) * 3L * n * 2 * ntr));
    TRY(launch_rhs(dc.v, CellSet{n, 0, nullptr}, ntr, dth.as<double>(), dph.as<double>(), k0, quirks, ref32 != 0,
                   dv.as<double2>(), 3L * n, ctx->st));
    MIDAS_HIP(hipMemcpy2DAsync(V, sizeof(double2) * ldv, dv.p, sizeof(double2) * 3 * n, sizeof(double2) * 3 * n,
                               2 * ntr, hipMemcpyDeviceToHost, ctx->st));
    MIDAS_HIP(hipStreamSynchronize(ctx->st));
    return 0;
}

int midas_spherical_wave(midas_ctx *ctx, const midas_cells_t *cells, double k0, int quirks, int ntr, const double *th,
                         const double *ph, int ref32, double *V, long ldv)
{
    MIDAS_CHECK(ctx && cells && th && ph && V && ntr > 0 && cells->n > 0, -1, "spherical_wave: bad arguments");
    MIDAS_HIP(hipSetDevice(ctx->device));
    const int n = cells->n;
    MIDAS_CHECK(ldv >= 3L * n, -1, "spherical_wave: ldv too small");
    DevCellsOwned dc;
    midas_cells_t geo = *cells;
    std::vector<double> zr;
    if (!(geo.kappa && geo.che && geo.znnpp) && !(geo.sc && geo.eps)) {
        zr.assign(2 * (size_t)n, 0.0);
        geo.kappa = zr.data();
        geo.che = zr.data();
        geo.znnpp = zr.data();
    }
    TRY(dc.from_host(&geo, k0, quirks, ctx->st));
    double cen[3], rs;
    sph_source_frame(cells->x, cells->y, cells->z, nullptr, n, k0, quirks, cen, &rs);
    DBuf dth, dph, dv;
    TRY(upload(dth, th, sizeof(double) * ntr, ctx->st));
    TRY(upload(dph, ph, sizeof(double) * ntr, ctx->st));
    TRY(dv.alloc(sizeof(double2) * 3L * n * 2 * ntr));
    TRY(launch_rhs_sph(dc.v, CellSet{n, 0, nullptr}, ntr, dth.as<double>(), dph.as<double>(), k0, quirks, ref32 != 0,
                       cen, rs, dv.as<double2>(), 3L * n, ctx->st));
    MIDAS_HIP(hipMemcpy2DAsync(V, sizeof(double2) * ldv, dv.p, sizeof(double2) * 3 * n, sizeof(double2) * 3 * n,
                               2 * ntr, hipMemcpyDeviceToHost, ctx->st));
    MIDAS_HIP(hipStreamSynchronize(ctx->st));
    return 0;
}

int midas_solve(midas_ctx *ctx, int n, int nrhs, const double *Z, long ldz, double *B, long ldb, int mode,
                const midas_solve_opts_t *opts, midas_solve_info_t *info)
{
    MIDAS_CHECK(ctx && Z && B && n > 0 && nrhs > 0 && ldz >= n && ldb >= n, -1, "solve: bad arguments");
    MIDAS_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->st;
    DBuf dz, db, dx;
    TRY(dz.alloc(sizeof(double2) * (size_t)n * n));
    TRY(db.alloc(sizeof(double2) * (size_t)n * nrhs));
    TRY(dx.alloc(sizeof(double2) * (size_t)n * nrhs));
    MIDAS_HIP(hipMemcpy2DAsync(dz.p, sizeof(double2) * n, Z, sizeof(double2) * ldz, sizeof(double2) * n, n,
                               hipMemcpyHostToDevice, st));
    MIDAS_HIP(hipMemcpy2DAsync(db.p, sizeof(double2) * n, B, sizeof(double2) * ldb, sizeof(double2) * n, nrhs,
                               hipMemcpyHostToDevice, st));
    // lu_pivot_search = 2 extends the diagonal-block pivot search to these
    // general solves (refinement and its FP64 fallback keep the answer)
    ctx->ws.pivot_search = tune(TUNE_LU_PIVOT_SEARCH) == 2;
    int r = solve_dev(ctx, n, nrhs, dz.as<double2>(), n, nullptr, 0, db.as<double2>(), n, dx.as<double2>(), n, mode,
                      opts, info, true);
    ctx->ws.pivot_search = 0;
    // not converged (no_fallback): the best iterate is still returned in B
    if (r < 0 && r != MIDAS_ERR_NOT_CONVERGED) return r;
    MIDAS_HIP(hipMemcpy2DAsync(B, sizeof(double2) * ldb, dx.p, sizeof(double2) * n, sizeof(double2) * n, nrhs,
                               hipMemcpyDeviceToHost, st));
    MIDAS_HIP(hipStreamSynchronize(st));
    return r;
}

int midas_scattering_matrix(midas_ctx *ctx, const midas_cells_t *cells, double k0, int quirks, const double *E,
                            long lde, int ntr, int nrx, const double *rth, const double *rph, int npol,
                            const double *beta, double *S, long lds)
{
    MIDAS_CHECK(ctx && cells && E && rth && rph && S && ntr > 0 && nrx > 0 && npol > 0, -1,
                "scattering_matrix: bad arguments");
    MIDAS_CHECK(lde >= 3L * cells->n && lds >= nrx, -1, "scattering_matrix: leading dimension too small");
    MIDAS_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->st;
    const int n = cells->n;
    DevCellsOwned dc;
    TRY(dc.from_host(cells, k0, quirks, st));
    std::vector<double> b0(npol, 0.0);
    DBuf dE, drt, drp, db, dS, FF, E3, G;
    TRY(dE.alloc(sizeof(double2) * 3L * n * 2 * ntr));
    MIDAS_HIP(hipMemcpy2DAsync(dE.p, sizeof(double2) * 3 * n, E, sizeof(double2) * lde, sizeof(double2) * 3 * n,
                               2 * ntr, hipMemcpyHostToDevice, st));
    TRY(upload(drt, rth, sizeof(double) * nrx, st));
    TRY(upload(drp, rph, sizeof(double) * nrx, st));
    TRY(upload(db, beta ? beta : b0.data(), sizeof(double) * npol, st));
    TRY(dS.alloc(sizeof(double2) * (size_t)nrx * 4 * ntr * npol));
    TRY(FF.alloc(sizeof(double2) * (size_t)nrx * n));
    TRY(E3.alloc(sizeof(double2) * (size_t)n * 6 * ntr));
    TRY(G.alloc(sizeof(double2) * (size_t)nrx * 6 * ntr));
    TRY(far_field(dc.v, dE.as<double2>(), 3L * n, ntr, nrx, drt.as<double>(), drp.as<double>(), npol,
                  db.as<double>(), k0, quirks, dS.as<double2>(), nrx, FF.as<double2>(), E3.as<double2>(),
                  G.as<double2>(), st));
    MIDAS_HIP(hipMemcpy2DAsync(S, sizeof(double2) * lds, dS.p, sizeof(double2) * nrx, sizeof(double2) * nrx,
                               4L * ntr * npol, hipMemcpyDeviceToHost, st));
    MIDAS_HIP(hipStreamSynchronize(st));
    return 0;
}

int midas_cext_cabs(midas_ctx *ctx, const midas_cells_t *cells, double k0, int quirks, const double *E, long lde,
                    int ntr, const double *th, const double *ph, double *C_ext, double *C_abs)
{
    MIDAS_CHECK(ctx && cells && E && th && ph && C_ext && C_abs && ntr > 0, -1, "cext_cabs: bad arguments");
    MIDAS_CHECK(lde >= 3L * cells->n, -1, "cext_cabs: lde too small");
    MIDAS_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->st;
    const int n = cells->n;
    DevCellsOwned dc;
    TRY(dc.from_host(cells, k0, quirks, st));
    DBuf dE, dth, dph, dce, dca;
    TRY(dE.alloc(sizeof(double2) * 3L * n * 2 * ntr));
    MIDAS_HIP(hipMemcpy2DAsync(dE.p, sizeof(double2) * 3 * n, E, sizeof(double2) * lde, sizeof(double2) * 3 * n,
                               2 * ntr, hipMemcpyHostToDevice, st));
    TRY(upload(dth, th, sizeof(double) * ntr, st));
    TRY(upload(dph, ph, sizeof(double) * ntr, st));
    TRY(dce.alloc(sizeof(double2) * ntr));
    TRY(dca.alloc(sizeof(double2) * ntr));
    TRY(cext_cabs(dc.v, dE.as<double2>(), 3L * n, ntr, dth.as<double>(), dph.as<double>(), k0, quirks,
                  dce.as<double>(), dca.as<double>(), st));
    MIDAS_HIP(hipMemcpyAsync(C_ext, dce.p, sizeof(double2) * ntr, hipMemcpyDeviceToHost, st));
    MIDAS_HIP(hipMemcpyAsync(C_abs, dca.p, sizeof(double2) * ntr, hipMemcpyDeviceToHost, st));
    MIDAS_HIP(hipStreamSynchronize(st));
    return 0;
}

int midas_near_field(midas_ctx *ctx, const midas_cells_t *cells, double k0, int quirks, const double *E, long lde,
                     int ntr, int nrx, const double *rth, const double *rph, double rso, double *Es, long lds)
{
    MIDAS_CHECK(ctx && cells && E && rth && rph && Es && ntr > 0 && nrx > 0 && rso > 0, -1,
                "near_field: bad arguments");
    MIDAS_CHECK(lde >= 3L * cells->n && lds >= nrx, -1, "near_field: lde / lds too small");
    MIDAS_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->st;
    const int n = cells->n;
    DevCellsOwned dc;
    TRY(dc.from_host(cells, k0, quirks, st));
    DBuf dE, drt, drp, dS;
    TRY(dE.alloc(sizeof(double2) * 3L * n * 2 * ntr));
    MIDAS_HIP(hipMemcpy2DAsync(dE.p, sizeof(double2) * 3 * n, E, sizeof(double2) * lde, sizeof(double2) * 3 * n,
                               2 * ntr, hipMemcpyHostToDevice, st));
    TRY(upload(drt, rth, sizeof(double) * nrx, st));
    TRY(upload(drp, rph, sizeof(double) * nrx, st));
    TRY(dS.alloc(sizeof(double2) * (size_t)nrx * 4 * ntr));
    TRY(near_field(dc.v, dE.as<double2>(), 3L * n, ntr, nrx, drt.as<double>(), drp.as<double>(), rso, k0, quirks,
                   dS.as<double2>(), nrx, st));
    MIDAS_HIP(hipMemcpy2DAsync(Es, sizeof(double2) * lds, dS.p, sizeof(double2) * nrx, sizeof(double2) * nrx, 4 * ntr,
                               hipMemcpyDeviceToHost, st));
    MIDAS_HIP(hipStreamSynchronize(st));
    return 0;
}

int midas_scattering_quantities(midas_ctx *ctx, const double *S, long lds, int nrx_tot, int nrx, int ntr,
                                const double *rx_th, const double *rx_ph, int rx_type, const double *rx_w,
                                const double *tr_th, const double *tr_ph, int tr_type, const double *tr_w, double k0,
                                double dm, int quirks, const double *C_ext, const double *C_abs, int cext_intfields,
                                double *Q, double *Q_av)
{
    MIDAS_CHECK(ctx && S && rx_th && rx_ph && tr_th && tr_ph && C_ext && C_abs && Q && Q_av, -1,
                "scattering_quantities: bad arguments");
    MIDAS_CHECK(nrx > 0 && ntr > 0 && lds >= nrx_tot && k0 > 0 && dm > 0, -1,
                "scattering_quantities: nrx %d, ntr %d, lds %ld, nrx_tot %d", nrx, ntr, lds, nrx_tot);
    MIDAS_CHECK(rx_type >= MIDAS_RX_SD && rx_type <= MIDAS_RX_LB && tr_type >= MIDAS_RX_SD && tr_type <= MIDAS_RX_LB,
                -1, "scattering_quantities: receiver / transmitter rule %d / %d", rx_type, tr_type);
    MIDAS_CHECK(rx_type != MIDAS_RX_LB || rx_w, -1, "scattering_quantities: lb receivers need weights");
    MIDAS_CHECK(tr_type != MIDAS_RX_LB || ntr == 1 || tr_w, -1, "scattering_quantities: lb transmitters need weights");
    // forward / backward receivers (Compute_Scattering_Quantities.f90:773-777, 867-871)
    const bool same = ntr == nrx;
    auto fw = [&](int t) { return same && (rx_type == MIDAS_RX_SD || rx_type == MIDAS_RX_LB) ? t : nrx + t; };
    auto bk = [&](int t) { return same ? nrx + t : nrx + ntr + t; };
    MIDAS_CHECK(bk(ntr - 1) < nrx_tot && fw(ntr - 1) < nrx_tot, -1,
                "scattering_quantities: nrx_tot %d has no forward / backward receivers", nrx_tot);
    MIDAS_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->st;
    // the context's buffers, grown when needed
    DBuf &dS = pooled(ctx, "q.S"), &drt = pooled(ctx, "q.rth"), &drp = pooled(ctx, "q.rph"), &dw = pooled(ctx, "q.w"),
         &dtt = pooled(ctx, "q.tth"), &dtp = pooled(ctx, "q.tph"), &dout = pooled(ctx, "q.sums");
    TRY(dS.reserve(sizeof(double2) * (size_t)nrx * 4 * ntr));
    MIDAS_HIP(hipMemcpy2DAsync(dS.p, sizeof(double2) * nrx, S, sizeof(double2) * lds, sizeof(double2) * nrx, 4 * ntr,
                               hipMemcpyHostToDevice, st));
    TRY(upload_keep(drt, rx_th, sizeof(double) * nrx, st));
    TRY(upload_keep(drp, rx_ph, sizeof(double) * nrx, st));
    if (rx_type == MIDAS_RX_LB) TRY(upload_keep(dw, rx_w, sizeof(double) * nrx, st));
    TRY(upload_keep(dtt, tr_th, sizeof(double) * ntr, st));
    TRY(upload_keep(dtp, tr_ph, sizeof(double) * ntr, st));
    TRY(dout.reserve(sizeof(double) * 2 * ntr));
    TRY(scatq_sums(dS.as<double2>(), nrx, nrx, ntr, drt.as<double>(), drp.as<double>(),
                   rx_type == MIDAS_RX_LB ? dw.as<double>() : nullptr, dtt.as<double>(), dtp.as<double>(), quirks,
                   dout.as<double>(), st));
    std::vector<double> sums(2 * ntr);
    TRY(copy_sync(sums.data(), dout.p, sizeof(double) * 2 * ntr, hipMemcpyDeviceToHost, st));
    // per transmitter, the reference's expressions and literals
    const double pi = host_pi(quirks), ap = dm / 2.0, X = k0 * dm / 2.0;
    auto sq = [](double re, double im) {
        const double m = std::hypot(re, im);
        return m * m;
    };
    auto s_at = [&](int r, int c) { return S + 2 * ((long)r + (long)c * lds); };
    for (int t = 0; t < ntr; ++t) {
        const double *f = s_at(fw(t), 4 * t), *fh = s_at(fw(t), 4 * t + 3);
        const double q_ext = (2 / (X * X)) * (std::fabs(fh[1]) + std::fabs(f[1]));
        // C / (pi ap^2) taken to its real part, as the real(8) assignment does
        const double q_ext_intf = C_ext[2 * t] / (pi * ap * ap);
        const double q_abs_intf = C_abs[2 * t] / (pi * ap * ap);
        double q_sca, g;
        if (rx_type == MIDAS_RX_LB) {
            const double q_sd = sums[2 * t];
            const double c_sca = 1.0 / (2 * k0 * k0) * q_sd;
            q_sca = (4 * pi) / (2 * pi * X * X) * q_sd;
            g = 1.0 / (2.0 * k0 * k0 * c_sca) * sums[2 * t + 1];
        } else {
            const double q_sd = (4 * pi / nrx) * sums[2 * t];
            const double c_sca = 1.0 / (2 * k0 * k0) * q_sd;
            q_sca = 1 / (2 * pi * X * X) * q_sd;
            g = 1.0 / (2.0 * k0 * k0 * c_sca) * (4 * pi / nrx) * sums[2 * t + 1];
        }
        const double *b1 = s_at(bk(t), 4 * t), *b2 = s_at(bk(t), 4 * t + 1), *b3 = s_at(bk(t), 4 * t + 2),
                     *b4 = s_at(bk(t), 4 * t + 3);
        const double q_bks = (1.0 / (2.0 * pi * X * X)) *
                             (((sq(b1[0], b1[1]) + sq(b4[0], b4[1])) + sq(b2[0], b2[1])) + sq(b3[0], b3[1]));
        double *q = Q + 5L * t;
        q[0] = cext_intfields ? q_ext_intf : q_ext;
        q[1] = q_abs_intf;   // Q_abs is always the internal-field value (:877)
        q[2] = q_sca;
        q[3] = q_bks;
        q[4] = g;
    }
    // orientation averages (:896-918)
    for (int j = 0; j < 5; ++j) {
        double a = 0.0;
        if (ntr == 1) {
            a = Q[j];
        } else if (tr_type != MIDAS_RX_LB) {
            for (int t = 0; t < ntr; ++t) a += Q[5L * t + j];
            a /= ntr;
        } else {
            for (int t = 0; t < ntr; ++t) a = a + tr_w[t] * Q[5L * t + j];
        }
        Q_av[j] = a;
    }
    return 0;
}

// Fill (FP64 Z, and the FP32 copy W for mixed modes) + RHS + solve, all on
// the device; X (3n x 2ntr).  Shared by the drop-in and the device entry point.
// The MoM mixed solves with few right-hand sides take their FP64 residuals
// from the cells (the FP64 matrix is not stored, only its FP32 copy).
static bool mom_matrix_free(const SolveWs &ws, int mode, int ntr)
{
    return mode != MIDAS_SOLVE_FP64 && 2 * ntr <= 4 && !ws.lent && tune(TUNE_MOM_MATRIX_FREE);
}

static int mom_solve_dev(midas_ctx *ctx, SolveWs &ws, const DevCells &cells, double k0, int quirks, int ntr,
                         const double *dth, const double *dph, int mode, const midas_solve_opts_t *opts, double2 *Z,
                         long ldz, double2 *X, long ldx, midas_solve_info_t *info, double *t_fill_ms)
{
    hipStream_t st = ctx->st;
    const int n = cells.n, n3 = 3 * n;
    const bool mf = mom_matrix_free(ws, mode, ntr);
    MIDAS_CHECK(Z || mf, -1, "mom_solve: no FP64 matrix buffer");
    float2 *w32 = nullptr;
    if (mode != MIDAS_SOLVE_FP64) {
        TRY(ws.W.reserve(sizeof(float2) * (size_t)n3 * n3));
        w32 = ws.W.as<float2>();
    }
    TRY(ws.V.reserve(sizeof(double2) * (size_t)n3 * 2 * ntr));
    Timer tm(st, !ws.lent);
    TRY(launch_fill(cells, CellSet{n, 0, nullptr}, CellSet{n, 0, nullptr}, k0, quirks, false, mf ? nullptr : Z, ldz,
                    w32, n3, st));
    TRY(launch_rhs(cells, CellSet{n, 0, nullptr}, ntr, dth, dph, k0, quirks, false, ws.V.as<double2>(), n3, st));
    double tf = tm.ms();
    if (t_fill_ms) *t_fill_ms = tf;
    if (mf) {
        ws.mf_cells = &cells;
        ws.mf_k0 = k0;
        ws.mf_pi = host_pi(quirks);
        ws.mf_quirks = quirks;
    }
    // The MoM matrix (1 - Znnpp on the diagonal, Green's couplings scaled by
    // the cells' contrast off it) takes its pivots on the diagonal under
    // partial pivoting (LAPACK's zgetrf makes no interchange on the dielectric
    // scatterers here): its mixed LU searches each panel's diagonal block only
    // and reports the certificate (max |L21|); refinement and its FP64
    // fallback keep the answer whatever the pivots.
    ws.pivot_search = tune(TUNE_LU_PIVOT_SEARCH);
    const int r = solve_dev(ctx, n3, 2 * ntr, Z, ldz, w32, n3, ws.V.as<double2>(), n3, X, ldx, mode, opts, info, true,
                            &ws);
    ws.pivot_search = 0;
    ws.mf_cells = nullptr;
    return r;
}

int midas_compute_efields_mom(midas_ctx *ctx, const midas_cells_t *cells, double k0, int quirks, int ntr,
                              const double *tr_th, const double *tr_ph, int nrx, const double *rx_th,
                              const double *rx_ph, int mode, const midas_solve_opts_t *opts, double *S_total,
                              double *C_ext, double *C_abs, double *E, midas_solve_info_t *info)
{
    MIDAS_CHECK(ctx && cells && tr_th && tr_ph && ntr > 0, -1, "compute_efields_mom: bad arguments");
    MIDAS_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->st;
    const int n = cells->n, n3 = 3 * n;
    // the context's buffers, grown when needed (none made or freed per call)
    DevCellsOwned &dc = ctx->call_cells;
    TRY(dc.from_host(cells, k0, quirks, st));
    DBuf &X = pooled(ctx, "mom.X"), &dth = pooled(ctx, "mom.th"), &dph = pooled(ctx, "mom.ph");
    DBuf Zloc;
    DBuf &Z = mom_matrix_free(ctx->ws, mode, ntr) ? Zloc : pooled(ctx, "mom.Z");
    // (a matrix-free mixed solve needs no FP64 matrix unless it falls back)
    if (!mom_matrix_free(ctx->ws, mode, ntr)) TRY(Z.reserve(sizeof(double2) * (size_t)n3 * n3));
    TRY(X.reserve(sizeof(double2) * (size_t)n3 * 2 * ntr));
    TRY(upload_keep(dth, tr_th, sizeof(double) * ntr, st));
    TRY(upload_keep(dph, tr_ph, sizeof(double) * ntr, st));
    int r = mom_solve_dev(ctx, ctx->ws, dc.v, k0, quirks, ntr, dth.as<double>(), dph.as<double>(), mode, opts,
                          Z.as<double2>(), n3, X.as<double2>(), n3, info, nullptr);
    // a refinement that missed its tolerance (no_fallback) still delivers the
    // outputs of its best iterate, then reports
    if (r < 0 && r != MIDAS_ERR_NOT_CONVERGED) return r;
    int singular = r;
    std::string nc_msg = r == MIDAS_ERR_NOT_CONVERGED ? midas::last_error() : "";
    if (S_total && nrx > 0) {
        DBuf &drt = pooled(ctx, "ff.rth"), &drp = pooled(ctx, "ff.rph"), &db = pooled(ctx, "ff.beta"),
             &dS = pooled(ctx, "ff.S"), &FF = pooled(ctx, "ff.coef"), &E3 = pooled(ctx, "ff.E3"),
             &G = pooled(ctx, "ff.G");
        double b0 = 0.0;
        TRY(upload_keep(drt, rx_th, sizeof(double) * nrx, st));
        TRY(upload_keep(drp, rx_ph, sizeof(double) * nrx, st));
        TRY(upload_keep(db, &b0, sizeof(double), st));
        TRY(dS.reserve(sizeof(double2) * (size_t)nrx * 4 * ntr));
        TRY(FF.reserve(sizeof(double2) * (size_t)nrx * n));
        TRY(E3.reserve(sizeof(double2) * (size_t)n * 6 * ntr));
        TRY(G.reserve(sizeof(double2) * (size_t)nrx * 6 * ntr));
        TRY(far_field(dc.v, X.as<double2>(), n3, ntr, nrx, drt.as<double>(), drp.as<double>(), 1, db.as<double>(), k0,
                      quirks, dS.as<double2>(), nrx, FF.as<double2>(), E3.as<double2>(), G.as<double2>(), st));
        MIDAS_HIP(hipMemcpyAsync(S_total, dS.p, sizeof(double2) * (size_t)nrx * 4 * ntr, hipMemcpyDeviceToHost, st));
        MIDAS_HIP(hipStreamSynchronize(st));
    }
    if (C_ext && C_abs) {
        DBuf &dce = pooled(ctx, "cext"), &dca = pooled(ctx, "cabs");
        TRY(dce.reserve(sizeof(double2) * ntr));
        TRY(dca.reserve(sizeof(double2) * ntr));
        TRY(cext_cabs(dc.v, X.as<double2>(), n3, ntr, dth.as<double>(), dph.as<double>(), k0, quirks, dce.as<double>(),
                      dca.as<double>(), st));
        MIDAS_HIP(hipMemcpyAsync(C_ext, dce.p, sizeof(double2) * ntr, hipMemcpyDeviceToHost, st));
        MIDAS_HIP(hipMemcpyAsync(C_abs, dca.p, sizeof(double2) * ntr, hipMemcpyDeviceToHost, st));
    }
    if (E) MIDAS_HIP(hipMemcpyAsync(E, X.p, sizeof(double2) * (size_t)n3 * 2 * ntr, hipMemcpyDeviceToHost, st));
    MIDAS_HIP(hipStreamSynchronize(st));
    if (singular == MIDAS_ERR_NOT_CONVERGED) midas::set_error("%s", nc_msg.c_str());
    return singular;
}

int midas_dev_mom_fill(midas_ctx *ctx, const midas_cells_t *dcells, double k0, int quirks, double *Z, long ldz,
                       float *W32, long ldw)
{
    MIDAS_CHECK(ctx && dcells && (Z || W32), -1, "dev_mom_fill: bad arguments");
    MIDAS_HIP(hipSetDevice(ctx->device));
    const int n = dcells->n;
    DevCellsOwned dc;
    TRY(dc.from_device(dcells, k0, quirks, ctx->st));
    TRY(launch_fill(dc.v, CellSet{n, 0, nullptr}, CellSet{n, 0, nullptr}, k0, quirks, false, (double2 *)Z, ldz,
                    (float2 *)W32, ldw, ctx->st));
    MIDAS_HIP(hipStreamSynchronize(ctx->st));
    return 0;
}

int midas_dev_mom_solve(midas_ctx *ctx, const midas_cells_t *dcells, double k0, int quirks, int ntr, const double *dth,
                        const double *dph, int mode, const midas_solve_opts_t *opts, double *Z, long ldz, double *X,
                        long ldx, midas_solve_info_t *info, double *t_fill_ms)
{
    MIDAS_CHECK(ctx && dcells && dth && dph && Z && X && ntr > 0, -1, "dev_mom_solve: bad arguments");
    MIDAS_CHECK(ldz >= 3L * dcells->n && ldx >= 3L * dcells->n, -1, "dev_mom_solve: leading dimension too small");
    MIDAS_HIP(hipSetDevice(ctx->device));
    DevCellsOwned &dc = ctx->cells;
    TRY(dc.from_device(dcells, k0, quirks, ctx->st));
    int r = mom_solve_dev(ctx, ctx->ws, dc.v, k0, quirks, ntr, dth, dph, mode, opts, (double2 *)Z, ldz, (double2 *)X,
                          ldx, info, t_fill_ms);
    MIDAS_HIP(hipStreamSynchronize(ctx->st));
    return r;
}

// ---------------------------------------------------------------------------
// Allocation-free variant of midas_dev_mom_solve: every work buffer is carved
// from the device's static scratch array (its address from midas_init).
// Unknowns n <= NOALLOC_MAX_N, right-hand sides <= NOALLOC_MAX_RHS; panels of
// 256 columns (mixed modes) / 128 (FP64 and the FP64 fallback); the trailing
// updates, residuals and FP64 products run on this library's kernels (no
// rocBLAS); the factorisation runs on the context's stream alone (no second
// stream).  Calls on one device share the array: calls that overlap race.
// ---------------------------------------------------------------------------
constexpr int NOALLOC_MAX_N = 2048, NOALLOC_MAX_RHS = 32, NOALLOC_NB = 256, NOALLOC_NB64 = 128;

struct NoallocPlan {
    size_t off[20];
    size_t total;
};
// offsets (256-byte aligned) of the buffers, in the order used by noalloc_lend
static NoallocPlan noalloc_plan(int ncell, int nrhs)
{
    const int n = 3 * ncell;
    const size_t d = sizeof(double) * ncell, z2 = 2 * d;
    const size_t sz[] = {
        d, z2, z2, z2, d, sizeof(int) * (size_t)ncell,                     // cells: kappa che znnpp kchi a id
        sizeof(float2) * (size_t)n * n,                                     // W
        trail16_scratch_bytes(n, n, NOALLOC_NB),                            // trail
        sizeof(int) * (size_t)n, sizeof(int) * (size_t)n,                   // ipiv perm
        sizeof(double) * (size_t)(n + 2 * nrhs + 4),                        // norms
        zx_scratch_bytes(n, nrhs <= 4 ? nrhs : 1),                          // zx
        sizeof(double2) * (size_t)n * nrhs,                                 // R
        sizeof(float2) * (size_t)n * nrhs,                                  // D
        sizeof(double2) * (size_t)getrs_tmp_elems(n, NOALLOC_NB, nrhs),     // t2 (FP32 or FP64 use)
        sizeof(double2) * (size_t)n * nrhs,                                 // xbest
        sizeof(double2) * (size_t)n * nrhs,                                 // V
        LuWork<float>::bytes_needed(n, NOALLOC_NB),                         // luf
        LuWork<double>::bytes_needed(n, NOALLOC_NB64),                      // lud
    };
    NoallocPlan p{};
    size_t o = 0;
    for (size_t i = 0; i < sizeof(sz) / sizeof(sz[0]); ++i) {
        p.off[i] = o;
        o += (sz[i] + 255) & ~(size_t)255;
    }
    p.off[19] = o;
    p.total = o;
    return p;
}

static int noalloc_lend(midas_ctx *ctx, int ncell, int nrhs)
{
    const NoallocPlan p = noalloc_plan(ncell, nrhs);
    MIDAS_CHECK(p.total <= NOALLOC_SCRATCH_BYTES, -1, "noalloc: %zu bytes of scratch needed, %zu held", p.total,
                NOALLOC_SCRATCH_BYTES);
    char *b = (char *)ctx->noalloc_base;
    auto len = [&](int i) { return p.off[i + 1] - p.off[i]; };
    DevCellsOwned &c = ctx->ncells;
    DBuf *cb[] = {&c.kappa, &c.che, &c.znnpp, &c.kchi, &c.a, &c.id};
    for (int i = 0; i < 6; ++i) cb[i]->borrow(b + p.off[i], len(i));
    SolveWs &w = ctx->nws;
    DBuf *wb[] = {&w.W, &w.trail, &w.ipiv, &w.perm, &w.norms, &w.zx, &w.R, &w.D, &w.t2, &w.xbest, &w.V};
    for (int i = 0; i < 11; ++i) wb[i]->borrow(b + p.off[6 + i], len(6 + i));
    w.zcopy.release();   // Z is factored in place: no copy is ever asked for
    w.lent = true;
    const int n = 3 * ncell;
    TRY(w.luf.lend(b + p.off[17], len(17), n, NOALLOC_NB, 1, ctx->st, ctx->nev[0], ctx->nev[1]));
    TRY(w.lud.lend(b + p.off[18], len(18), n, NOALLOC_NB64, 1, ctx->st, ctx->nev[0], ctx->nev[1]));
    return 0;
}

int midas_dev_mom_solve_noalloc(midas_ctx *ctx, const midas_cells_t *dcells, double k0, int quirks, int ntr,
                                const double *dth, const double *dph, int mode, const midas_solve_opts_t *opts,
                                double *Z, long ldz, double *X, long ldx, midas_solve_info_t *info)
{
    MIDAS_CHECK(ctx && dcells && dth && dph && Z && X && ntr > 0, -1, "dev_mom_solve_noalloc: bad arguments");
    MIDAS_CHECK(ctx->noalloc_base, -1, "dev_mom_solve_noalloc: no scratch address (midas_init)");
    const int ncell = dcells->n;
    MIDAS_CHECK(ncell > 0 && 3L * ncell <= NOALLOC_MAX_N && 2L * ntr <= NOALLOC_MAX_RHS, -1,
                "dev_mom_solve_noalloc: %d unknowns and %d right-hand sides; at most %d and %d", 3 * ncell, 2 * ntr,
                NOALLOC_MAX_N, NOALLOC_MAX_RHS);
    MIDAS_CHECK(ldz >= 3L * ncell && ldx >= 3L * ncell, -1, "dev_mom_solve_noalloc: leading dimension too small");
    MIDAS_CHECK(mode >= 0 && mode <= 5, -1, "dev_mom_solve_noalloc: unknown mode %d", mode);
    MIDAS_HIP(hipSetDevice(ctx->device));
    TRY(noalloc_lend(ctx, ncell, 2 * ntr));
    midas_solve_opts_t o{};
    if (opts) o = *opts;
    o.nb = mode == MIDAS_SOLVE_FP64 ? NOALLOC_NB64 : NOALLOC_NB;
    TRY(ctx->ncells.from_device(dcells, k0, quirks, ctx->st));
    int r = mom_solve_dev(ctx, ctx->nws, ctx->ncells.v, k0, quirks, ntr, dth, dph, mode, &o, (double2 *)Z, ldz,
                          (double2 *)X, ldx, info, nullptr);
    MIDAS_HIP(hipStreamSynchronize(ctx->st));
    return r;
}

// ---------------------------------------------------------------------------
// kernel-level test hooks (include/midas_gpu_test.h)
// ---------------------------------------------------------------------------
__global__ void k_c2z(const float2 *S, double2 *D, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) D[i] = make_double2(S[i].x, S[i].y);
}
__global__ void k_z2c1(const double2 *S, float2 *D, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) D[i] = make_float2((float)S[i].x, (float)S[i].y);
}

static int upload_mat(DBuf &b, const double *h, int rows, int cols, long ld, hipStream_t st)
{
    TRY(b.alloc(sizeof(double2) * (size_t)rows * cols));
    if ((size_t)rows * cols == 0) return 0;
    MIDAS_HIP(hipMemcpy2DAsync(b.p, sizeof(double2) * rows, h, sizeof(double2) * ld, sizeof(double2) * rows, cols,
                               hipMemcpyHostToDevice, st));
    return 0;
}
static int to_c64(DBuf &d32, const DBuf &d64, long cnt, hipStream_t st)
{
    TRY(d32.alloc(sizeof(float2) * cnt));
    hipLaunchKernelGGL(k_z2c1, dim3((cnt + 255) / 256), dim3(256), 0, st, d64.as<double2>(), d32.as<float2>(), cnt);
    MIDAS_HIP(hipGetLastError());
    return 0;
}

int midas_test_gemm(midas_ctx *ctx, int prec, int m, int n, int k, const double *alpha, const double *A, long lda,
                    int transA, const double *B, long ldb, int transB, const double *beta, double *C, long ldc)
{
    MIDAS_CHECK(ctx && A && B && C && alpha && beta && (prec == 64 || prec == 32 || prec == 65 || prec == 66), -1,
                "test_gemm: bad arguments");
    MIDAS_CHECK(prec < 65 || (!transA && !transB), -1, "test_gemm: prec 65 and 66 take no transposes");
    MIDAS_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->st;
    const int nbat = prec == 66 ? 3 : 1;
    int ar = transA ? k : m, ac = transA ? m : k, br = transB ? n : k, bc = transB ? k : n;
    DBuf dA, dB, dC;
    TRY(upload_mat(dA, A, ar, ac * nbat, lda, st));
    TRY(upload_mat(dB, B, br, bc * nbat, ldb, st));
    TRY(upload_mat(dC, C, m, n * nbat, ldc, st));
    if (prec >= 65) {
        TRY(launch_zgemm_nn(m, n, k, z64{alpha[0], alpha[1]}, dA.p, ar, (long)ar * ac, dB.p, br, (long)br * bc,
                            z64{beta[0], beta[1]}, dC.p, m, (long)m * n, nbat, st, tune(TUNE_ZGEMM_TILE)));
    } else if (prec == 64) {
        TRY(launch_cgemm<double>(m, n, k, z64{alpha[0], alpha[1]}, dA.p, ar, transA, dB.p, br, transB,
                                 z64{beta[0], beta[1]}, dC.p, m, st));
    } else {
        DBuf a32, b32, c32;
        TRY(to_c64(a32, dA, (long)ar * ac, st));
        TRY(to_c64(b32, dB, (long)br * bc, st));
        TRY(to_c64(c32, dC, (long)m * n, st));
        TRY(launch_cgemm<float>(m, n, k, midas::c32{(float)alpha[0], (float)alpha[1]}, a32.p, ar, transA, b32.p, br,
                                transB, midas::c32{(float)beta[0], (float)beta[1]}, c32.p, m, st));
        hipLaunchKernelGGL(k_c2z, dim3(((long)m * n + 255) / 256), dim3(256), 0, st, c32.as<float2>(),
                           dC.as<double2>(), (long)m * n);
    }
    MIDAS_HIP(hipMemcpy2DAsync(C, sizeof(double2) * ldc, dC.p, sizeof(double2) * m, sizeof(double2) * m, n * nbat,
                               hipMemcpyDeviceToHost, st));
    MIDAS_HIP(hipStreamSynchronize(st));
    return 0;
}

}   // extern "C"

template <typename T>
static int test_getrf_t(midas_ctx *ctx, int n, typename cvec<T>::t *dA, int nb, int pivot_search, int *dipiv,
                        double *max_l21)
{
    LuWork<T> w;
    w.share_stream(ctx->ps, ctx->pev[0], ctx->pev[1]);
    int r = w.alloc(n, nb);
    DBuf l21;
    if (!r && pivot_search) {
        const size_t sl = sizeof(cplx<T>) * (size_t)n * nb;
        r = l21.alloc(sl + 256);
        w.diag_pivot = true;
        w.l21 = l21.as<cplx<T>>();
        w.lmax = (unsigned *)((char *)l21.p + sl);
    }
    TrailingGemm<T> tg;
    if (!r) r = getrf<T>(dA, n, n, dipiv, w, ctx->st, tg);
    unsigned lm = 0;
    if (!r && pivot_search) r = copy_sync(&lm, w.lmax, sizeof(unsigned), hipMemcpyDeviceToHost, ctx->st);
    if (max_l21) *max_l21 = (double)__builtin_bit_cast(float, lm);
    w.release();
    return r;
}

extern "C" {

int midas_test_getrf(midas_ctx *ctx, int prec, int n, double *A, long lda, int nb, int pivot_search, int *ipiv,
                     double *max_l21)
{
    MIDAS_CHECK(ctx && A && ipiv && n > 0 && lda >= n && nb >= 16 && nb <= 512 && (prec == 64 || prec == 32), -1,
                "test_getrf: bad arguments");
    MIDAS_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->st;
    DBuf dA, dp;
    TRY(upload_mat(dA, A, n, n, lda, st));
    TRY(dp.alloc(sizeof(int) * n));
    if (prec == 64) {
        TRY(test_getrf_t<double>(ctx, n, dA.as<double2>(), nb, pivot_search, dp.as<int>(), max_l21));
    } else {
        DBuf a32;
        TRY(to_c64(a32, dA, (long)n * n, st));
        TRY(test_getrf_t<float>(ctx, n, a32.as<float2>(), nb, pivot_search, dp.as<int>(), max_l21));
        hipLaunchKernelGGL(k_c2z, dim3(((long)n * n + 255) / 256), dim3(256), 0, st, a32.as<float2>(),
                           dA.as<double2>(), (long)n * n);
        MIDAS_HIP(hipGetLastError());
    }
    MIDAS_HIP(hipMemcpy2DAsync(A, sizeof(double2) * lda, dA.p, sizeof(double2) * n, sizeof(double2) * n, n,
                               hipMemcpyDeviceToHost, st));
    MIDAS_HIP(hipMemcpyAsync(ipiv, dp.p, sizeof(int) * n, hipMemcpyDeviceToHost, st));
    MIDAS_HIP(hipStreamSynchronize(st));
    return 0;
}

int midas_test_heev(midas_ctx *ctx, int n, int batch, const double *G, double *D, double *V, double cut, int path)
{
    MIDAS_CHECK(ctx && G && D && V && n > 0 && batch > 0, -1, "test_heev: bad arguments");
    MIDAS_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->st;
    const size_t nn = (size_t)n * n;
    DBuf dG, dD, dI, W;
    TRY(dG.alloc(sizeof(double2) * nn * batch));
    TRY(dD.alloc(sizeof(double) * (size_t)n * batch));
    TRY(dI.alloc(sizeof(int) * batch));
    TRY(W.alloc(heev_work_bytes(n, batch)));
    MIDAS_HIP(hipMemcpyAsync(dG.p, G, sizeof(double2) * nn * batch, hipMemcpyHostToDevice, st));
    TRY(launch_heev_batched(n, dG.as<double2>(), n, (long)nn, dD.as<double>(), n, dI.as<int>(), batch, W.p, st, cut,
                            path));
    std::vector<int> info(batch);
    MIDAS_HIP(hipMemcpyAsync(info.data(), dI.p, sizeof(int) * batch, hipMemcpyDeviceToHost, st));
    MIDAS_HIP(hipMemcpyAsync(D, dD.p, sizeof(double) * (size_t)n * batch, hipMemcpyDeviceToHost, st));
    MIDAS_HIP(hipMemcpyAsync(V, dG.p, sizeof(double2) * nn * batch, hipMemcpyDeviceToHost, st));
    MIDAS_HIP(hipStreamSynchronize(st));
    for (int b = 0; b < batch; ++b) MIDAS_CHECK(info[b] == 0, -2, "test_heev: matrix %d did not converge", b);
    return 0;
}

int midas_test_pchol(midas_ctx *ctx, int n, int batch, const double *G, double *L, int *rank, double tolrel)
{
    MIDAS_CHECK(ctx && G && L && rank && n > 0 && batch > 0, -1, "test_pchol: bad arguments");
    MIDAS_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->st;
    const size_t nn = (size_t)n * n;
    DBuf dG, dL, dR;
    TRY(dG.alloc(sizeof(double2) * nn * batch));
    TRY(dL.alloc(sizeof(double2) * nn * batch));
    TRY(dR.alloc(sizeof(int) * batch));
    MIDAS_HIP(hipMemcpyAsync(dG.p, G, sizeof(double2) * nn * batch, hipMemcpyHostToDevice, st));
    TRY(launch_pchol_batched(n, dG.as<double2>(), n, (long)nn, dL.as<double2>(), n, (long)nn, dR.as<int>(), batch,
                             tolrel, st));
    MIDAS_HIP(hipMemcpyAsync(L, dL.p, sizeof(double2) * nn * batch, hipMemcpyDeviceToHost, st));
    MIDAS_HIP(hipMemcpyAsync(rank, dR.p, sizeof(int) * batch, hipMemcpyDeviceToHost, st));
    MIDAS_HIP(hipStreamSynchronize(st));
    return 0;
}

int midas_test_trailing(midas_ctx *ctx, int mode, int m, int n, int k, const double *A, long lda, const double *B,
                        long ldb, double *C, long ldc)
{
    MIDAS_CHECK(ctx && A && B && C, -1, "test_trailing: bad arguments");
    MIDAS_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->st;
    DBuf dA, dB, dC, a32, b32, c32, sc;
    MIDAS_CHECK(lda >= m && ldc >= m && ldb >= k, -1, "test_trailing: leading dimensions below the block");
    // A and C keep their leading dimensions on the device (the rows past m
    // travel along, rounded to FP32), so the kernel sees lda, ldc > m
    TRY(upload_mat(dA, A, (int)lda, k, lda, st));
    TRY(upload_mat(dB, B, k, n, ldb, st));
    TRY(upload_mat(dC, C, (int)ldc, n, ldc, st));
    TRY(to_c64(a32, dA, lda * k, st));
    TRY(to_c64(b32, dB, (long)k * n, st));
    TRY(to_c64(c32, dC, ldc * n, st));
    TRY(sc.alloc(trail16_scratch_bytes(m, n, k)));
    TRY(launch_trail16(mode, m, n, k, a32.as<float2>(), lda, b32.as<float2>(), k, c32.as<float2>(), ldc, sc.p, st));
    hipLaunchKernelGGL(k_c2z, dim3((ldc * n + 255) / 256), dim3(256), 0, st, c32.as<float2>(), dC.as<double2>(),
                       ldc * n);
    MIDAS_HIP(hipMemcpyAsync(C, dC.p, sizeof(double2) * ldc * n, hipMemcpyDeviceToHost, st));
    MIDAS_HIP(hipStreamSynchronize(st));
    return 0;
}

int midas_test_mfma_probe(midas_ctx *ctx, int which, double *out)
{
    MIDAS_CHECK(ctx && out && which >= 0 && which <= 5, -1, "mfma_probe: bad arguments");
    MIDAS_HIP(hipSetDevice(ctx->device));
    DBuf d;
    const int cnt = which >= 4 ? 1024 : 256;
    TRY(d.alloc(sizeof(double) * cnt));
    TRY(mfma_probe(which, d.as<double>(), ctx->st));
    MIDAS_HIP(hipMemcpyAsync(out, d.p, sizeof(double) * cnt, hipMemcpyDeviceToHost, ctx->st));
    MIDAS_HIP(hipStreamSynchronize(ctx->st));
    return 0;
}

}  // extern "C"
