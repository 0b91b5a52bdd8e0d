// Host-side plumbing shared by the C-ABI translation units: device buffers,
// the context, the device cell table and the dense solver entry.
#pragma once
#include "midas_common.h"
#include "midas_lu.h"
#include "../../include/midas_gpu.h"

#include <algorithm>
#include <map>
#include <string>
#include <vector>

namespace midas {

// ---------------------------------------------------------------------------
// small device buffer owner
// ---------------------------------------------------------------------------
struct DBuf {
    void *p = nullptr;
    size_t bytes = 0;
    bool borrowed = false;   // memory lent by its owner (never allocated or freed here)
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    int alloc(size_t b)
    {
        MIDAS_CHECK(!borrowed, -1, "internal: allocation of a borrowed buffer (%zu bytes)", b);
        release();
        if (b == 0) b = 16;
        hipError_t e = hipMalloc(&p, b);
        if (e != hipSuccess) {
            p = nullptr;
            set_error("hipMalloc(%zu bytes): %s", b, hipGetErrorString(e));
            return -100 - (int)e;
        }
        bytes = b;
        return 0;
    }
    // grow-only: keeps the allocation when it is large enough (contents are
    // not preserved when it grows; the slack is a quarter, at most 256 MiB);
    // hipFree waits for work still using it.  A borrowed buffer never grows:
    // asking for more than it holds is an error.
    int reserve(size_t b)
    {
        if (p && b <= bytes) return 0;
        MIDAS_CHECK(!borrowed, -1, "internal: borrowed buffer of %zu bytes asked for %zu", bytes, b);
        return alloc(b + std::min(b / 4, (size_t)256 << 20));
    }
    // lend `cap` bytes at q (not owned: release forgets them)
    void borrow(void *q, size_t cap)
    {
        release();
        p = q;
        bytes = cap;
        borrowed = true;
    }
    void release()
    {
        if (p && !borrowed) hipFree(p);
        p = nullptr;
        bytes = 0;
        borrowed = false;
    }
    template <typename X> X *as() const { return (X *)p; }
};

#define TRY(x)                  \
    do {                        \
        int r_ = (x);           \
        if (r_) return r_;      \
    } while (0)

// blocking copy ordered after the work already queued on st
inline int copy_sync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, hipStream_t st)
{
    MIDAS_HIP(hipMemcpyAsync(dst, src, bytes, kind, st));
    MIDAS_HIP(hipStreamSynchronize(st));
    return 0;
}

inline int upload(DBuf &b, const void *h, size_t bytes, hipStream_t st)
{
    TRY(b.alloc(bytes));
    if (h && bytes) MIDAS_HIP(hipMemcpyAsync(b.p, h, bytes, hipMemcpyHostToDevice, st));
    return 0;
}
// the same into a grow-only buffer (one kept between calls)
inline int upload_keep(DBuf &b, const void *h, size_t bytes, hipStream_t st)
{
    TRY(b.reserve(bytes));
    if (h && bytes) MIDAS_HIP(hipMemcpyAsync(b.p, h, bytes, hipMemcpyHostToDevice, st));
    return 0;
}

int launch_iota1(int *id, int n, hipStream_t st);   // id[i] = i + 1 (midas_resid.hip)

}  // namespace midas

using namespace midas;

// Device-resident cell table with derived constants.
struct DevCellsOwned {
    DBuf x, y, z, sc, eps, kappa, che, znnpp, kchi, id, a;
    DevCells v{};
    // host cells -> device; constants computed on the GPU unless supplied
    // (grow-only buffers: a DevCellsOwned kept by the context serves repeated calls)
    int from_host(const midas_cells_t *c, double k0, int quirks, hipStream_t st)
    {
        MIDAS_CHECK(c && c->n > 0 && c->x && c->y && c->z, -1, "cells: empty or missing coordinates");
        const int n = c->n;
        const size_t d = sizeof(double) * n, z2 = 2 * d;
        TRY(upload_keep(x, c->x, d, st));
        TRY(upload_keep(y, c->y, d, st));
        TRY(upload_keep(this->z, c->z, d, st));
        TRY(upload_keep(sc, c->sc, d, st));
        std::vector<int> ids;
        if (!c->id) {
            ids.resize(n);
            for (int i = 0; i < n; ++i) ids[i] = i + 1;
        }
        TRY(upload_keep(id, c->id ? (const void *)c->id : (const void *)ids.data(), sizeof(int) * n, st));
        TRY(kappa.reserve(d));
        TRY(che.reserve(z2));
        TRY(znnpp.reserve(z2));
        TRY(kchi.reserve(z2));
        TRY(a.reserve(d));
        if (c->kappa && c->che && c->znnpp) {
            MIDAS_HIP(hipMemcpyAsync(kappa.p, c->kappa, d, hipMemcpyHostToDevice, st));
            MIDAS_HIP(hipMemcpyAsync(che.p, c->che, z2, hipMemcpyHostToDevice, st));
            MIDAS_HIP(hipMemcpyAsync(znnpp.p, c->znnpp, z2, hipMemcpyHostToDevice, st));
            std::vector<double> kc(2 * n);
            for (int i = 0; i < n; ++i) {
                kc[2 * i] = c->kappa[i] * c->che[2 * i];
                kc[2 * i + 1] = c->kappa[i] * c->che[2 * i + 1];
            }
            MIDAS_HIP(hipMemcpyAsync(kchi.p, kc.data(), z2, hipMemcpyHostToDevice, st));
            MIDAS_HIP(hipStreamSynchronize(st));
        } else {
            MIDAS_CHECK(c->sc && c->eps, -1, "cells: need sc and eps (or kappa, che, znnpp)");
            TRY(upload_keep(eps, c->eps, z2, st));
            TRY(launch_cell_params(n, sc.as<double>(), eps.as<double>(), k0, quirks, a.as<double>(),
                                   che.as<double>(), znnpp.as<double>(), kappa.as<double>(), kchi.as<double>(), st));
        }
        v = DevCells{n, x.as<double>(), y.as<double>(), this->z.as<double>(), sc.as<double>(), kappa.as<double>(),
                     che.as<double>(), znnpp.as<double>(), kchi.as<double>(), id.as<int>()};
        return 0;
    }
    // device pointers supplied by the caller; constants computed here
    int from_device(const midas_cells_t *c, double k0, int quirks, hipStream_t st)
    {
        MIDAS_CHECK(c && c->n > 0 && c->x && c->y && c->z && c->sc && c->eps, -1, "dev cells: missing arrays");
        const int n = c->n;
        const size_t d = sizeof(double) * n, z2 = 2 * d;
        // grow-only: a DevCellsOwned kept by the context serves repeated calls
        TRY(kappa.reserve(d));
        TRY(che.reserve(z2));
        TRY(znnpp.reserve(z2));
        TRY(kchi.reserve(z2));
        TRY(a.reserve(d));
        const int *ids = c->id;
        if (!ids) {
            TRY(id.reserve(sizeof(int) * n));
            TRY(launch_iota1(id.as<int>(), n, st));
            ids = id.as<int>();
        }
        TRY(launch_cell_params(n, c->sc, c->eps, k0, quirks, a.as<double>(), che.as<double>(), znnpp.as<double>(),
                               kappa.as<double>(), kchi.as<double>(), st));
        v = DevCells{n, c->x, c->y, c->z, c->sc, kappa.as<double>(), che.as<double>(), znnpp.as<double>(),
                     kchi.as<double>(), ids};
        return 0;
    }
};

// Device workspace of a context, grow-only, reused by every solve on it.
struct SolveWs {
    DBuf W;            // complex-FP32 copy of the matrix (mixed modes)
    DBuf trail;        // split-precision operand planes of the trailing update
    DBuf trailp;       // the same for the panel stream's update (two-panel steps)
    DBuf trail2;       // the main stream's U12 and second-panel products (two-panel steps)
    DBuf ipiv, perm;   // interchanges and the final permutation
    DBuf norms;        // row sums, column maxima, scalars
    DBuf zx;           // one-pass residual / row-sum scratch
    DBuf R, D, t2;     // refinement: FP64 residual, FP32 correction, substitution scratch
    DBuf xbest;        // refinement: iterate with the smallest backward error
    DBuf zcopy;        // FP64 copy of Z when Z must be kept
    DBuf V;            // excitation of the MoM solve
    DBuf l21;          // diagonal-block pivot search: L21 staging (n x nb) and max |L21|
    int pivot_search = 0;   // the next mixed solve restricts its pivot search to diagonal blocks (MoM)
    // matrix-free FP64 residuals (MoM mixed solves, <= 4 right-hand sides): the
    // cells the matrix is recomputed from; the FP64 matrix is then filled only
    // for the FP64 fallback
    const DevCells *mf_cells = nullptr;
    double mf_k0 = 0.0, mf_pi = 0.0;
    int mf_quirks = 0;
    bool mf_active = false;   // set by the solve while it uses them
    DBuf mfs;                 // their per-chunk partial sums
    LuWork<float> luf;
    LuWork<double> lud;
    bool lent = false; // every buffer lent from a fixed scratch: nothing may be allocated
    void release()
    {
        DBuf *b[] = {&W, &trail, &trailp, &trail2, &ipiv, &perm, &norms, &zx, &R, &D, &t2, &xbest, &zcopy, &V, &l21, &mfs};
        for (DBuf *x : b) x->release();
        luf.release();
        lud.release();
    }
};

// Buffers of the CBF block batches' mixed-precision solves
struct BlockMixedWs {
    DBuf W, E, R, D, tmp, scr, rs, anorm, cm, l21;
    LuWork<float> lu;
    void release()
    {
        for (DBuf *b : {&W, &E, &R, &D, &tmp, &scr, &rs, &anorm, &cm, &l21}) b->release();
        lu.release();
    }
};

// The ACA pipeline's buffers of its two batch parities, and the pinned host
// words that bring a batch's ranks back
struct AcaBufs {
    DBuf dj[2], Ub[2], Vb[2], dk[2], Gb[2], LR[2], dgj[2];
    int *hk[2] = {};
    size_t hk_n = 0;
    int reserve_ranks(size_t n)
    {
        if (n <= hk_n) return 0;
        release_ranks();
        for (int q = 0; q < 2; ++q) MIDAS_HIP(hipHostMalloc((void **)&hk[q], sizeof(int) * n));
        hk_n = n;
        return 0;
    }
    void release_ranks()
    {
        for (int *&h : hk)
            if (h) {
                hipHostFree(h);
                h = nullptr;
            }
        hk_n = 0;
    }
    void release()
    {
        for (int q = 0; q < 2; ++q)
            for (DBuf *b : {&dj[q], &Ub[q], &Vb[q], &dk[q], &Gb[q], &LR[q], &dgj[q]}) b->release();
        release_ranks();
    }
};

struct midas_ctx {
    int device = 0;
    hipStream_t st = nullptr;
    // the factorisations' panel stream and the CBFM pipeline's second stream,
    // with their events, made once by midas_init: streams made call after call
    // can land on the main stream's hardware queue and serialise with it
    hipStream_t ps = nullptr, s2 = nullptr;
    hipEvent_t pev[2] = {}, aev[5] = {};
    SolveWs ws;
    DevCellsOwned cells;          // cell constants of the device-pointer entry points
    // the allocation-free entry point (midas_dev_mom_solve_noalloc): its
    // workspace, lent from the device's static scratch array, and the events
    // its factorisation uses, made once by midas_init
    void *noalloc_base = nullptr;
    SolveWs nws;
    DevCellsOwned ncells;
    hipEvent_t nev[2] = {};
    double cbfm_budget[4] = {};   // bytes; 0 = default (midas_set_cbfm_budgets)
    // executable graphs of the CBFM block batches (one per batch position),
    // updated in place by the next call whose batch has the same launches
    std::vector<hipGraphExec_t> cbfm_graphs;
    // the device buffers of the MoM and CBFM calls (host-pointer entry points),
    // kept between calls and grown when a call needs more (hipMalloc and
    // hipFree wait for the device: a call that made and freed its buffers left
    // the GPU idle around each): named buffers, the cells, the CBF block
    // batches' LU workspaces and the ACA pipeline's buffers
    // (midas_release_workspace frees them)
    std::map<std::string, DBuf> call_pool;
    DevCellsOwned call_cells;
    BlockMixedWs cbfm_mixed;
    LuWork<double> cbfm_lu;
    AcaBufs cbfm_aca;
    void release_calls()
    {
        call_pool.clear();
        for (DBuf *b : {&call_cells.x, &call_cells.y, &call_cells.z, &call_cells.sc, &call_cells.eps,
                        &call_cells.kappa, &call_cells.che, &call_cells.znnpp, &call_cells.kchi, &call_cells.id,
                        &call_cells.a})
            b->release();
        cbfm_mixed.release();
        cbfm_lu.release();
        cbfm_aca.release();
    }
};

// a named call buffer of the context (grow it with reserve)
inline DBuf &pooled(midas_ctx *ctx, const char *name) { return ctx->call_pool[name]; }

// Event timer on a stream; a disabled timer creates no events and reads 0.
struct Timer {
    hipEvent_t a{}, b{};
    hipStream_t st;
    bool on;
    explicit Timer(hipStream_t s, bool enabled = true) : st(s), on(enabled)
    {
        if (!on) return;
        hipEventCreate(&a);
        hipEventCreate(&b);
        hipEventRecord(a, st);
    }
    double ms()
    {
        if (!on) return 0.0;
        hipEventRecord(b, st);
        hipEventSynchronize(b);
        float t = 0;
        hipEventElapsedTime(&t, a, b);
        hipEventRecord(a, st);
        return t;
    }
    ~Timer()
    {
        if (!on) return;
        hipEventDestroy(a);
        hipEventDestroy(b);
    }
};

// One-pass kernels over the FP64 matrix (midas_resid.hip)
namespace midas {
size_t zx_scratch_bytes(int n, int nrhs);
int launch_residual_few(const double2 *Z, long ldz, int n, const double2 *X, long ldx, int nrhs, const double2 *B,
                        long ldb, double2 *R, long ldr, void *scratch, hipStream_t st);
int launch_rowabs(const double2 *Z, long ldz, int n, double *rowsum, void *scratch, hipStream_t st);
// the same two from the cells (the MoM matrix recomputed, never stored)
size_t mf_scratch_bytes(int ncells, int nrhs);
int launch_residual_free(const DevCells &cells, int nc, double k0, double pi, const double2 *X, long ldx, int nrhs,
                         const double2 *B, long ldb, double2 *R, long ldr, void *scratch, hipStream_t st);
int launch_rowabs_free(const DevCells &cells, int nc, double k0, double pi, double *rowsum, void *scratch,
                       hipStream_t st);
}  // namespace midas

// Dense solve on the device (FP64 LU or mixed-precision LU + FP64 refinement):
// Z (n x n, ldz) FP64; W32 optional FP32 copy of Z for mixed modes; X = Z^-1 B.
// ws: the workspace (ctx->ws when NULL); a lent workspace (ws->lent) makes
// the solve allocate nothing: no timers, hand-written FP64 GEMMs only.
int solve_dev(midas_ctx *ctx, int n, int nrhs, double2 *Z, long ldz, float2 *W32, long ldw, const double2 *B,
              long ldb, double2 *X, long ldx, int mode, const midas_solve_opts_t *opts, midas_solve_info_t *info,
              bool z_inplace, SolveWs *ws = nullptr);
