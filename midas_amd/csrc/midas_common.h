// Shared device/host helpers for the MIDAS gfx950 library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#define MIDAS_PI_EXACT 3.14159265358979323846
// Quirk bits (the reference's constant/rounding quirks, switchable)
#define MIDAS_Q_PI 1          // Pi = acos(-1.0f) widened to double (Initialization.f90:39)
#define MIDAS_Q_LITERALS 2    // (1./3.), (2./3.) single literals (SetCellsParams.f90:37-44)
#define MIDAS_Q_CEXT_SUM 4    // sum-of-components Cext/Cabs (Compute_ExtAbsCsec_fromIntField.f90:53-60)
#define MIDAS_Q_FP32_BLOCKS 8 // FP32 intermediates in CBFM/ACA fills and RHS (Green_functions.f90:109)

namespace midas {

void set_error(const char *fmt, ...);
const char *last_error();

#define MIDAS_HIP(call)                                                                       \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) {                                                               \
            ::midas::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
            return -100 - (int)e_;                                                            \
        }                                                                                     \
    } while (0)

#define MIDAS_CHECK(cond, code, ...)             \
    do {                                         \
        if (!(cond)) {                           \
            ::midas::set_error(__VA_ARGS__);     \
            return (code);                       \
        }                                        \
    } while (0)

inline double host_pi(int quirks) { return (quirks & MIDAS_Q_PI) ? (double)acosf(-1.0f) : MIDAS_PI_EXACT; }

// ---------------------------------------------------------------------------
// complex arithmetic on plain structs (interleaved re, im)
// ---------------------------------------------------------------------------
template <typename T> struct cplx { T r, i; };
typedef cplx<double> z64;
typedef cplx<float> c32;

template <typename T> __host__ __device__ inline cplx<T> cmk(T r, T i) { return {r, i}; }
template <typename T> __host__ __device__ inline cplx<T> cadd(cplx<T> a, cplx<T> b) { return {a.r + b.r, a.i + b.i}; }
template <typename T> __host__ __device__ inline cplx<T> csub(cplx<T> a, cplx<T> b) { return {a.r - b.r, a.i - b.i}; }
template <typename T> __host__ __device__ inline cplx<T> cmul(cplx<T> a, cplx<T> b)
{
    return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
}
template <typename T> __host__ __device__ inline cplx<T> cscale(cplx<T> a, T s) { return {a.r * s, a.i * s}; }
template <typename T> __host__ __device__ inline T cabs2(cplx<T> a) { return a.r * a.r + a.i * a.i; }
template <typename T> __host__ __device__ inline cplx<T> cinv(cplx<T> a)
{
    T d = (T)1 / (a.r * a.r + a.i * a.i);
    return {a.r * d, -a.i * d};
}
template <typename T> __host__ __device__ inline cplx<T> cconj(cplx<T> a) { return {a.r, -a.i}; }

// Device cell table (structure of arrays) with per-cell EM constants.
struct DevCells {
    int n;
    const double *x, *y, *z;   // centres
    const double *sc;          // cell size
    const double *kappa;       // Kappa_n
    const double *che;         // Che_n = eps - 1, interleaved complex
    const double *znnpp;       // self term, interleaved complex
    const double *kchi;        // Kappa_n * Che_n, interleaved complex
    const int *id;             // cell identifier (n_cell)
};

// Live kernel accounting (HIP events around selected launches), used by the
// benchmark to time the dominant kernel on the stream it runs on.
struct KernelProfile {
    bool on = false;
    struct Rec { hipEvent_t a, b; double flops, bytes; int kind; };
    std::vector<Rec> recs;
};
KernelProfile &profile();
// kinds
enum { PROF_TRAIL = 0, PROF_FILL = 1, PROF_PANEL = 2, PROF_ACA = 3 };
struct ProfScope {
    hipEvent_t a{}, b{};
    hipStream_t st;
    double flops, bytes;
    int kind;
    bool on;
    ProfScope(hipStream_t s, int k, double f, double by) : st(s), flops(f), bytes(by), kind(k), on(profile().on)
    {
        // launches recorded into a graph are not timed (a timing event cannot
        // be recorded on a capturing stream)
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (on && hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) on = false;
        if (on) {
            hipEventCreate(&a);
            hipEventCreate(&b);
            hipEventRecord(a, st);
        }
    }
    ~ProfScope()
    {
        if (on) {
            hipEventRecord(b, st);
            profile().recs.push_back({a, b, flops, bytes, kind});
        }
    }
};

// A set of cells: contiguous range [start, start+n) or an index list.
struct CellSet {
    int n;          // cells in the set
    int start;      // first cell when idx == nullptr
    const int *idx; // optional 0-based cell indices
    __host__ __device__ int cell(int k) const { return idx ? idx[k] : start + k; }
};

// One matrix of a grouped block launch (blockIdx.z = matrix of the batch):
// the extended block's cells are idx[off, off + n); dst is the block's number,
// fsr its SR mask factor, (cx, cy, cz, rs) its spherical-wave source frame.
struct BlockJob {
    int n, off, dst, pad_;
    double fsr, cx, cy, cz, rs;
};

// launchers (kernel files)
int launch_cell_params(int n, const double *sc, const double *eps, double k0, int quirks, double *a, double *che,
                       double *znnpp, double *kappa, double *kchi, hipStream_t st);
int launch_fill(const DevCells &cells, CellSet R, CellSet S, double k0, int quirks, bool ref32, double2 *Z64,
                long ld64, float2 *W32, long ld32, hipStream_t st);
int launch_rhs(const DevCells &cells, CellSet R, int ntr, const double *th, const double *ph, double k0, int quirks,
               bool ref32, double2 *V, long ld, hipStream_t st);
void sph_source_frame(const double *x, const double *y, const double *z, const int *idx, int n, double k0, int quirks,
                      double c[3], double *rs);
int launch_rhs_sph(const DevCells &cells, CellSet R, int ntr, const double *th, const double *ph, double k0,
                   int quirks, bool ref32, const double c[3], double rs, double2 *V, long ld, hipStream_t st);
// Grouped forms for a batch of nq diagonal blocks (jobs and idx in device
// memory; nmax = the largest jobs[q].n): matrix q is the (3n x 3n) block of its
// cell set at Z + q zstride (ld), its excitations the 3n x 2 ntr columns at
// V + q vstride (ld); sph picks the spherical waves of each job's frame.
int launch_fill_group(const DevCells &cells, const BlockJob *jobs, const int *idx, int nq, int nmax, double k0,
                      int quirks, bool ref32, double2 *Z, long ld, long zstride, hipStream_t st);
int launch_rhs_group(const DevCells &cells, const BlockJob *jobs, const int *idx, int nq, int nmax, int ntr,
                     const double *th, const double *ph, double k0, int quirks, bool ref32, bool sph, double2 *V,
                     long ld, long vstride, hipStream_t st);
int near_field(const DevCells &cells, const double2 *E, long ld, int T, int nrx, const double *rth, const double *rph,
               double rso, double k0, int quirks, double2 *S, long lds, hipStream_t st);
int scatq_sums(const double2 *S, long lds, int nrx, int ntr, const double *rth, const double *rph, const double *w,
               const double *tth, const double *tph, int quirks, double *out, hipStream_t st);
int far_field(const DevCells &cells, const double2 *E, long ld, int T, int nrx, const double *rth,
              const double *rph, int npol, const double *beta, double k0, int quirks, double2 *S, long lds,
              double2 *FF, double2 *E3, double2 *G, hipStream_t st);
int cext_cabs(const DevCells &cells, const double2 *E, long ld, int T, const double *th, const double *ph, double k0,
              int quirks, double *cext, double *cabs, hipStream_t st);

}  // namespace midas
