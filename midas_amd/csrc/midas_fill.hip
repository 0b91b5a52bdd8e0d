// Matrix fill, cell constants and plane-wave right-hand sides (gfx950).
//
//  k_cell_params : SetCellsParams.f90:37-48 (per cell a_n, Che_n, Znnpp, Kappa_n)
//  k_fill        : compute_Green_s_tr_single_element (Compute_EFields_MoM.f90:433-523)
//                  for every pair of a receiver set x source set, and
//                  Green_s_tr_partial (Green_functions.f90:96-173) with FP32
//                  intermediates when `ref32` is set.
//  k_rhs         : Incident_Field_single_element (Compute_EFields_MoM.f90:529-598)
//                  and Incident_Field (Incident_Field.f90:1-63, FP32 locals).
//
// One thread per cell pair computes the whole 3x3 block: the pair's distance,
// e^{ikr} and dyad factors are computed once instead of nine times as in the
// reference's per-entry routine.  A wavefront runs along receiver cells, so
// each of the 3 columns a wave touches is one contiguous run of 192 entries;
// the wave stages it in LDS and writes it with three fully coalesced stores
// (64 consecutive elements each).  The source cell's constants are
// wave-uniform.
#include "midas_common.h"
#include "midas_elem.h"

#include <algorithm>
#include <cmath>

namespace midas {

__global__ void k_cell_params(int n, const double *__restrict__ sc, const double *__restrict__ eps, double k0,
                              double pi, double third, double twothird, double *a_out, double *che_out,
                              double *znnpp_out, double *kappa_out, double *kchi_out)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double a = sc[c] * pow(0.75 / pi, third);
    z64 che = {eps[2 * c] - 1.0, eps[2 * c + 1]};
    double s, co;
    sincos(k0 * a, &s, &co);
    // ((2/3) e^{ika} (1 - ika) - 1) * che
    z64 e = {twothird * co, twothird * s};
    z64 t = cmul(e, z64{1.0, -k0 * a});
    t.r -= 1.0;
    z64 zn = cmul(t, che);
    double kap = 4.0 * pi * (s - k0 * a * co) / pow(k0, 3.0);
    a_out[c] = a;
    che_out[2 * c] = che.r;
    che_out[2 * c + 1] = che.i;
    znnpp_out[2 * c] = zn.r;
    znnpp_out[2 * c + 1] = zn.i;
    kappa_out[c] = kap;
    kchi_out[2 * c] = kap * che.r;
    kchi_out[2 * c + 1] = kap * che.i;
}

// Z64 (ld64) and/or W32 (interleaved complex float, ld32) receive the block
// (3*R.n x 3*S.n).  Diagonal cell pairs are detected by identifier equality.
// Tile (bx, by) of a fill: 64 receiver cells x 4 source cells (one per wave).
template <bool REF32>
__device__ __forceinline__ void fill_tile(const DevCells &cells, const CellSet &R, const CellSet &S, double k0,
                                          double pi, double2 *__restrict__ Z64, long ld64, float2 *__restrict__ W32,
                                          long ld32, int bx, int by)
{
    // per wave: one column's 192 entries, staged so that each store
    // instruction writes 64 consecutive elements
    __shared__ double2 st64[4][3 * 64];
    __shared__ float2 st32[4][3 * 64];
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int o = bx * 64 + tx;
    const int sl = by * 4 + ty;
    if (sl >= S.n) return;   // wave-uniform
    const bool valid = o < R.n;
    const int rows = 3 * min(64, R.n - bx * 64);
    const int go = R.cell(valid ? o : bx * 64), gs = S.cell(sl);
    z64 e[9];
    if (cells.id[go] == cells.id[gs]) {
        z64 d = {1.0 - cells.znnpp[2 * gs], -cells.znnpp[2 * gs + 1]};
#pragma unroll
        for (int q = 0; q < 9; ++q) e[q] = {0.0, 0.0};
        e[0] = e[4] = e[8] = d;
    } else {
        double rx = cells.x[go] - cells.x[gs], ry = cells.y[go] - cells.y[gs], rz = cells.z[go] - cells.z[gs];
        if (REF32) {
            dyad_ref32(rx, ry, rz, k0, pi, cells.kappa[gs], z64{cells.che[2 * gs], cells.che[2 * gs + 1]}, e);
        } else {
            Dyad64 d = dyad64(rx, ry, rz, k0, 0.25 / pi, z64{cells.kchi[2 * gs], cells.kchi[2 * gs + 1]});
#pragma unroll
            for (int ps = 0; ps < 3; ++ps)
#pragma unroll
                for (int po = 0; po < 3; ++po) e[3 * ps + po] = dyad_entry(d, po, ps);
        }
    }
#pragma unroll
    for (int ps = 0; ps < 3; ++ps) {
        const long col = 3L * sl + ps;
        if (Z64) {
            if (valid) {
#pragma unroll
                for (int po = 0; po < 3; ++po) st64[ty][3 * tx + po] = make_double2(e[3 * ps + po].r, e[3 * ps + po].i);
            }
            asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
            double2 *dst = Z64 + col * ld64 + 3L * bx * 64;
#pragma unroll
            for (int k = 0; k < 3; ++k)
                if (64 * k + tx < rows) dst[64 * k + tx] = st64[ty][64 * k + tx];
            asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        }
        if (W32) {
            if (valid) {
#pragma unroll
                for (int po = 0; po < 3; ++po)
                    st32[ty][3 * tx + po] = make_float2((float)e[3 * ps + po].r, (float)e[3 * ps + po].i);
            }
            asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
            float2 *dst = W32 + col * ld32 + 3L * bx * 64;
#pragma unroll
            for (int k = 0; k < 3; ++k)
                if (64 * k + tx < rows) dst[64 * k + tx] = st32[ty][64 * k + tx];
            asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        }
    }
}

template <bool REF32>
__global__ __launch_bounds__(256) void k_fill(DevCells cells, CellSet R, CellSet S, double k0, double pi,
                                              double2 *__restrict__ Z64, long ld64, float2 *__restrict__ W32,
                                              long ld32)
{
    fill_tile<REF32>(cells, R, S, k0, pi, Z64, ld64, W32, ld32, blockIdx.x, blockIdx.y);
}

// Grouped diagonal-block fill: matrix blockIdx.z of a batch is the block of
// the cell set jobs[z] (Green_s_tr_partial of an extended block,
// Compute_EFields_CBFME.f90:375-377), one launch for all blocks of the batch.
template <bool REF32>
__global__ __launch_bounds__(256) void k_fill_group(DevCells cells, const BlockJob *__restrict__ jobs,
                                                    const int *__restrict__ idx, double k0, double pi,
                                                    double2 *__restrict__ Z, long ld, long zstride)
{
    const BlockJob j = jobs[blockIdx.z];
    if ((int)blockIdx.x * 64 >= j.n) return;   // workgroup-uniform
    const CellSet S{j.n, 0, idx + j.off};
    fill_tile<REF32>(cells, S, S, k0, pi, Z + blockIdx.z * zstride, ld, nullptr, 0, blockIdx.x, blockIdx.y);
}

// Plane-wave excitation: column t (V-pol) and t+ntr (H-pol) of V (3*R.n x 2*ntr).
// ref32: K11 and the field components are single precision (Incident_Field.f90:17).
template <bool REF32>
__device__ __forceinline__ void rhs_plane_at(const DevCells &cells, const CellSet &R, int c, int t, int ntr,
                                             const double *__restrict__ th_deg, const double *__restrict__ ph_deg,
                                             double k0, double pi, double2 *__restrict__ V, long ld)
{
#pragma clang fp contract(off)
    int g = R.cell(c);
    double th = th_deg[t] * pi / 180.0, ph = ph_deg[t] * pi / 180.0;
    double sth, cth, sph, cph;
    sincos(th, &sth, &cth);
    sincos(ph, &sph, &cph);
    double kx = k0 * cth, ky = k0 * sth * cph, kz = k0 * sth * sph;
    if (REF32) {
        kx = (double)(float)kx;
        ky = (double)(float)ky;
        kz = (double)(float)kz;
    }
    // exp(i kx X) exp(i ky Y) exp(i kz Z) as the reference forms it
    double s1, c1, s2, c2, s3, c3;
    sincos(kx * cells.x[g], &s1, &c1);
    sincos(ky * cells.y[g], &s2, &c2);
    sincos(kz * cells.z[g], &s3, &c3);
    z64 e = cmul(cmul(z64{c1, s1}, z64{c2, s2}), z64{c3, s3});
    z64 v[3] = {cscale(e, -sth), cscale(cscale(e, cth), cph), cscale(cscale(e, cth), sph)};
    z64 h[3] = {{0.0, 0.0}, cscale(e, -sph), cscale(e, cph)};
    double2 *dv = V + (long)t * ld + 3L * c;
    double2 *dh = V + (long)(t + ntr) * ld + 3L * c;
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        if (REF32) {
            dv[p] = make_double2((double)(float)v[p].r, (double)(float)v[p].i);
            dh[p] = make_double2((double)(float)h[p].r, (double)(float)h[p].i);
        } else {
            dv[p] = make_double2(v[p].r, v[p].i);
            dh[p] = make_double2(h[p].r, h[p].i);
        }
    }
}

// Spherical-wave excitation (Incident_Field_Spherical, Incident_Field.f90:65-166):
// the source of direction t sits at (cx, cy, cz) + rs (cos th, sin th cos ph,
// sin th sin ph), rs = 10 max(R_block, 2 pi / k0) around the centroid of the
// extended block (computed by the host).  Field at a cell, with R, u the
// distance and unit vector from the source, theta = acos(ux),
// phi = atan2(uz, uy), phase = e^{i k0 R} / R:
//   V-pol (-phase sin theta, phase cos theta cos phi, phase cos theta sin phi)
//   H-pol (0, -phase sin phi, phase cos phi)
// REF32: phase and every component are complex(4), as the reference declares
// them; the products are formed in double first (mixed-kind promotion).
template <bool REF32>
__device__ __forceinline__ void rhs_sph_at(const DevCells &cells, const CellSet &R, int c, int t, int ntr,
                                           const double *__restrict__ th_deg, const double *__restrict__ ph_deg,
                                           double k0, double pi, double cx, double cy, double cz, double rs,
                                           double2 *__restrict__ V, long ld)
{
#pragma clang fp contract(off)
    const int g = R.cell(c);
    const double th = th_deg[t] * pi / 180.0, ph = ph_deg[t] * pi / 180.0;
    const double xs = cx + rs * cos(th);
    const double ys = cy + rs * sin(th) * cos(ph);
    const double zs = cz + rs * sin(th) * sin(ph);
    const double dx = cells.x[g] - xs, dy = cells.y[g] - ys, dz = cells.z[g] - zs;
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    const double ux = dx / r, uy = dy / r, uz = dz / r;
    double sk, ck;
    sincos(k0 * r, &sk, &ck);
    double pr = ck / r, pim = sk / r;
    if (REF32) {
        pr = (double)(float)pr;
        pim = (double)(float)pim;
    }
    const double theta = acos(ux), phi = atan2(uz, uy);
    const double st = sin(theta), ct = cos(theta), sp = sin(phi), cp = cos(phi);
    z64 v[3] = {{-(pr * st), -(pim * st)}, {pr * ct * cp, pim * ct * cp}, {pr * ct * sp, pim * ct * sp}};
    z64 h[3] = {{0.0, 0.0}, {-(pr * sp), -(pim * sp)}, {pr * cp, pim * cp}};
    double2 *dv = V + (long)t * ld + 3L * c;
    double2 *dh = V + (long)(t + ntr) * ld + 3L * c;
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        if (REF32) {
            dv[p] = make_double2((double)(float)v[p].r, (double)(float)v[p].i);
            dh[p] = make_double2((double)(float)h[p].r, (double)(float)h[p].i);
        } else {
            dv[p] = make_double2(v[p].r, v[p].i);
            dh[p] = make_double2(h[p].r, h[p].i);
        }
    }
}

template <bool REF32>
__global__ void k_rhs(DevCells cells, CellSet R, int ntr, const double *__restrict__ th_deg,
                      const double *__restrict__ ph_deg, double k0, double pi, double2 *__restrict__ V, long ld)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < R.n) rhs_plane_at<REF32>(cells, R, c, blockIdx.y, ntr, th_deg, ph_deg, k0, pi, V, ld);
}

template <bool REF32>
__global__ void k_rhs_sph(DevCells cells, CellSet R, int ntr, const double *__restrict__ th_deg,
                          const double *__restrict__ ph_deg, double k0, double pi, double cx, double cy, double cz,
                          double rs, double2 *__restrict__ V, long ld)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < R.n) rhs_sph_at<REF32>(cells, R, c, blockIdx.y, ntr, th_deg, ph_deg, k0, pi, cx, cy, cz, rs, V, ld);
}

// Grouped excitations of a batch of extended blocks (blockIdx.z = matrix):
// plane waves (Incident_Field) or each block's spherical waves
// (Incident_Field_Spherical), Compute_EFields_CBFME.f90:363-369.
template <bool REF32, bool SPH>
__global__ void k_rhs_group(DevCells cells, const BlockJob *__restrict__ jobs, const int *__restrict__ idx, int ntr,
                            const double *__restrict__ th_deg, const double *__restrict__ ph_deg, double k0, double pi,
                            double2 *__restrict__ V, long ld, long vstride)
{
    const BlockJob j = jobs[blockIdx.z];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= j.n) return;
    const CellSet R{j.n, 0, idx + j.off};
    V += blockIdx.z * vstride;
    if (SPH)
        rhs_sph_at<REF32>(cells, R, c, blockIdx.y, ntr, th_deg, ph_deg, k0, pi, j.cx, j.cy, j.cz, j.rs, V, ld);
    else
        rhs_plane_at<REF32>(cells, R, c, blockIdx.y, ntr, th_deg, ph_deg, k0, pi, V, ld);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
int launch_cell_params(int n, const double *sc, const double *eps, double k0, int quirks, double *a, double *che,
                       double *znnpp, double *kappa, double *kchi, hipStream_t st)
{
    double pi = host_pi(quirks);
    double third = (quirks & MIDAS_Q_LITERALS) ? (double)(1.0f / 3.0f) : 1.0 / 3.0;
    double twothird = (quirks & MIDAS_Q_LITERALS) ? (double)(2.0f / 3.0f) : 2.0 / 3.0;
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_cell_params, dim3((n + 255) / 256), dim3(256), 0, st, n, sc, eps, k0, pi, third, twothird,
                       a, che, znnpp, kappa, kchi);
    MIDAS_HIP(hipGetLastError());
    return 0;
}

int launch_fill(const DevCells &cells, CellSet R, CellSet S, double k0, int quirks, bool ref32, double2 *Z64,
                long ld64, float2 *W32, long ld32, hipStream_t st)
{
    if (R.n <= 0 || S.n <= 0) return 0;
    MIDAS_CHECK(!Z64 || ld64 >= 3L * R.n, -1, "fill: ld64 %ld < 3*%d", ld64, R.n);
    MIDAS_CHECK(!W32 || ld32 >= 3L * R.n, -1, "fill: ld32 %ld < 3*%d", ld32, R.n);
    MIDAS_CHECK((S.n + 3) / 4 <= 65535 * 64, -1, "fill: source set too large");
    double pi = host_pi(quirks);
    const double elems = 9.0 * R.n * (double)S.n;
    ProfScope ps(st, PROF_FILL, 125.0 * R.n * (double)S.n, elems * ((Z64 ? 16.0 : 0.0) + (W32 ? 8.0 : 0.0)));
    dim3 blk(64, 4);
    // grid.y is limited to 65535: split the source set in chunks
    const int ymax = 65535;
    for (int s0 = 0; s0 < S.n; s0 += ymax * 4) {
        CellSet Sc = S;
        int cnt = S.n - s0 < ymax * 4 ? S.n - s0 : ymax * 4;
        Sc.n = cnt;
        if (S.idx) Sc.idx = S.idx + s0; else Sc.start = S.start + s0;
        dim3 grd((R.n + 63) / 64, (cnt + 3) / 4);
        double2 *z = Z64 ? Z64 + 3L * s0 * ld64 : nullptr;
        float2 *w = W32 ? W32 + 3L * s0 * ld32 : nullptr;
        if (ref32)
            hipLaunchKernelGGL(k_fill<true>, grd, blk, 0, st, cells, R, Sc, k0, pi, z, ld64, w, ld32);
        else
            hipLaunchKernelGGL(k_fill<false>, grd, blk, 0, st, cells, R, Sc, k0, pi, z, ld64, w, ld32);
        MIDAS_HIP(hipGetLastError());
    }
    return 0;
}

int launch_rhs(const DevCells &cells, CellSet R, int ntr, const double *th, const double *ph, double k0, int quirks,
               bool ref32, double2 *V, long ld, hipStream_t st)
{
    if (R.n <= 0 || ntr <= 0) return 0;
    MIDAS_CHECK(ntr <= 65535, -1, "rhs: too many directions in one launch");
    double pi = host_pi(quirks);
    dim3 grd((R.n + 255) / 256, ntr);
    if (ref32)
        hipLaunchKernelGGL(k_rhs<true>, grd, dim3(256), 0, st, cells, R, ntr, th, ph, k0, pi, V, ld);
    else
        hipLaunchKernelGGL(k_rhs<false>, grd, dim3(256), 0, st, cells, R, ntr, th, ph, k0, pi, V, ld);
    MIDAS_HIP(hipGetLastError());
    return 0;
}

// Centre and source radius of Incident_Field_Spherical for the cells idx[0..n)
// (an extended block, block cells first): centroid as sequential sums / n,
// R_block = the largest distance to it, rs = 10 max(R_block, 2 pi / k0).
void sph_source_frame(const double *x, const double *y, const double *z, const int *idx, int n, double k0, int quirks,
                      double c[3], double *rs)
{
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int i = 0; i < n; ++i) {
        const int g = idx ? idx[i] : i;
        sx += x[g];
        sy += y[g];
        sz += z[g];
    }
    c[0] = sx / n;
    c[1] = sy / n;
    c[2] = sz / n;
    double m = 0.0;
    for (int i = 0; i < n; ++i) {
        const int g = idx ? idx[i] : i;
        const double d = (x[g] - c[0]) * (x[g] - c[0]) + (y[g] - c[1]) * (y[g] - c[1]) + (z[g] - c[2]) * (z[g] - c[2]);
        m = d > m ? d : m;
    }
    const double pi = host_pi(quirks);
    *rs = 10.0 * std::max(std::sqrt(m), 2.0 * pi / k0);
}

int launch_rhs_sph(const DevCells &cells, CellSet R, int ntr, const double *th, const double *ph, double k0,
                   int quirks, bool ref32, const double c[3], double rs, double2 *V, long ld, hipStream_t st)
{
    if (R.n <= 0 || ntr <= 0) return 0;
    MIDAS_CHECK(ntr <= 65535, -1, "rhs_sph: too many directions in one launch");
    double pi = host_pi(quirks);
    dim3 grd((R.n + 255) / 256, ntr);
    if (ref32)
        hipLaunchKernelGGL(k_rhs_sph<true>, grd, dim3(256), 0, st, cells, R, ntr, th, ph, k0, pi, c[0], c[1], c[2], rs,
                           V, ld);
    else
        hipLaunchKernelGGL(k_rhs_sph<false>, grd, dim3(256), 0, st, cells, R, ntr, th, ph, k0, pi, c[0], c[1], c[2],
                           rs, V, ld);
    MIDAS_HIP(hipGetLastError());
    return 0;
}

int launch_fill_group(const DevCells &cells, const BlockJob *jobs, const int *idx, int nq, int nmax, double k0,
                      int quirks, bool ref32, double2 *Z, long ld, long zstride, hipStream_t st)
{
    if (nq <= 0 || nmax <= 0) return 0;
    MIDAS_CHECK(ld >= 3L * nmax && zstride >= ld * 3L * nmax && nq <= 65535 && (nmax + 3) / 4 <= 65535, -1,
                "fill_group: %d blocks of up to %d cells do not fit ld %ld, stride %ld", nq, nmax, ld, zstride);
    const double pi = host_pi(quirks);
    const dim3 grd((nmax + 63) / 64, (nmax + 3) / 4, nq), blk(64, 4);
    if (ref32)
        hipLaunchKernelGGL(k_fill_group<true>, grd, blk, 0, st, cells, jobs, idx, k0, pi, Z, ld, zstride);
    else
        hipLaunchKernelGGL(k_fill_group<false>, grd, blk, 0, st, cells, jobs, idx, k0, pi, Z, ld, zstride);
    MIDAS_HIP(hipGetLastError());
    return 0;
}

int launch_rhs_group(const DevCells &cells, const BlockJob *jobs, const int *idx, int nq, int nmax, int ntr,
                     const double *th, const double *ph, double k0, int quirks, bool ref32, bool sph, double2 *V,
                     long ld, long vstride, hipStream_t st)
{
    if (nq <= 0 || nmax <= 0 || ntr <= 0) return 0;
    MIDAS_CHECK(ntr <= 65535 && nq <= 65535 && ld >= 3L * nmax && vstride >= ld * 2L * ntr, -1,
                "rhs_group: %d blocks of up to %d cells, %d directions do not fit ld %ld, stride %ld", nq, nmax, ntr,
                ld, vstride);
    const double pi = host_pi(quirks);
    const dim3 grd((nmax + 255) / 256, ntr, nq);
#define RHS_GROUP(R32, SP)                                                                                        \
    hipLaunchKernelGGL((k_rhs_group<R32, SP>), grd, dim3(256), 0, st, cells, jobs, idx, ntr, th, ph, k0, pi, V, ld, \
                       vstride)
    if (ref32) {
        if (sph) RHS_GROUP(true, true); else RHS_GROUP(true, false);
    } else {
        if (sph) RHS_GROUP(false, true); else RHS_GROUP(false, false);
    }
#undef RHS_GROUP
    MIDAS_HIP(hipGetLastError());
    return 0;
}

}  // namespace midas
