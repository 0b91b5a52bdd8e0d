// Far-field scattering matrix and extinction / absorption cross sections.
//
//  ff_n(r)   = e^{-ik u_r . p_n} k^3 kappa_n chi_n / (4 pi)      GetFFieldCoeff (Green_functions.f90:1399-1438)
//  E_{v,h}(r)= sum_n ff_n(r) E_n                                  Compute_EFields_MoM.f90:314-330
//  S         = polarisation projection with beta rotation         Compute_Scattering_Matrices.f90:148-207
//  Cext/Cabs = sums over cells of the internal field              Compute_ExtAbsCsec_fromIntField.f90:45-64
//
// The per-receiver loops of the reference become one complex GEMM
// FF(nrx x N) * E3(N x 6T) (E3 holds the x, y, z components of all 2T
// solutions side by side), followed by an elementwise projection kernel.
#include "midas_common.h"
#include "midas_host.h"
#include "midas_lu.h"

#include <algorithm>

namespace midas {

__global__ void k_ffcoef(DevCells cells, int nrx, const double *rth, const double *rph, double k0, double pi,
                         double2 *FF, long ld)
{
    int r = blockIdx.x * 64 + threadIdx.x;
    int c = blockIdx.y * 4 + threadIdx.y;
    if (r >= nrx || c >= cells.n) return;
    double st, ct, sp, cp;
    sincos(rth[r] * pi / 180.0, &st, &ct);
    sincos(rph[r] * pi / 180.0, &sp, &cp);
    double ux = ct, uy = st * cp, uz = st * sp;
    double ph = -k0 * (ux * cells.x[c] + uy * cells.y[c] + uz * cells.z[c]);
    double s, co;
    sincos(ph, &s, &co);
    z64 f = {cells.kchi[2 * c], cells.kchi[2 * c + 1]};
    z64 v = cscale(cmul(cscale(z64{co, s}, k0 * k0), f), 1.0 / (4.0 * pi));
    v = cscale(v, k0);
    FF[(long)r + (long)c * ld] = make_double2(v.r, v.i);
}

// E (3N x 2T, ld) -> E3 (N x 6T): column p*2T + t holds component p of solution t.
__global__ void k_split_components(const double2 *E, long ld, int N, int twoT, double2 *E3)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    int t = blockIdx.y;
    if (c >= N) return;
#pragma unroll
    for (int p = 0; p < 3; ++p) E3[(long)c + (long)(p * twoT + t) * N] = E[3L * c + p + (long)t * ld];
}

// G (nrx x 6T): projection to S (nrx x 4*T*npol), columns 4*npol*t + 4*ip + {Vv,Vh,Hv,Hh}.
__global__ void k_project(const double2 *G, int nrx, int T, const double *rth, const double *rph, int npol,
                          const double *beta, double pi, double2 *S, long lds)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    int t = blockIdx.y;
    if (r >= nrx) return;
    const int twoT = 2 * T;
    z64 ev[3], eh[3];
    for (int p = 0; p < 3; ++p) {
        double2 a = G[(long)r + (long)(p * twoT + t) * nrx];
        double2 b = G[(long)r + (long)(p * twoT + T + t) * nrx];
        ev[p] = {a.x, a.y};
        eh[p] = {b.x, b.y};
    }
    double th = rth[r] * pi / 180.0, ph = rph[r] * pi / 180.0;
    double st = sin(th), ct = cos(th), sp = sin(ph), cp = cos(ph);
    for (int ip = 0; ip < npol; ++ip) {
        double b = beta[ip] * pi / 180.0, cb = cos(b), sb = sin(b);
        z64 evp[3], ehp[3];
        for (int p = 0; p < 3; ++p) {
            evp[p] = cadd(cscale(ev[p], cb), cscale(eh[p], sb));
            ehp[p] = cadd(cscale(ev[p], -sb), cscale(eh[p], cb));
        }
        // theta-hat and phi-hat rotated by beta
        double tv0 = -cb * st, tv1 = cb * ct * cp - sb * sp, tv2 = cb * ct * sp + sb * cp;
        double th0 = sb * st, th1 = -(sb * ct * cp + cb * sp), th2 = cb * cp - sb * ct * sp;
        z64 vv = cadd(cadd(cscale(evp[0], tv0), cscale(evp[1], tv1)), cscale(evp[2], tv2));
        z64 vh = cadd(cadd(cscale(evp[0], th0), cscale(evp[1], th1)), cscale(evp[2], th2));
        z64 hv = cadd(cadd(cscale(ehp[0], tv0), cscale(ehp[1], tv1)), cscale(ehp[2], tv2));
        z64 hh = cadd(cadd(cscale(ehp[0], th0), cscale(ehp[1], th1)), cscale(ehp[2], th2));
        long col = 4L * npol * t + 4L * ip;
        S[(long)r + (col + 0) * lds] = make_double2(vv.r, vv.i);
        S[(long)r + (col + 1) * lds] = make_double2(vh.r, vh.i);
        S[(long)r + (col + 2) * lds] = make_double2(hv.r, hv.i);
        S[(long)r + (col + 3) * lds] = make_double2(hh.r, hh.i);
    }
}

// One workgroup per incidence direction t: C_ext(t), C_abs(t).
// Einc is the plane wave of Incident_Field (FP32 locals when ref32).
template <bool REF32, bool QSUM>
__global__ __launch_bounds__(256) void k_cext_cabs(DevCells cells, const double2 *E, long ld, int T,
                                                   const double *th_deg, const double *ph_deg, double k0, double pi,
                                                   double *cext, double *cabs)
{
#pragma clang fp contract(off)
    const int t = blockIdx.x;
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
    double ext = 0.0, ab = 0.0;
    for (int c = threadIdx.x; c < cells.n; c += blockDim.x) {
        double s1, c1, s2, c2, s3, c3;
        sincos(kx * cells.x[c], &s1, &c1);
        sincos(ky * cells.y[c], &s2, &c2);
        sincos(kz * cells.z[c], &s3, &c3);
        z64 e = cmul(cmul(z64{c1, s1}, z64{c2, s2}), z64{c3, s3});
        z64 inc[2][3] = {{cscale(e, -sth), cscale(cscale(e, cth), cph), cscale(cscale(e, cth), sph)},
                         {{0.0, 0.0}, cscale(e, -sph), cscale(e, cph)}};
        if (REF32)
            for (int a = 0; a < 2; ++a)
                for (int p = 0; p < 3; ++p) inc[a][p] = {(double)(float)inc[a][p].r, (double)(float)inc[a][p].i};
        z64 chi = {cells.che[2 * c], cells.che[2 * c + 1]};
        double v = cells.sc[c] * cells.sc[c] * cells.sc[c];
        for (int pol = 0; pol < 2; ++pol) {
            const double2 *ec = E + (long)(t + pol * T) * ld + 3L * c;
            z64 ep[3] = {{ec[0].x, ec[0].y}, {ec[1].x, ec[1].y}, {ec[2].x, ec[2].y}};
            if (QSUM) {
                z64 se = cadd(cadd(ep[0], ep[1]), ep[2]);
                z64 si = cadd(cadd(inc[pol][0], inc[pol][1]), inc[pol][2]);
                ab += chi.i * cabs2(se) * v;
                ext += cmul(cmul(chi, se), cconj(si)).i * v;
            } else {
                double ee = cabs2(ep[0]) + cabs2(ep[1]) + cabs2(ep[2]);
                z64 ei = cadd(cadd(cmul(ep[0], cconj(inc[pol][0])), cmul(ep[1], cconj(inc[pol][1]))),
                              cmul(ep[2], cconj(inc[pol][2])));
                ab += chi.i * ee * v;
                ext += cmul(chi, ei).i * v;
            }
        }
    }
    __shared__ double re[256], ra[256];
    re[threadIdx.x] = ext;
    ra[threadIdx.x] = ab;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            re[threadIdx.x] += re[threadIdx.x + s];
            ra[threadIdx.x] += ra[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        cext[2 * t] = k0 * re[0] / 2.0;
        cext[2 * t + 1] = 0.0;
        cabs[2 * t] = k0 * ra[0] / 2.0;
        cabs[2 * t + 1] = 0.0;
    }
}

// S (nrx x 4*T*npol, lds) from E (3N x 2T).  Workspace: FF (nrx x N), E3 (N x 6T), G (nrx x 6T).
// Near-field dyad (Green_s_dt, Green_functions.f90:1335-1397), FP64: row
// 3r+a, column 3n+q of G (3 nrx x 3N, ld) is +F_aq(r, n) kappa_n chi_n for a
// receiver at rso (cos th, sin th cos ph, sin th sin ph), with
// Gr = e^{ik r} / (4 Pi r^2), Tau = ik - 1/r,
//   F_aa = Gr (Tau + k^2 (r - r_a^2 / r) - 3 r_a^2 / r^2 Tau)
//   F_ab = Gr / r (r_a r_b) (-k^2 - 3 Tau / r).
__global__ void k_green_dt(DevCells cells, int nrx, const double *__restrict__ rth, const double *__restrict__ rph,
                           double rso, double k0, double pi, double2 *__restrict__ G, long ld)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (n >= cells.n) return;
    const double t = rth[r] * pi / 180.0, p = rph[r] * pi / 180.0;
    const double ra[3] = {rso * cos(t) - cells.x[n], rso * sin(t) * cos(p) - cells.y[n],
                          rso * sin(t) * sin(p) - cells.z[n]};
    const double rr = sqrt(ra[0] * ra[0] + ra[1] * ra[1] + ra[2] * ra[2]);
    double s, c;
    sincos(k0 * rr, &s, &c);
    const double den = 4.0 * pi * (rr * rr);
    const z64 gr = {c / den, s / den};
    const z64 tau = {-1.0 / rr, k0};
    const z64 kchi = {cells.kchi[2 * n], cells.kchi[2 * n + 1]};
    const z64 grr = {gr.r / rr, gr.i / rr};
    const z64 offc = {-k0 * k0 - 3.0 * tau.r / rr, -3.0 * tau.i / rr};
    z64 F[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double q = 3.0 * (ra[a] * ra[a]) / (rr * rr);
        const z64 b = {tau.r + k0 * k0 * (rr - ra[a] * ra[a] / rr) - q * tau.r, tau.i - q * tau.i};
        F[a][a] = cmul(gr, b);
    }
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int q = a + 1; q < 3; ++q) F[a][q] = F[q][a] = cmul(cscale(grr, ra[a] * ra[q]), offc);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        double2 *col = G + (long)(3 * n + q) * ld + 3L * r;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const z64 v = cmul(F[a][q], kchi);
            col[a] = make_double2(v.r, v.i);
        }
    }
}

// Es (nrx x 4T): the field Es3 (3 nrx x 2T, rows 3r+a) projected on the
// receiver's theta / phi directions (Compute_Scattered_Fields.f90:100-128).
__global__ void k_project_nf(const double2 *__restrict__ Es3, long ld3, int nrx, int T, int r0,
                             const double *__restrict__ rth, const double *__restrict__ rph, double pi,
                             double2 *__restrict__ S, long lds)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int t = blockIdx.y;
    if (r >= nrx) return;
    const double th = rth[r0 + r] * pi / 180.0, ph = rph[r0 + r] * pi / 180.0;
    const double st = sin(th), ct = cos(th), sp = sin(ph), cp = cos(ph);
    const double2 *ev = Es3 + (long)t * ld3 + 3L * r, *eh = Es3 + (long)(t + T) * ld3 + 3L * r;
    auto theta_part = [&](const double2 *e) {
        return z64{-e[0].x * st + e[1].x * ct * cp + e[2].x * ct * sp, -e[0].y * st + e[1].y * ct * cp + e[2].y * ct * sp};
    };
    auto phi_part = [&](const double2 *e) { return z64{-e[1].x * sp + e[2].x * cp, -e[1].y * sp + e[2].y * cp}; };
    const z64 vv = theta_part(ev), vh = phi_part(ev), hv = theta_part(eh), hh = phi_part(eh);
    double2 *o = S + (long)(r0 + r);
    o[(long)(4 * t) * lds] = make_double2(vv.r, vv.i);
    o[(long)(4 * t + 1) * lds] = make_double2(vh.r, vh.i);
    o[(long)(4 * t + 2) * lds] = make_double2(hv.r, hv.i);
    o[(long)(4 * t + 3) * lds] = make_double2(hh.r, hh.i);
}

int near_field(const DevCells &cells, const double2 *E, long ld, int T, int nrx, const double *rth, const double *rph,
               double rso, double k0, int quirks, double2 *S, long lds, hipStream_t st)
{
    if (T <= 0 || nrx <= 0) return 0;
    MIDAS_CHECK(T <= 65535, -1, "near_field: too many transmitters for one launch");
    const double pi = host_pi(quirks);
    const int N = cells.n;
    // receivers in chunks so that the dyads take at most ~4 GB
    const long per_rx = 9L * N * (long)sizeof(double2);
    const int rc = (int)std::max(1L, std::min((long)nrx, (long)(4e9 / (double)per_rx)));
    DBuf G, Es3;
    TRY(G.alloc(per_rx * rc));
    TRY(Es3.alloc(sizeof(double2) * 3L * rc * 2 * T));
    for (int r0 = 0; r0 < nrx; r0 += rc) {
        const int nr = std::min(rc, nrx - r0);
        hipLaunchKernelGGL(k_green_dt, dim3((N + 255) / 256, nr), dim3(256), 0, st, cells, nr, rth + r0, rph + r0, rso,
                           k0, pi, G.as<double2>(), 3L * nr);
        MIDAS_HIP(hipGetLastError());
        int r = launch_cgemm<double>(3 * nr, 2 * T, 3 * N, z64{1.0, 0.0}, G.as<double2>(), 3L * nr, 0, E, ld, 0,
                                     z64{0.0, 0.0}, Es3.as<double2>(), 3L * nr, st);
        if (r) return r;
        hipLaunchKernelGGL(k_project_nf, dim3((nr + 255) / 256, T), dim3(256), 0, st, Es3.as<double2>(), 3L * nr, nr, T,
                           r0, rth, rph, pi, S, lds);
        MIDAS_HIP(hipGetLastError());
    }
    MIDAS_HIP(hipStreamSynchronize(st));   // G and Es3 are released on return
    return 0;
}

// Orientation-averaged quantities (Compute_Scattering_Quantities.f90:593-1013):
// per transmitter t (one workgroup), over the NRx scattering receivers,
//   f(r)  = w_r (|S_hh|^2 + |S_vv|^2 + |S_vh|^2 + |S_hv|^2)    (w_r = 1 unless weighted)
//   out   = (sum_r f(r), sum_r f(r) k_i . k_s(r))
// with k = -(sin th cos ph, sin th sin ph, cos th) for the incident and the
// scattered direction (:784-788, :799-805).  The host turns the sums into
// Q_sca and g with the reference's constants.
template <bool WEIGHTED>
__global__ __launch_bounds__(256) void k_scatq_sums(const double2 *__restrict__ S, long lds, int nrx,
                                                    const double *__restrict__ rth, const double *__restrict__ rph,
                                                    const double *__restrict__ w, const double *__restrict__ tth,
                                                    const double *__restrict__ tph, double pi,
                                                    double *__restrict__ out)
{
    __shared__ double r1[256], r2[256];
    const int t = blockIdx.x;
    const double thi = tth[t] * pi / 180.0, phi = tph[t] * pi / 180.0;
    const double kix = -(sin(thi) * cos(phi)), kiy = -(sin(thi) * sin(phi)), kiz = -cos(thi);
    const double2 *vv = S + (long)(4 * t) * lds, *vh = vv + lds, *hv = vh + lds, *hh = hv + lds;
    auto a2 = [](double2 z) {
        const double m = hypot(z.x, z.y);
        return m * m;
    };
    double s1 = 0.0, s2 = 0.0;
    for (int r = threadIdx.x; r < nrx; r += 256) {
        double f = ((a2(hh[r]) + a2(vv[r])) + a2(vh[r])) + a2(hv[r]);
        if (WEIGHTED) f = w[r] * f;
        const double ths = rth[r] * pi / 180.0, phs = rph[r] * pi / 180.0;
        const double ksx = -(sin(ths) * cos(phs)), ksy = -(sin(ths) * sin(phs)), ksz = -cos(ths);
        const double cg = kix * ksx + kiy * ksy + kiz * ksz;
        s1 += f;
        s2 += f * cg;
    }
    r1[threadIdx.x] = s1;
    r2[threadIdx.x] = s2;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            r1[threadIdx.x] += r1[threadIdx.x + h];
            r2[threadIdx.x] += r2[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * t] = r1[0];
        out[2 * t + 1] = r2[0];
    }
}

int scatq_sums(const double2 *S, long lds, int nrx, int ntr, const double *rth, const double *rph, const double *w,
               const double *tth, const double *tph, int quirks, double *out, hipStream_t st)
{
    if (ntr <= 0) return 0;
    const double pi = host_pi(quirks);
    if (w)
        hipLaunchKernelGGL(k_scatq_sums<true>, dim3(ntr), dim3(256), 0, st, S, lds, nrx, rth, rph, w, tth, tph, pi,
                           out);
    else
        hipLaunchKernelGGL(k_scatq_sums<false>, dim3(ntr), dim3(256), 0, st, S, lds, nrx, rth, rph, w, tth, tph, pi,
                           out);
    MIDAS_HIP(hipGetLastError());
    return 0;
}

int far_field(const DevCells &cells, const double2 *E, long ld, int T, int nrx, const double *rth,
              const double *rph, int npol, const double *beta, double k0, int quirks, double2 *S, long lds,
              double2 *FF, double2 *E3, double2 *G, hipStream_t st)
{
    if (nrx <= 0 || T <= 0) return 0;
    MIDAS_CHECK(6 * T <= 65535 * 64, -1, "far_field: too many directions");
    double pi = host_pi(quirks);
    MIDAS_CHECK((cells.n + 3) / 4 <= 65535, -1, "far_field: too many cells for one launch");
    hipLaunchKernelGGL(k_ffcoef, dim3((nrx + 63) / 64, (cells.n + 3) / 4), dim3(64, 4), 0, st, cells, nrx, rth, rph,
                       k0, pi, FF, (long)nrx);
    hipLaunchKernelGGL(k_split_components, dim3((cells.n + 255) / 256, 2 * T), dim3(256), 0, st, E, ld, cells.n, 2 * T,
                       E3);
    MIDAS_HIP(hipGetLastError());
    int r = launch_cgemm<double>(nrx, 6 * T, cells.n, z64{1.0, 0.0}, FF, nrx, 0, E3, cells.n, 0, z64{0.0, 0.0}, G,
                                 nrx, st);
    if (r) return r;
    hipLaunchKernelGGL(k_project, dim3((nrx + 255) / 256, T), dim3(256), 0, st, G, nrx, T, rth, rph, npol, beta, pi, S,
                       lds);
    MIDAS_HIP(hipGetLastError());
    return 0;
}

int cext_cabs(const DevCells &cells, const double2 *E, long ld, int T, const double *th, const double *ph, double k0,
              int quirks, double *cext, double *cabs, hipStream_t st)
{
    if (T <= 0) return 0;
    double pi = host_pi(quirks);
    bool ref32 = (quirks & MIDAS_Q_FP32_BLOCKS) != 0, qsum = (quirks & MIDAS_Q_CEXT_SUM) != 0;
    if (ref32 && qsum)
        hipLaunchKernelGGL((k_cext_cabs<true, true>), dim3(T), dim3(256), 0, st, cells, E, ld, T, th, ph, k0, pi, cext, cabs);
    else if (ref32)
        hipLaunchKernelGGL((k_cext_cabs<true, false>), dim3(T), dim3(256), 0, st, cells, E, ld, T, th, ph, k0, pi, cext, cabs);
    else if (qsum)
        hipLaunchKernelGGL((k_cext_cabs<false, true>), dim3(T), dim3(256), 0, st, cells, E, ld, T, th, ph, k0, pi, cext, cabs);
    else
        hipLaunchKernelGGL((k_cext_cabs<false, false>), dim3(T), dim3(256), 0, st, cells, E, ld, T, th, ph, k0, pi, cext, cabs);
    MIDAS_HIP(hipGetLastError());
    return 0;
}

}  // namespace midas
