// Dyadic Green's-function element for one (receiver cell, source cell) pair.
//
// Reference formula (Compute_EFields_MoM.f90:474-519, Green_functions.f90:139-167):
//   F_pp = Gr (Tau + k^2 (r - r_p^2/r) - 3 r_p^2/r^2 Tau)
//   F_pq = Gr/r r_p r_q (-k^2 - 3 Tau/r),   Gr = e^{ikr}/(4 pi r^2),  Tau = ik - 1/r
//   Z    = -F kappa_s chi_s   (material factor of the SOURCE cell)
// Written as F = A I + C r r^T with A = Gr (Tau + k^2 r) and C = Gr/r (-k^2 - 3 Tau/r),
// so the nine entries cost one complex exp and six real products.
#pragma once
#include "midas_common.h"

namespace midas {

struct Dyad64 {
    z64 a;   // -(A kchi)      : isotropic part
    z64 c;   // -(C kchi)      : r r^T part
    double rx, ry, rz;
};

__device__ inline Dyad64 dyad64(double rx, double ry, double rz, double k0, double inv4pi, z64 kchi)
{
    double r2 = rx * rx + ry * ry + rz * rz;
    double r = sqrt(r2);
    double rinv = 1.0 / r;
    double s, c;
    sincos(k0 * r, &s, &c);
    double g = rinv * rinv * inv4pi;
    z64 gr = {c * g, s * g};
    z64 A = cmul(gr, z64{k0 * k0 * r - rinv, k0});
    z64 gri = cscale(gr, rinv);
    z64 Cc = cmul(gri, z64{-k0 * k0 + 3.0 * rinv * rinv, -3.0 * k0 * rinv});
    z64 af = cmul(A, kchi), cf = cmul(Cc, kchi);
    return {{-af.r, -af.i}, {-cf.r, -cf.i}, rx, ry, rz};
}

// Entry (row component po, column component ps) of the 3x3 block.
__device__ inline z64 dyad_entry(const Dyad64 &d, int po, int ps)
{
    double rp = po == 0 ? d.rx : (po == 1 ? d.ry : d.rz);
    double rq = ps == 0 ? d.rx : (ps == 1 ? d.ry : d.rz);
    double w = rp * rq;
    z64 e = {d.c.r * w, d.c.i * w};
    if (po == ps) { e.r += d.a.r; e.i += d.a.i; }
    return e;
}

// ---------------------------------------------------------------------------
// Reference rounding: Gr_mn, Tau_mn, f_kapChe and F** are single-precision
// `Complex` in Green_functions.f90 (quirk Q3).  Mixed real(8)/complex(4)
// expressions promote to double, assignments round to float, and the final
// -F*f_kapChe is a complex(4) product.  out[3*ps+po].
// ---------------------------------------------------------------------------
__device__ inline float2 f2(double r, double i) { return make_float2((float)r, (float)i); }

// The single-precision dyad F (row p, column q) and f_kapChe of one pair.
__device__ inline void dyad_ref32_F(double rx, double ry, double rz, double k0, double pi, double kappa, z64 che,
                                    float2 F[3][3], float2 &f)
{
#pragma clang fp contract(off)
    double r = sqrt(rx * rx + ry * ry + rz * rz);
    double s, c;
    sincos(k0 * r, &s, &c);
    double den = 4.0 * pi * (r * r);
    float2 gr = f2(c / den, s / den);
    float2 tau = f2(-1.0 / r, k0);
    f = f2(kappa * che.r, kappa * che.i);
    double rv[3] = {rx, ry, rz};
    for (int p = 0; p < 3; ++p) {
        double rp = rv[p];
        double t1 = k0 * k0 * (r - rp * rp / r);
        double t2 = 3.0 * rp * rp / (r * r);
        // (Tau + t1) - t2*Tau, then Gr * (.)
        double br = (double)tau.x + t1 - t2 * (double)tau.x;
        double bi = (double)tau.y - t2 * (double)tau.y;
        F[p][p] = f2((double)gr.x * br - (double)gr.y * bi, (double)gr.x * bi + (double)gr.y * br);
    }
    // plain operators: this function runs with contraction off, so every
    // product and sum is rounded separately as in the reference
    float2 tau3 = make_float2(3.0f * tau.x, 3.0f * tau.y);
    for (int p = 0; p < 3; ++p)
        for (int q = p + 1; q < 3; ++q) {
            double w = rv[p] * rv[q];
            double gr_r = (double)gr.x / r * w, gr_i = (double)gr.y / r * w;
            double br = -k0 * k0 - (double)tau3.x / r, bi = -(double)tau3.y / r;
            F[p][q] = F[q][p] = f2(gr_r * br - gr_i * bi, gr_r * bi + gr_i * br);
        }
}

// -F * f_kapChe as a complex(4) product, widened (Green_functions.f90:160-167)
__device__ inline z64 neg_fprod(float2 a, float2 f)
{
#pragma clang fp contract(off)
    float re = a.x * f.x - a.y * f.y;
    float im = a.x * f.y + a.y * f.x;
    return {-(double)re, -(double)im};
}

__device__ inline void dyad_ref32(double rx, double ry, double rz, double k0, double pi, double kappa, z64 che,
                                  z64 out[9])
{
    float2 F[3][3], f;
    dyad_ref32_F(rx, ry, rz, k0, pi, kappa, che, F, f);
    for (int ps = 0; ps < 3; ++ps)
        for (int po = 0; po < 3; ++po) out[3 * ps + po] = neg_fprod(F[po][ps], f);
}

}  // namespace midas
