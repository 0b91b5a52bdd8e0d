// Batched Hermitian eigensolver for the CBF subspaces (Compute_EFields_CBFME.f90:418-444,
// Compute_EFields_CBFME_ACA.f90:356-381: the reference's GESVD of every block's
// solutions; here the eigen-decomposition of their Gram matrix).
//
// One workgroup per matrix, all three stages in one launch:
//   1. Householder reduction to real symmetric tridiagonal form,
//      Q^H G Q = T (Golub & Van Loan, Matrix Computations, 4th ed., 8.3.1):
//      for column j the reflector H_j = I - tau v v^H with H_j^H x = beta e_1,
//      beta real (the complex reflector of LAPACK's zlarfg convention, so T is
//      real), then the two-sided update of the trailing block in rank-2 form
//      A -= w v^H + v w^H, y = tau A v, w = y - (conj(tau) v^H y / 2) v.
//   2. Implicit QL with Wilkinson-type shifts on T, the rotations accumulated
//      into Z (the tql2 algorithm of Bowdler, Martin, Reinsch & Wilkinson,
//      Numer. Math. 11 (1968) 293-306, in the formulation of Press et al.,
//      Numerical Recipes, 11.3), deflating where |e_m| <= eps |T|.  Lane 0 runs each sweep's scalar recurrence
//      and records its rotations in LDS; then every lane applies them to its
//      own rows of Z.
//   3. Eigenvalues sorted ascending, eigenvectors U = H_0 H_1 ... H_{n-2} Z
//      written over G (column c belongs to eigenvalue c), as LAPACK's heev.
// Every global row of G and Z belongs to one lane (rows tid, tid + 256, ...)
// in stages 1 and 2, so no lane reads another lane's global writes there; the
// back-transformation works on U^T (lane per column, coalesced) in a
// workspace.  Workspace per matrix: V (the reflectors, n^2 complex), Z (n^2
// real), U^T (n^2 complex).
#include "midas_common.h"
#include "midas_lu.h"

#include <cfloat>

namespace midas {

constexpr int HEEV_TH = 256;
constexpr int HEEV_MAXIT = 60;   // QL iterations per eigenvalue

__device__ inline double2 zmul2(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
// conj(a) b
__device__ inline double2 zcmul2(double2 a, double2 b) { return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x); }

// sum over the workgroup (every lane gets it); red holds HEEV_TH / 64 partials
__device__ inline double2 heev_sum(double2 v, double2 *red)
{
    for (int o = 32; o > 0; o >>= 1) {
        v.x += __shfl_xor(v.x, o);
        v.y += __shfl_xor(v.y, o);
    }
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double2 s = red[0];
    for (int i = 1; i < HEEV_TH / 64; ++i) {
        s.x += red[i].x;
        s.y += red[i].y;
    }
    return s;
}

__global__ __launch_bounds__(HEEV_TH) void k_heev(int n, double2 *Gall, long ld, long gstride, double *Dall,
                                                  long dstride, int *info, double2 *Vall, double *Zall, double2 *Uall)
{
    extern __shared__ double2 smem[];
    double2 *sv = smem, *sw = sv + n, *sx = sw + n, *stau = sx + n;
    double *sd = (double *)(stau + n), *se = sd + n, *sc = se + n, *ss = sc + n;
    int *sperm = (int *)(ss + n);
    __shared__ double2 red[HEEV_TH / 64];
    __shared__ double2 s_tau, s_scale;
    __shared__ int s_flag, s_m, s_iend;
    __shared__ double s_tn;

    const int tid = threadIdx.x;
    const long b = blockIdx.x, nn = (long)n * n;
    double2 *G = Gall + b * gstride;
    double2 *V = Vall + b * nn, *UT = Uall + b * nn;
    double *Z = Zall + b * nn, *D = Dall + b * dstride;

    // ---- 1. tridiagonalisation; sx holds the current column j (rows >= j)
    for (int g = tid; g < n; g += HEEV_TH) sx[g] = G[g];
    __syncthreads();
    for (int j = 0; j + 1 < n; ++j) {
        double2 part = make_double2(0.0, 0.0);
        for (int g = tid; g < n; g += HEEV_TH)
            if (g >= j + 2) part.x += sx[g].x * sx[g].x + sx[g].y * sx[g].y;
        const double xn2 = heev_sum(part, red).x;
        if (tid == 0) {
            sd[j] = sx[j].x;
            const double2 al = sx[j + 1];
            if (xn2 == 0.0 && al.y == 0.0) {
                s_tau = make_double2(0.0, 0.0);
                se[j] = al.x;
            } else {
                const double beta = -copysign(sqrt(al.x * al.x + al.y * al.y + xn2), al.x);
                s_tau = make_double2((beta - al.x) / beta, -al.y / beta);
                // 1 / (alpha - beta)
                const double2 q = make_double2(al.x - beta, al.y);
                const double q2 = q.x * q.x + q.y * q.y;
                s_scale = make_double2(q.x / q2, -q.y / q2);
                se[j] = beta;
            }
            stau[j] = s_tau;
        }
        __syncthreads();
        const double2 tau = s_tau, scale = s_scale;
        if (tau.x == 0.0 && tau.y == 0.0) {
            // H_j = I: the trailing block stays; its first column is next
            for (int g = tid; g < n; g += HEEV_TH)
                if (g >= j + 1) sx[g] = G[g + (long)(j + 1) * ld];
            __syncthreads();
            continue;
        }
        for (int g = tid; g < n; g += HEEV_TH)
            if (g >= j + 1) {
                const double2 v = g == j + 1 ? make_double2(1.0, 0.0) : zmul2(sx[g], scale);
                sv[g] = v;
                V[g + (long)j * n] = v;
            }
        __syncthreads();
        // y = tau A v on this lane's rows
        double2 dot = make_double2(0.0, 0.0);
        for (int g = tid; g < n; g += HEEV_TH)
            if (g >= j + 1) {
                double2 a0 = make_double2(0.0, 0.0);
                const double2 *row = G + g;
#pragma unroll 4
                for (int c = j + 1; c < n; ++c) {
                    const double2 p0 = zmul2(row[(long)c * ld], sv[c]);
                    a0.x += p0.x;
                    a0.y += p0.y;
                }
                const double2 y = zmul2(tau, a0);
                sw[g] = y;
                const double2 t = zcmul2(sv[g], y);
                dot.x += t.x;
                dot.y += t.y;
            }
        dot = heev_sum(dot, red);
        // gamma = conj(tau) (v^H y) / 2 (real in exact arithmetic)
        const double2 gam = zcmul2(tau, dot);
        for (int g = tid; g < n; g += HEEV_TH)
            if (g >= j + 1) {
                const double2 t = zmul2(make_double2(0.5 * gam.x, 0.5 * gam.y), sv[g]);
                sw[g] = make_double2(sw[g].x - t.x, sw[g].y - t.y);
            }
        __syncthreads();
        // A -= w v^H + v w^H on this lane's rows; the new first column is next
        for (int g = tid; g < n; g += HEEV_TH)
            if (g >= j + 1) {
                const double2 wg = sw[g], vg = sv[g];
                double2 *row = G + g;
#pragma unroll 4
                for (int c = j + 1; c < n; ++c) {
                    const double2 a = row[(long)c * ld];
                    const double2 t1 = zcmul2(sv[c], wg);   // w_g conj(v_c)
                    const double2 t2 = zcmul2(sw[c], vg);   // v_g conj(w_c)
                    const double2 r = make_double2(a.x - t1.x - t2.x, a.y - t1.y - t2.y);
                    row[(long)c * ld] = r;
                    if (c == j + 1) sx[g] = r;
                }
            }
        __syncthreads();
    }
    if (tid == 0) {
        sd[n - 1] = sx[n - 1].x;
        se[n - 1] = 0.0;
        // |T| for the deflation test: an off-diagonal below eps |T| is zero
        // (absolute, as tql2's running norm; a test relative to the two
        // neighbouring diagonal entries never passes in the rounding-level
        // tail of a rank-deficient Gram matrix)
        double tn = 0.0;
        for (int i = 0; i < n; ++i) tn = fmax(tn, fabs(sd[i]) + fabs(se[i]));
        s_tn = tn;
    }

    // ---- 2. implicit QL on (sd, se), rotations into Z (column-major, n x n)
    for (int g = tid; g < n; g += HEEV_TH)
        for (int i = 0; i < n; ++i) Z[g + (long)i * n] = g == i ? 1.0 : 0.0;
    __syncthreads();
    int fail = 0;
    for (int l = 0; l < n && !fail; ++l) {
        int iter = 0;
        for (;;) {
            if (tid == 0) {
                int m = l;
                for (; m + 1 < n; ++m)
                    if (fabs(se[m]) <= DBL_EPSILON * s_tn) break;
                if (m == l) {
                    s_flag = 0;
                } else if (iter++ >= HEEV_MAXIT) {
                    s_flag = 2;
                } else {
                    double g = (sd[l + 1] - sd[l]) / (2.0 * se[l]);
                    double r = hypot(g, 1.0);
                    g = sd[m] - sd[l] + se[l] / (g + copysign(r, g));
                    double s = 1.0, c = 1.0, p = 0.0;
                    int i = m - 1;
                    bool early = false;
                    for (; i >= l; --i) {
                        const double f = s * se[i], bb = c * se[i];
                        r = hypot(f, g);
                        se[i + 1] = r;
                        if (r == 0.0) {
                            sd[i + 1] -= p;
                            se[m] = 0.0;
                            early = true;
                            break;
                        }
                        s = f / r;
                        c = g / r;
                        g = sd[i + 1] - p;
                        r = (sd[i] - g) * s + 2.0 * c * bb;
                        p = s * r;
                        sd[i + 1] = g + p;
                        g = c * r - bb;
                        sc[i] = c;
                        ss[i] = s;
                    }
                    if (!early) {
                        sd[l] -= p;
                        se[l] = g;
                        se[m] = 0.0;
                    }
                    s_m = m;
                    s_iend = early ? i + 1 : l;   // rotations i = m-1 .. s_iend
                    s_flag = 1;
                }
            }
            __syncthreads();
            const int flag = s_flag;
            if (flag == 1) {
                const int m = s_m, iend = s_iend;
                if (iend <= m - 1)
                    for (int g = tid; g < n; g += HEEV_TH) {
                        double car = Z[g + (long)m * n];
#pragma unroll 4
                        for (int i = m - 1; i >= iend; --i) {
                            const double zi = Z[g + (long)i * n];
                            Z[g + (long)(i + 1) * n] = ss[i] * zi + sc[i] * car;
                            car = sc[i] * zi - ss[i] * car;
                        }
                        Z[g + (long)iend * n] = car;
                    }
            }
            // every lane has read the flags and the rotations before lane 0
            // writes the next ones
            __syncthreads();
            if (flag == 2) fail = 1;
            if (flag != 1) break;
        }
    }

    // ---- 3. ascending order, U^T = (Q Z)^T, U written over G
    for (int i = tid; i < n; i += HEEV_TH) {
        const double di = sd[i];
        int rank = 0;
        for (int q = 0; q < n; ++q) rank += (sd[q] < di) || (sd[q] == di && q < i);
        sperm[rank] = i;
    }
    __syncthreads();
    for (int c = tid; c < n; c += HEEV_TH) {
        D[c] = sd[sperm[c]];
        const double *zc = Z + (long)sperm[c] * n;
        for (int g = 0; g < n; ++g) UT[c + (long)g * n] = make_double2(zc[g], 0.0);
    }
    for (int j = n - 2; j >= 0; --j) {
        const double2 tau = stau[j];
        if (tau.x == 0.0 && tau.y == 0.0) continue;
        __syncthreads();
        for (int g = tid; g < n; g += HEEV_TH)
            if (g >= j + 1) sv[g] = V[g + (long)j * n];
        __syncthreads();
        for (int c = tid; c < n; c += HEEV_TH) {
            double2 s = make_double2(0.0, 0.0);
#pragma unroll 4
            for (int g = j + 1; g < n; ++g) {
                const double2 t = zcmul2(sv[g], UT[c + (long)g * n]);
                s.x += t.x;
                s.y += t.y;
            }
            s = zmul2(tau, s);
#pragma unroll 4
            for (int g = j + 1; g < n; ++g) {
                const double2 t = zmul2(sv[g], s);
                double2 &u = UT[c + (long)g * n];
                u = make_double2(u.x - t.x, u.y - t.y);
            }
        }
    }
    __syncthreads();
    for (int g = tid; g < n; g += HEEV_TH)
        for (int c = 0; c < n; ++c) G[g + (long)c * ld] = UT[c + (long)g * n];
    if (tid == 0) info[b] = fail;
}

// ---------------------------------------------------------------------------
// Large orders (the Gram matrices of 400+ CBF directions, n up to 1024), where
// the O(n^3) global traffic of the kernel above (unblocked two-sided updates,
// rotations accumulated for all n vectors) dominates.  One 1024-thread
// workgroup per matrix, lane g owning row g:
//   1. blocked tridiagonalisation (the panel scheme of LAPACK's latrd, with the
//      formulas above): per column, the column corrected by the panel's
//      earlier (V, W), its reflector, y = tau (A v - V W^H v - W V^H v) with
//      the trailing matrix as of the panel's start (one read of it per
//      column), w; per panel of HL_NB columns one update A -= V W^H + W V^H;
//      the reflectors stay in G's columns.
//   2. eigenvalues by bisection on Sturm counts (Barth, Martin & Wilkinson,
//      Numer. Math. 9 (1967) 386-393): the largest, then only those at or
//      above cut * lambda_max (cut <= 0: all), one lane each;
//   3. their eigenvectors of T by inverse iteration (Wilkinson; the scheme of
//      LAPACK's stein): tridiagonal LU with partial pivoting, three solves from
//      a fixed pseudo-random start, eigenvalues closer than 1e-3 |T| taken as a
//      cluster whose vectors are orthogonalised (modified Gram-Schmidt) against
//      its earlier ones after every solve; one wave per cluster;
//   4. back-transformation of those vectors only, U = H_0 ... H_{n-2} Z.
// Outputs as the kernel above for the computed eigenpairs (ascending, in the
// last kv places); the eigenvalues not computed are 0 and their columns zero.
constexpr int HL_TH = 1024, HL_NB = 8, HL_WAVES = HL_TH / 64;
constexpr int HL_CHUNK = 64;   // columns of V, W staged per step of the panel update

__device__ inline double hl_wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// sums of cnt complex values over the workgroup into out[0..cnt) (LDS)
__device__ inline void hl_block_sum(const double2 *vals, int cnt, double2 *scratch, double2 *out)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int k = 0; k < cnt; ++k) {
        double2 v = vals[k];
        v.x = hl_wave_sum(v.x);
        v.y = hl_wave_sum(v.y);
        if (lane == 0) scratch[w * 2 * HL_NB + k] = v;
    }
    __syncthreads();
    if ((int)threadIdx.x < cnt) {
        double2 t = make_double2(0.0, 0.0);
        for (int q = 0; q < HL_WAVES; ++q) {
            t.x += scratch[q * 2 * HL_NB + threadIdx.x].x;
            t.y += scratch[q * 2 * HL_NB + threadIdx.x].y;
        }
        out[threadIdx.x] = t;
    }
    __syncthreads();
}

// #eigenvalues of T below x
__device__ inline int hl_sturm(const double *d, const double *e, int n, double x, double pivmin)
{
    int cnt = 0;
    double q = d[0] - x;
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += q < 0.0;
    for (int i = 1; i < n; ++i) {
        q = d[i] - x - e[i - 1] * e[i - 1] / q;
        if (fabs(q) < pivmin) q = -pivmin;
        cnt += q < 0.0;
    }
    return cnt;
}

// eigenvalue k (ascending) of T in [lo, hi)
__device__ inline double hl_bisect(const double *d, const double *e, int n, int k, double lo, double hi,
                                   double pivmin)
{
    const double tol = 2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)) + 2.0 * pivmin;
    for (int it = 0; it < 256 && hi - lo > tol; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (hl_sturm(d, e, n, mid, pivmin) <= k)
            lo = mid;
        else
            hi = mid;
    }
    return 0.5 * (lo + hi);
}

__global__ __launch_bounds__(HL_TH) void k_heev_large(int n, double2 *Gall, long ld, long gstride, double *Dall,
                                                      long dstride, int *info, double cut, double2 *Wpall,
                                                      double2 *UTall, double *Zvall, double *LUall)
{
    extern __shared__ double2 hsm[];
    double2 *sx = hsm, *sv = sx + n, *sy = sv + n, *stau = sy + n;
    double2 *stage = stau + n;   // 2 HL_NB HL_CHUNK values
    double *sd = (double *)(stage + 2 * HL_NB * HL_CHUNK), *se = sd + n, *sev = se + n;
    int *scl = (int *)(sev + n);   // cluster starts, n + 1
    __shared__ double2 red[HL_WAVES * 2 * HL_NB], rsum[2 * HL_NB];
    __shared__ double2 s_tau, s_scale;
    __shared__ double s_gl, s_gu, s_piv, s_nrm;
    __shared__ int s_kv, s_ncl;

    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int g = tid;   // this lane's row
    const bool row = g < n;
    const long b = blockIdx.x, nn = (long)n * n;
    double2 *G = Gall + b * gstride, *Wp = Wpall + b * (long)n * HL_NB, *UT = UTall + b * nn;
    double *Zv = Zvall + b * nn, *D = Dall + b * dstride;
    double *LU = LUall + (b * HL_WAVES + wv) * 6L * n;   // this wave's u0 u1 u2 l piv x
    auto V = [&](int r, int k) { return G[r + (long)k * ld]; };

    // ---- 1. blocked tridiagonalisation
    for (int j0 = 0; j0 < n - 1; j0 += HL_NB) {
        const int jend = min(j0 + HL_NB, n - 1);
        for (int j = j0; j < jend; ++j) {
            const int jj = j - j0;
            // (a) column j (rows >= j) as of now: start value minus the panel's terms
            if (row && g >= j) {
                double2 a = G[g + (long)j * ld];
                for (int k = 0; k < jj; ++k) {
                    const double2 t1 = zcmul2(Wp[j + (long)k * n], V(g, j0 + k));   // V(g,k) conj(W(j,k))
                    const double2 t2 = zcmul2(V(j, j0 + k), Wp[g + (long)k * n]);   // W(g,k) conj(V(j,k))
                    a.x -= t1.x + t2.x;
                    a.y -= t1.y + t2.y;
                }
                sx[g] = a;
            }
            __syncthreads();
            // (b) reflector of sx[j+1 .. n-1]
            double2 part = make_double2(row && g >= j + 2 ? sx[g].x * sx[g].x + sx[g].y * sx[g].y : 0.0, 0.0);
            hl_block_sum(&part, 1, red, rsum);
            if (tid == 0) {
                const double xn2 = rsum[0].x;
                sd[j] = sx[j].x;
                const double2 al = sx[j + 1];
                if (xn2 == 0.0 && al.y == 0.0) {
                    s_tau = make_double2(0.0, 0.0);
                    s_scale = make_double2(0.0, 0.0);
                    se[j] = al.x;
                } else {
                    const double beta = -copysign(sqrt(al.x * al.x + al.y * al.y + xn2), al.x);
                    s_tau = make_double2((beta - al.x) / beta, -al.y / beta);
                    const double2 q = make_double2(al.x - beta, al.y);
                    const double q2 = q.x * q.x + q.y * q.y;
                    s_scale = make_double2(q.x / q2, -q.y / q2);
                    se[j] = beta;
                }
                stau[j] = s_tau;
            }
            __syncthreads();
            const double2 tau = s_tau, scale = s_scale;
            const bool act = row && g >= j + 1;
            if (act) {
                const double2 v = g == j + 1 ? make_double2(1.0, 0.0) : zmul2(sx[g], scale);
                sv[g] = v;
                G[g + (long)j * ld] = v;
            }
            __syncthreads();
            if (tau.x == 0.0 && tau.y == 0.0) {
                if (act) Wp[g + (long)jj * n] = make_double2(0.0, 0.0);
                __syncthreads();
                continue;
            }
            // (c) y = tau (A v - V W^H v - W V^H v), w = y - conj(tau) (v^H y) v / 2;
            // A v by all lanes: lane (row r, part p) sums the columns c = p mod parts
            // (all lanes busy while the trailing block shrinks)
            const int L = n - j - 1, parts = max(1, min(HL_TH / L, 16));
            if (tid < L * parts) {
                const int r = j + 1 + tid % L, p = tid / L;
                const double2 *rp = G + r;
                double2 a = make_double2(0.0, 0.0), a2 = a;
                int c = j + 1 + p;
                // eight loads in flight per lane
                for (; c + 7 * parts < n; c += 8 * parts) {
                    double2 x[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) x[u] = rp[(long)(c + u * parts) * ld];
#pragma unroll
                    for (int u = 0; u < 8; u += 2) {
                        const double2 q = zmul2(x[u], sv[c + u * parts]);
                        const double2 q2 = zmul2(x[u + 1], sv[c + (u + 1) * parts]);
                        a.x += q.x;
                        a.y += q.y;
                        a2.x += q2.x;
                        a2.y += q2.y;
                    }
                }
                for (; c < n; c += parts) {
                    const double2 q = zmul2(rp[(long)c * ld], sv[c]);
                    a.x += q.x;
                    a.y += q.y;
                }
                stage[tid] = make_double2(a.x + a2.x, a.y + a2.y);
            }
            __syncthreads();
            double2 acc = make_double2(0.0, 0.0);
            if (act)
                for (int p = 0; p < parts; ++p) {
                    acc.x += stage[p * L + (g - j - 1)].x;
                    acc.y += stage[p * L + (g - j - 1)].y;
                }
            if (jj > 0) {
                // (W^H v)_k and (V^H v)_k, k < jj, summed over the rows
                const double2 vg = act ? sv[g] : make_double2(0.0, 0.0);
                for (int k = 0; k < 2 * jj; ++k) {
                    double2 t = make_double2(0.0, 0.0);
                    if (act) t = k < jj ? zcmul2(Wp[g + (long)k * n], vg) : zcmul2(V(g, j0 + k - jj), vg);
                    t.x = hl_wave_sum(t.x);
                    t.y = hl_wave_sum(t.y);
                    if (lane == 0) red[wv * 2 * HL_NB + k] = t;
                }
                __syncthreads();
                if (tid < 2 * jj) {
                    double2 t = make_double2(0.0, 0.0);
                    for (int q = 0; q < HL_WAVES; ++q) {
                        t.x += red[q * 2 * HL_NB + tid].x;
                        t.y += red[q * 2 * HL_NB + tid].y;
                    }
                    rsum[tid] = t;
                }
                __syncthreads();
            }
            double2 dot = make_double2(0.0, 0.0);
            double2 y = make_double2(0.0, 0.0);
            if (act) {
                for (int k = 0; k < jj; ++k) {
                    const double2 t1 = zmul2(V(g, j0 + k), rsum[k]);
                    const double2 t2 = zmul2(Wp[g + (long)k * n], rsum[jj + k]);
                    acc.x -= t1.x + t2.x;
                    acc.y -= t1.y + t2.y;
                }
                y = zmul2(tau, acc);
                dot = zcmul2(sv[g], y);
            }
            hl_block_sum(&dot, 1, red, rsum);
            const double2 gam = zcmul2(tau, rsum[0]);
            if (act) {
                const double2 t = zmul2(make_double2(0.5 * gam.x, 0.5 * gam.y), sv[g]);
                Wp[g + (long)jj * n] = make_double2(y.x - t.x, y.y - t.y);
            }
            __syncthreads();
        }
        // (d) trailing update A -= V W^H + W V^H over rows / columns >= jend
        const int nbk = jend - j0;
        double2 vg[HL_NB], wg[HL_NB];
        const bool tr = row && g >= jend;
#pragma unroll
        for (int k = 0; k < HL_NB; ++k) {
            vg[k] = tr && k < nbk ? V(g, j0 + k) : make_double2(0.0, 0.0);
            wg[k] = tr && k < nbk ? Wp[g + (long)k * n] : make_double2(0.0, 0.0);
        }
        for (int c0 = jend; c0 < n; c0 += HL_CHUNK) {
            for (int idx = tid; idx < HL_CHUNK * HL_NB; idx += HL_TH) {
                const int cc = idx % HL_CHUNK, k = idx / HL_CHUNK, c = c0 + cc;
                const bool ok = c < n && k < nbk;
                stage[k * HL_CHUNK + cc] = ok ? V(c, j0 + k) : make_double2(0.0, 0.0);
                stage[(HL_NB + k) * HL_CHUNK + cc] = ok ? Wp[c + (long)k * n] : make_double2(0.0, 0.0);
            }
            __syncthreads();
            if (tr) {
                const int ce = min(c0 + HL_CHUNK, n);
                double2 *rp = G + g;
#pragma unroll 2
                for (int c = c0; c < ce; ++c) {
                    const int cc = c - c0;
                    double2 a = rp[(long)c * ld];
#pragma unroll
                    for (int k = 0; k < HL_NB; ++k) {
                        const double2 t1 = zcmul2(stage[(HL_NB + k) * HL_CHUNK + cc], vg[k]);   // V(g,k) conj(W(c,k))
                        const double2 t2 = zcmul2(stage[k * HL_CHUNK + cc], wg[k]);             // W(g,k) conj(V(c,k))
                        a.x -= t1.x + t2.x;
                        a.y -= t1.y + t2.y;
                    }
                    rp[(long)c * ld] = a;
                }
            }
            __syncthreads();
        }
    }
    if (tid == 0) {
        sd[n - 1] = G[(n - 1) + (long)(n - 1) * ld].x;
        se[n - 1] = 0.0;
        double gl = sd[0], gu = sd[0], nrm = 0.0, emax2 = 0.0;
        for (int i = 0; i < n; ++i) {
            const double r = (i > 0 ? fabs(se[i - 1]) : 0.0) + (i + 1 < n ? fabs(se[i]) : 0.0);
            gl = fmin(gl, sd[i] - r);
            gu = fmax(gu, sd[i] + r);
            nrm = fmax(nrm, fabs(sd[i]) + r);
            if (i + 1 < n) emax2 = fmax(emax2, se[i] * se[i]);
        }
        const double w = 2.0 * DBL_EPSILON * fmax(nrm, DBL_MIN) * n + DBL_MIN;
        s_gl = gl - w;
        s_gu = gu + w;
        s_nrm = nrm;
        s_piv = DBL_MIN * fmax(1.0, emax2);
        // ---- 2a. the largest eigenvalue, then how many are wanted
        const double lmax = hl_bisect(sd, se, n, n - 1, s_gl, s_gu, s_piv);
        int kv = n;
        if (cut > 0.0) kv = lmax > 0.0 ? n - hl_sturm(sd, se, n, cut * lmax, s_piv) : 0;
        s_kv = kv;
    }
    __syncthreads();
    const int kv = s_kv, k0 = n - kv;
    // ---- 2b. eigenvalues k0 .. n-1
    if (tid < kv) sev[k0 + tid] = hl_bisect(sd, se, n, k0 + tid, s_gl, s_gu, s_piv);
    __syncthreads();
    if (tid == 0) {
        // separate coincident eigenvalues slightly (as stein), then clusters
        const double ortol = 1e-3 * s_nrm;
        int ncl = 0;
        for (int i = k0; i < n; ++i) {
            if (i > k0) {
                const double pert = 10.0 * DBL_EPSILON * fmax(fabs(sev[i]), s_nrm * 1e-3);
                if (sev[i] - sev[i - 1] < pert) sev[i] = sev[i - 1] + pert;
            }
            if (i == k0 || sev[i] - sev[i - 1] > ortol) scl[ncl++] = i;
        }
        scl[ncl] = n;
        s_ncl = ncl;
    }
    __syncthreads();
    // ---- 3. inverse iteration, one wave per cluster
    {
        double *u0 = LU, *u1 = LU + n, *u2 = LU + 2L * n, *lm = LU + 3L * n, *pv = LU + 4L * n, *x = LU + 5L * n;
        const double tiny = DBL_EPSILON * fmax(s_nrm, DBL_MIN);
        for (int cl = wv; cl < s_ncl; cl += HL_WAVES)
            for (int i = scl[cl]; i < scl[cl + 1]; ++i) {
                const double lam = sev[i];
                const int c = i - k0;
                if (lane == 0) {
                    // LU with partial pivoting of T - lam I
                    double c0 = sd[0] - lam, c1 = n > 1 ? se[0] : 0.0, c2 = 0.0;
                    for (int q = 0; q + 1 < n; ++q) {
                        const double r0 = se[q], r1 = sd[q + 1] - lam, r2 = q + 2 < n ? se[q + 1] : 0.0;
                        double nc0, nc1;
                        if (fabs(c0) >= fabs(r0)) {
                            if (fabs(c0) < tiny) c0 = copysign(tiny, c0);
                            const double l = r0 / c0;
                            u0[q] = c0;
                            u1[q] = c1;
                            u2[q] = c2;
                            pv[q] = 0.0;
                            lm[q] = l;
                            nc0 = r1 - l * c1;
                            nc1 = r2 - l * c2;
                        } else {
                            const double l = c0 / r0;
                            u0[q] = r0;
                            u1[q] = r1;
                            u2[q] = r2;
                            pv[q] = 1.0;
                            lm[q] = l;
                            nc0 = c1 - l * r1;
                            nc1 = c2 - l * r2;
                        }
                        c0 = nc0;
                        c1 = nc1;
                        c2 = 0.0;
                    }
                    u0[n - 1] = fabs(c0) < tiny ? copysign(tiny, c0) : c0;
                }
                // fixed pseudo-random start
                for (int q = lane; q < n; q += 64) {
                    unsigned h = (unsigned)(q + 1) * 2654435761u ^ (unsigned)(c + 7) * 40503u;
                    h ^= h >> 15;
                    h *= 2246822519u;
                    h ^= h >> 13;
                    x[q] = (double)(h & 0xffffff) / 16777216.0 - 0.5;
                }
                __threadfence_block();
                __builtin_amdgcn_wave_barrier();
                for (int it = 0; it < 3; ++it) {
                    if (lane == 0) {
                        for (int q = 0; q + 1 < n; ++q) {
                            if (pv[q] != 0.0) {
                                const double t = x[q];
                                x[q] = x[q + 1];
                                x[q + 1] = t;
                            }
                            x[q + 1] -= lm[q] * x[q];
                        }
                        x[n - 1] /= u0[n - 1];
                        if (n > 1) x[n - 2] = (x[n - 2] - u1[n - 2] * x[n - 1]) / u0[n - 2];
                        for (int q = n - 3; q >= 0; --q) x[q] = (x[q] - u1[q] * x[q + 1] - u2[q] * x[q + 2]) / u0[q];
                    }
                    __threadfence_block();
                    __builtin_amdgcn_wave_barrier();
                    // orthogonal to the cluster's earlier vectors, unit length
                    for (int p = scl[cl]; p < i; ++p) {
                        const double *zp = Zv + (long)(p - k0) * n;
                        double dp = 0.0;
                        for (int q = lane; q < n; q += 64) dp += x[q] * zp[q];
                        dp = hl_wave_sum(dp);
                        for (int q = lane; q < n; q += 64) x[q] -= dp * zp[q];
                    }
                    double nr = 0.0;
                    for (int q = lane; q < n; q += 64) nr += x[q] * x[q];
                    nr = hl_wave_sum(nr);
                    const double inv = nr > 0.0 ? 1.0 / sqrt(nr) : 0.0;
                    for (int q = lane; q < n; q += 64) x[q] *= inv;
                    __threadfence_block();
                    __builtin_amdgcn_wave_barrier();
                }
                for (int q = lane; q < n; q += 64) Zv[(long)c * n + q] = x[q];
                __threadfence_block();
                __builtin_amdgcn_wave_barrier();
            }
    }
    __syncthreads();
    // ---- 4. U = H_0 ... H_{n-2} Z for the kv vectors, U^T layout (row g: kv values)
    for (long idx = tid; idx < (long)n * kv; idx += HL_TH) {
        const int gg = (int)(idx / kv), c = (int)(idx % kv);
        UT[idx] = make_double2(Zv[(long)c * n + gg], 0.0);
    }
    const int parts = kv > 0 ? max(1, min(HL_TH / kv, 2 * HL_NB * HL_CHUNK / kv)) : 1;
    for (int j = n - 2; j >= 0 && kv > 0; --j) {
        const double2 tau = stau[j];
        if (tau.x == 0.0 && tau.y == 0.0) continue;
        __syncthreads();
        if (row && g >= j + 1) sv[g] = G[g + (long)j * ld];
        __syncthreads();
        if (tid < kv * parts) {
            const int c = tid % kv, p = tid / kv;
            double2 s = make_double2(0.0, 0.0);
            for (int q = j + 1 + p; q < n; q += parts) {
                const double2 t = zcmul2(sv[q], UT[(long)q * kv + c]);
                s.x += t.x;
                s.y += t.y;
            }
            stage[tid] = s;
        }
        __syncthreads();
        if (tid < kv) {
            double2 s = make_double2(0.0, 0.0);
            for (int p = 0; p < parts; ++p) {
                s.x += stage[p * kv + tid].x;
                s.y += stage[p * kv + tid].y;
            }
            sy[tid] = zmul2(tau, s);
        }
        __syncthreads();
        for (long idx = tid; idx < (long)(n - j - 1) * kv; idx += HL_TH) {
            const int q = j + 1 + (int)(idx / kv), c = (int)(idx % kv);
            const double2 t = zmul2(sv[q], sy[c]);
            double2 &u = UT[(long)q * kv + c];
            u = make_double2(u.x - t.x, u.y - t.y);
        }
    }
    __syncthreads();
    for (int i = tid; i < n; i += HL_TH) D[i] = i >= k0 ? sev[i] : 0.0;
    for (long idx = tid; idx < nn; idx += HL_TH) {
        const int gg = (int)(idx % n), col = (int)(idx / n);
        G[gg + (long)col * ld] = col >= k0 ? UT[(long)gg * kv + (col - k0)] : make_double2(0.0, 0.0);
    }
    if (tid == 0) info[b] = 0;
}

// ---------------------------------------------------------------------------
// Pivoted Cholesky factorisation of a batch of Hermitian positive
// semi-definite matrices, stopped at numerical rank: G ~= L L^H with L n x r,
// r the first step whose largest remaining diagonal is <= tolrel times G's
// largest diagonal (the remainder S = G - L L^H is semi-definite with
// |S| <= (n - r) * that diagonal; Higham, Accuracy and Stability of Numerical
// Algorithms, 2nd ed., 10.3).  The CBF Gram matrices have a numerical rank
// well below their order, so their eigenpairs come from the r x r matrix
// L^H L instead (same nonzero eigenvalues; eigenvectors L w / sqrt(lambda)).
// One 1024-thread workgroup per matrix, lane i owning row i (n <= 1024);
// columns r .. n-1 of L are zeroed; rank[b] = r.
constexpr int PC_TH = 1024;

__global__ __launch_bounds__(PC_TH) void k_pchol(int n, const double2 *Gall, long ld, long gstride, double2 *Lall,
                                                 long ldl, long lstride, int *rank, double tolrel)
{
    __shared__ double2 lp[PC_TH];   // row p of L, columns 0 .. k-1
    __shared__ double wv_d[PC_TH / 64];
    __shared__ int wv_i[PC_TH / 64];
    __shared__ double s_d;
    __shared__ int s_p;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const long b = blockIdx.x;
    const double2 *G = Gall + b * gstride;
    double2 *L = Lall + b * lstride;
    const bool row = tid < n;
    double d = row ? G[tid + (long)tid * ld].x : 0.0;
    bool piv = false;
    // largest remaining diagonal (ties: the lower row) over the workgroup
    auto argmax = [&](double v, int &p) -> double {
        int idx = tid;
        for (int o = 32; o > 0; o >>= 1) {
            const double v2 = __shfl_xor(v, o);
            const int i2 = __shfl_xor(idx, o);
            if (v2 > v || (v2 == v && i2 < idx)) {
                v = v2;
                idx = i2;
            }
        }
        if (lane == 0) {
            wv_d[w] = v;
            wv_i[w] = idx;
        }
        __syncthreads();
        if (tid == 0) {
            double bv = wv_d[0];
            int bi = wv_i[0];
            for (int q = 1; q < PC_TH / 64; ++q)
                if (wv_d[q] > bv || (wv_d[q] == bv && wv_i[q] < bi)) {
                    bv = wv_d[q];
                    bi = wv_i[q];
                }
            s_d = bv;
            s_p = bi;
        }
        __syncthreads();
        p = s_p;
        return s_d;
    };
    int p0;
    const double tol = tolrel * fmax(argmax(row ? d : -1.0, p0), 0.0);
    int k = 0;
    for (; k < n; ++k) {
        int p;
        const double dp = argmax(row && !piv ? d : -1.0, p);
        if (!(dp > tol)) break;
        for (int j = tid; j < k; j += PC_TH) lp[j] = L[p + (long)j * ldl];
        __syncthreads();
        if (row) {
            double2 l;
            if (tid == p) {
                l = make_double2(sqrt(dp), 0.0);
                d = 0.0;
                piv = true;
            } else if (piv) {
                l = make_double2(0.0, 0.0);
            } else {
                double2 a = G[tid + (long)p * ld], a2 = make_double2(0.0, 0.0);
                const double2 *lr = L + tid;
                int j = 0;
                for (; j + 1 < k; j += 2) {
                    const double2 t = zcmul2(lp[j], lr[(long)j * ldl]);
                    const double2 t2 = zcmul2(lp[j + 1], lr[(long)(j + 1) * ldl]);
                    a.x -= t.x;
                    a.y -= t.y;
                    a2.x -= t2.x;
                    a2.y -= t2.y;
                }
                if (j < k) {
                    const double2 t = zcmul2(lp[j], lr[(long)j * ldl]);
                    a.x -= t.x;
                    a.y -= t.y;
                }
                const double is = 1.0 / sqrt(dp);
                l = make_double2((a.x + a2.x) * is, (a.y + a2.y) * is);
                d -= l.x * l.x + l.y * l.y;
            }
            L[tid + (long)k * ldl] = l;
        }
        __syncthreads();
    }
    for (long idx = tid; idx < (long)(n - k) * n; idx += PC_TH)
        L[(idx % n) + (k + idx / n) * ldl] = make_double2(0.0, 0.0);
    if (tid == 0) rank[b] = k;
}

int launch_pchol_batched(int n, const double2 *G, long ld, long stride, double2 *L, long ldl, long lstride, int *rank,
                         int batch, double tolrel, hipStream_t st)
{
    if (n <= 0 || batch <= 0) return 0;
    MIDAS_CHECK(n <= PC_TH && ld >= n && ldl >= n && stride >= ld * n && lstride >= ldl * n, -1,
                "pchol: bad layout (n %d, ld %ld, ldl %ld)", n, ld, ldl);
    hipLaunchKernelGGL(k_pchol, dim3(batch), dim3(PC_TH), 0, st, n, G, ld, stride, L, ldl, lstride, rank, tolrel);
    MIDAS_HIP(hipGetLastError());
    return 0;
}

constexpr int HEEV_SMALL_MAX = 256;   // orders the one-stage kernel takes (path 0)

static size_t heev_large_bytes(int n) { return (size_t)n * n * (sizeof(double2) + sizeof(double)) + (size_t)n * HL_NB * sizeof(double2) + (size_t)HL_WAVES * 6 * n * sizeof(double); }

size_t heev_work_bytes(int n, int batch)
{
    const size_t small = (size_t)n * n * (2 * sizeof(double2) + sizeof(double));
    return (size_t)batch * std::max(small, heev_large_bytes(n));
}

int launch_heev_batched(int n, double2 *G, long ld, long stride, double *D, long dstride, int *info, int batch,
                        void *work, hipStream_t st, double cut, int path)
{
    if (n <= 0 || batch <= 0) return 0;
    MIDAS_CHECK(ld >= n && stride >= ld * n && dstride >= n && work, -1,
                "heev: bad layout (n %d, ld %ld, stride %ld, dstride %ld)", n, ld, stride, dstride);
    if (path == 2 || (path == 0 && n > HEEV_SMALL_MAX)) {
        MIDAS_CHECK(n <= HL_TH, -1, "heev: order %d above %d", n, HL_TH);
        const size_t lds = (size_t)n * (4 * sizeof(double2) + 3 * sizeof(double) + sizeof(int)) + sizeof(int) +
                           2 * HL_NB * HL_CHUNK * sizeof(double2);
        if (lds > 64 * 1024)
            MIDAS_HIP(hipFuncSetAttribute((const void *)k_heev_large, hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)lds));
        const size_t nn = (size_t)n * n;
        double2 *UT = (double2 *)work, *Wp = UT + nn * batch;
        double *Zv = (double *)(Wp + (size_t)n * HL_NB * batch), *LU = Zv + nn * batch;
        hipLaunchKernelGGL(k_heev_large, dim3(batch), dim3(HL_TH), lds, st, n, G, ld, stride, D, dstride, info, cut,
                           Wp, UT, Zv, LU);
        MIDAS_HIP(hipGetLastError());
        return 0;
    }
    const size_t lds = (size_t)n * (4 * sizeof(double2) + 4 * sizeof(double) + sizeof(int));
    MIDAS_CHECK(lds <= 150 * 1024, -1, "heev: order %d too large for one workgroup", n);
    if (lds > 64 * 1024)
        MIDAS_HIP(hipFuncSetAttribute((const void *)k_heev, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const size_t nn = (size_t)n * n;
    double2 *V = (double2 *)work, *U = V + nn * batch;
    double *Z = (double *)(U + nn * batch);
    hipLaunchKernelGGL(k_heev, dim3(batch), dim3(HEEV_TH), lds, st, n, G, ld, stride, D, dstride, info, V, Z, U);
    MIDAS_HIP(hipGetLastError());
    return 0;
}

}  // namespace midas
