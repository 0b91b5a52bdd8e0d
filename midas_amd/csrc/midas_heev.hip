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

size_t heev_work_bytes(int n, int batch) { return (size_t)batch * n * n * (2 * sizeof(double2) + sizeof(double)); }

int launch_heev_batched(int n, double2 *G, long ld, long stride, double *D, long dstride, int *info, int batch,
                        void *work, hipStream_t st)
{
    if (n <= 0 || batch <= 0) return 0;
    MIDAS_CHECK(ld >= n && stride >= ld * n && dstride >= n && work, -1,
                "heev: bad layout (n %d, ld %ld, stride %ld, dstride %ld)", n, ld, stride, dstride);
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
