// Orientation-averaged scattering quantities over theta x phi grids (host code):
// the numbers of Compute_Scattering_Quantities_1
// (Compute_Scattering_Quantities.f90:1-589), used when the transmitters are
// integrated by 'un' (adaptive quadrature on a uniform grid), 'gl'
// (Gauss-Legendre), 'tr' (trapezoid cubature) or 'sm'.
//
//   qk61_grid, qage_grid  the reference's QUADPACK qag / qk61 (quadpack.f90),
//                         single precision, 61-point Gauss-Kronrod, whose
//                         integrand is read from a grid at the nearest node
//                         (getAdaptiveQuadPts2, Numerical_Integration.f90:213-245)
//   trap_2dc              trap_2Dc (Numerical_Integration.f90:251-305)
//   gauss_legendre        the rule of getGaussLegendreQuadPts3 (nodes in
//                         descending order), computed by Newton iteration
//
// Everything here is O(receivers x transmitters) on a few thousand values, so
// it runs on the host after S has been computed on the GPU.
#include "midas_common.h"
#include "midas_host.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#pragma clang fp contract(off)   // separate multiplies and adds, as the reference's compiled code

using midas::host_pi;

namespace {

// 61-point Gauss-Kronrod rule (QUADPACK qk61, Piessens et al. 1983): Kronrod
// abscissae (the even ones are the 30-point Gauss abscissae), Kronrod weights,
// Gauss weights; the reference holds them as single-precision literals.
const float XGK[31] = {
    0.999484410050490637571325895705811f, 0.996893484074649540271630050918695f, 0.991630996870404594858628366109486f,
    0.983668123279747209970032581605663f, 0.973116322501126268374693868423707f, 0.960021864968307512216871025581798f,
    0.944374444748559979415831324037439f, 0.926200047429274325879324277080474f, 0.905573307699907798546522558925958f,
    0.882560535792052681543116462530226f, 0.857205233546061098958658510658944f, 0.829565762382768397442898119732502f,
    0.799727835821839083013668942322683f, 0.767777432104826194917977340974503f, 0.733790062453226804726171131369528f,
    0.697850494793315796932292388026640f, 0.660061064126626961370053668149271f, 0.620526182989242861140477556431189f,
    0.579345235826361691756024932172540f, 0.536624148142019899264169793311073f, 0.492480467861778574993693061207709f,
    0.447033769538089176780609900322854f, 0.400401254830394392535476211542661f, 0.352704725530878113471037207089374f,
    0.304073202273625077372677107199257f, 0.254636926167889846439805129817805f, 0.204525116682309891438957671002025f,
    0.153869913608583546963794672743256f, 0.102806937966737030147096751318001f, 0.051471842555317695833025213166723f,
    0.0f};
const float WGK[31] = {
    0.001389013698677007624551591226760f, 0.003890461127099884051267201844516f, 0.006630703915931292173319826369750f,
    0.009273279659517763428441146892024f, 0.011823015253496341742232898853251f, 0.014369729507045804812451432443580f,
    0.016920889189053272627572289420322f, 0.019414141193942381173408951050128f, 0.021828035821609192297167485738339f,
    0.024191162078080601365686370725232f, 0.026509954882333101610601709335075f, 0.028754048765041292843978785354334f,
    0.030907257562387762472884252943092f, 0.032981447057483726031814191016854f, 0.034979338028060024137499670731468f,
    0.036882364651821229223911065617136f, 0.038678945624727592950348651532281f, 0.040374538951535959111995279752468f,
    0.041969810215164246147147541285970f, 0.043452539701356069316831728117073f, 0.044814800133162663192355551616723f,
    0.046059238271006988116271735559374f, 0.047185546569299153945261478181099f, 0.048185861757087129140779492298305f,
    0.049055434555029778887528165367238f, 0.049795683427074206357811569379942f, 0.050405921402782346840893085653585f,
    0.050881795898749606492297473049805f, 0.051221547849258772170656282604944f, 0.051426128537459025933862879215781f,
    0.051494729429451567558340433647099f};
const float WG[15] = {
    0.007968192496166605615465883474674f, 0.018466468311090959142302131912047f, 0.028784707883323369349719179611292f,
    0.038799192569627049596801936446348f, 0.048402672830594052902938140422808f, 0.057493156217619066481721689402056f,
    0.065974229882180495128128515115962f, 0.073755974737705206268243850022191f, 0.080755895229420215354694938460530f,
    0.086899787201082979802387530715126f, 0.092122522237786128717632707087619f, 0.096368737174644259639468626351810f,
    0.099593420586795267062780282103569f, 0.101762389748405504596428952168554f, 0.102852652893558840341285636705415f};

// Fortran NINT of a single-precision value
int nint_f(float v) { return v >= 0.0f ? (int)std::floor(v + 0.5f) : -(int)std::floor(-v + 0.5f); }

// An integrand sampled on a grid: column jj (1-based) of f (d1 x d2,
// column-major); a point x maps to the node nint(x (n - 1) / (b0 - a0)) + 1
// (the reference's mapping, exact for grids that start at 0).
struct GridFn {
    const double *f;
    int d1, d2, jj, n;   // n = nodes along the integration variable
    float a0, b0;
    bool bad = false;    // a point mapped outside the grid
    float at(float x)
    {
        const int i = nint_f(x * (float)(n - 1) / (b0 - a0)) + 1;
        if (i < 1 || i > d1) {
            bad = true;
            return 0.0f;
        }
        return (float)f[(i - 1) + (long)(jj - 1) * d1];
    }
};

// qk61 over [a, b] (quadpack.f90:486-725): result, abserr, resabs, resasc
void qk61_grid(GridFn &F, float a, float b, float &result, float &abserr, float &resabs, float &resasc)
{
    const float centr = 0.5f * (b + a), hlgth = 0.5f * (b - a), dhlgth = std::fabs(hlgth);
    float resg = 0.0f;
    const float fc = F.at(centr);
    float resk = WGK[30] * fc;
    resabs = std::fabs(resk);
    float fv1[30], fv2[30];
    for (int j = 1; j <= 15; ++j) {
        const int jtw = 2 * j;
        const float absc = hlgth * XGK[jtw - 1];
        const float f1 = F.at(centr - absc), f2 = F.at(centr + absc);
        fv1[jtw - 1] = f1;
        fv2[jtw - 1] = f2;
        const float fsum = f1 + f2;
        resg = resg + WG[j - 1] * fsum;
        resk = resk + WGK[jtw - 1] * fsum;
        resabs = resabs + WGK[jtw - 1] * (std::fabs(f1) + std::fabs(f2));
    }
    for (int j = 1; j <= 15; ++j) {
        const int jtwm1 = 2 * j - 1;
        const float absc = hlgth * XGK[jtwm1 - 1];
        const float f1 = F.at(centr - absc), f2 = F.at(centr + absc);
        fv1[jtwm1 - 1] = f1;
        fv2[jtwm1 - 1] = f2;
        const float fsum = f1 + f2;
        resk = resk + WGK[jtwm1 - 1] * fsum;
        resabs = resabs + WGK[jtwm1 - 1] * (std::fabs(f1) + std::fabs(f2));
    }
    const float reskh = resk * 0.5f;
    resasc = WGK[30] * std::fabs(fc - reskh);
    for (int j = 0; j < 30; ++j) resasc = resasc + WGK[j] * (std::fabs(fv1[j] - reskh) + std::fabs(fv2[j] - reskh));
    result = resk * hlgth;
    resabs = resabs * dhlgth;
    resasc = resasc * dhlgth;
    abserr = std::fabs((resk - resg) * hlgth);
    if (resasc != 0.0f && abserr != 0.0f) abserr = resasc * std::min(1.0f, std::pow(200.0f * abserr / resasc, 1.5f));
    if (resabs > FLT_MIN / (50.0f * FLT_EPSILON)) abserr = std::max((FLT_EPSILON * 50.0f) * resabs, abserr);
}

// qsort of quadpack.f90 (QUADPACK qpsrt): keeps iord descending by error;
// 1-based indices as the reference
void qpsrt(int limit, int last, int &maxerr, float &ermax, const float *elist, int *iord, int &nrmax)
{
    auto E = [&](int i) { return elist[i - 1]; };
    auto O = [&](int i) -> int & { return iord[i - 1]; };
    if (last <= 2) {
        O(1) = 1;
        O(2) = 2;
    } else {
        const float errmax = E(maxerr);
        for (int i = 1; i <= nrmax - 1; ++i) {
            const int isucc = O(nrmax - 1);
            if (errmax <= E(isucc)) break;
            O(nrmax) = isucc;
            nrmax = nrmax - 1;
        }
        int jupbn = last;
        if (limit / 2 + 2 < last) jupbn = limit + 3 - last;
        const float errmin = E(last);
        const int jbnd = jupbn - 1, ibeg = nrmax + 1;
        int i = ibeg;
        bool found = false;
        for (; i <= jbnd; ++i) {
            const int isucc = O(i);
            if (E(isucc) <= errmax) {
                found = true;
                break;
            }
            O(i - 1) = isucc;
        }
        if (!found) {
            O(jbnd) = maxerr;
            O(jupbn) = last;
        } else {
            O(i - 1) = maxerr;
            int k = jbnd;
            bool placed = false;
            for (int j = i; j <= jbnd; ++j) {
                const int isucc = O(k);
                if (errmin < E(isucc)) {
                    O(k + 1) = last;
                    placed = true;
                    break;
                }
                O(k + 1) = isucc;
                k = k - 1;
            }
            if (!placed) O(i) = last;
        }
    }
    maxerr = O(nrmax);
    ermax = E(maxerr);
}

// qag with key 6, epsabs 0, epsrel 1e-3, limit 500 (quadpack.f90:1-425),
// single precision throughout
float qage_grid(GridFn &F, float a, float b)
{
    constexpr int LIMIT = 500;
    const float epsabs = 0.0f, epsrel = 0.001f, c = 6.0f;
    std::vector<float> alist(LIMIT), blist(LIMIT), rlist(LIMIT), elist(LIMIT);
    std::vector<int> iord(LIMIT);
    int ier = 0;
    alist[0] = a;
    blist[0] = b;
    float result, abserr, defabs, resabs;
    qk61_grid(F, a, b, result, abserr, defabs, resabs);
    int last = 1;
    rlist[0] = result;
    elist[0] = abserr;
    iord[0] = 1;
    float errbnd = std::max(epsabs, epsrel * std::fabs(result));
    if (abserr <= 50.0f * FLT_EPSILON * defabs && errbnd < abserr) ier = 2;
    if (ier != 0 || (abserr <= errbnd && abserr != resabs) || abserr == 0.0f) return result;
    float errmax = abserr, area = result, errsum = abserr;
    int maxerr = 1, nrmax = 1, iroff1 = 0, iroff2 = 0;
    for (last = 2; last <= LIMIT; ++last) {
        const float a1 = alist[maxerr - 1], b1 = 0.5f * (alist[maxerr - 1] + blist[maxerr - 1]), a2 = b1,
                    b2 = blist[maxerr - 1];
        float area1, error1, defab1, area2, error2, defab2, ra;
        qk61_grid(F, a1, b1, area1, error1, ra, defab1);
        qk61_grid(F, a2, b2, area2, error2, ra, defab2);
        const float area12 = area1 + area2, erro12 = error1 + error2;
        errsum = errsum + erro12 - errmax;
        area = area + area12 - rlist[maxerr - 1];
        if (defab1 != error1 && defab2 != error2) {
            if (std::fabs(rlist[maxerr - 1] - area12) <= 1.0e-5f * std::fabs(area12) && 0.99f * errmax <= erro12)
                iroff1 = iroff1 + 1;
            if (10 < last && errmax < erro12) iroff2 = iroff2 + 1;
        }
        rlist[maxerr - 1] = area1;
        rlist[last - 1] = area2;
        errbnd = std::max(epsabs, epsrel * std::fabs(area));
        if (errbnd < errsum) {
            if (6 <= iroff1 || 20 <= iroff2) ier = 2;
            if (last == LIMIT) ier = 1;
            if (std::max(std::fabs(a1), std::fabs(b2)) <=
                (1.0f + c * 1.0e3f * FLT_EPSILON) * (std::fabs(a2) + 1.0e4f * FLT_MIN))
                ier = 3;
        }
        if (error2 <= error1) {
            alist[last - 1] = a2;
            blist[maxerr - 1] = b1;
            blist[last - 1] = b2;
            elist[maxerr - 1] = error1;
            elist[last - 1] = error2;
        } else {
            alist[maxerr - 1] = a2;
            alist[last - 1] = a1;
            blist[last - 1] = b1;
            rlist[maxerr - 1] = area2;
            rlist[last - 1] = area1;
            elist[maxerr - 1] = error2;
            elist[last - 1] = error1;
        }
        qpsrt(LIMIT, last, maxerr, errmax, elist.data(), iord.data(), nrmax);
        if (ier != 0 || errsum <= errbnd) break;
    }
    float s = 0.0f;
    for (int i = 0; i < last; ++i) s = s + rlist[i];
    return s;
}

// getAdaptiveQuadPts2: integral over [a0, b0] of column jj of f (d1 x d2);
// 'th' grids have nth nodes, 'ph' grids nph
int adaptive_quad(int nth, int nph, int d1, int d2, const double *f, bool ph, int jj, double a0, double b0, double &I)
{
    GridFn F{f, d1, d2, jj, ph ? nph : nth, (float)a0, (float)b0};
    const float r = qage_grid(F, F.a0, F.b0);
    MIDAS_CHECK(!F.bad, -1, "scattering quantities: a quadrature point falls outside the grid (range %g..%g)", a0, b0);
    I = (double)r;
    return 0;
}

// trap_2Dc: cubature trapezoid rule over nx x ny values
double trap_2dc(const double *f, double a, double b, double c, double d, int nx, int ny)
{
    auto F = [&](int i, int j) { return f[(i - 1) + (long)(j - 1) * nx]; };
    const double hx = (b - a) / (double)(nx - 1), hy = (d - c) / (double)(ny - 1);
    double sum = F(1, 1) + F(1, ny) + F(nx, 1) + F(nx, ny);
    for (int i = 2; i <= nx - 1; ++i) sum = sum + 2.0 * (F(i, 1) + F(i, ny));
    for (int j = 2; j <= ny - 1; ++j) sum = sum + 2.0 * (F(1, j) + F(nx, j));
    for (int i = 2; i <= nx - 1; ++i)
        for (int j = 2; j <= ny - 1; ++j) sum = sum + 4.0 * F(i, j);
    return 0.25 * hx * hy * sum;
}

// n-point Gauss-Legendre rule on [a, b], nodes in descending order (as the
// reference's getGaussLegendreQuadPts3 returns them): Newton iteration on
// P_n (three-term recurrence) from cos(pi (i + 3/4) / (n + 1/2)); weights
// 2 / ((1 - z^2) P_n'(z)^2) scaled to [a, b].  The rule is symmetric, so the
// weights need no reordering.
void gauss_legendre(int n, double a, double b, double *x, double *w)
{
    const double pi = MIDAS_PI_EXACT;
    for (int i = 0; i < (n + 1) / 2; ++i) {
        double z = std::cos(pi * (i + 0.75) / (n + 0.5)), pp = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= n; ++j) {
                const double p3 = p2;
                p2 = p1;
                p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            const double z1 = z;
            z = z1 - p1 / pp;
            if (std::fabs(z - z1) < 1e-15) break;
        }
        {   // P_n' at the converged node
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= n; ++j) {
                const double p3 = p2;
                p2 = p1;
                p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
        }
        const double wt = 2.0 / ((1.0 - z * z) * pp * pp);
        x[i] = (a * (1.0 - z) + b * (1.0 + z)) / 2.0;
        x[n - 1 - i] = (a * (1.0 + z) + b * (1.0 - z)) / 2.0;
        w[i] = w[n - 1 - i] = wt * (b - a) / 2.0;
    }
}

// unique values in order of first appearance (Unique1DArray_D)
std::vector<double> unique_first(const double *v, int n)
{
    std::vector<double> u;
    for (int i = 0; i < n; ++i)
        if (std::find(u.begin(), u.end(), v[i]) == u.end()) u.push_back(v[i]);
    return u;
}

// getTxRxIndex (get_trans_Receiv.f90:690-713): first receiver within 1e-2
// degree (single-precision literal) of (th, ph); -1 when none
int txrx_index(int n, const double *th, const double *ph, double t, double p)
{
    const double eps = (double)1e-2f;
    for (int i = 0; i < n; ++i)
        if (std::fabs(th[i] - t) <= eps && std::fabs(ph[i] - p) <= eps) return i;
    return -1;
}

}  // namespace

extern "C" int midas_quad_adaptive(int nth, int nph, int d1, int d2, const double *f, int ph, int jj, double a0,
                                   double b0, double *I)
{
    MIDAS_CHECK(f && I && d1 > 0 && d2 > 0 && jj >= 1 && jj <= d2 && nth > 0 && nph > 0, -1, "quad_adaptive: bad arguments");
    return adaptive_quad(nth, nph, d1, d2, f, ph != 0, jj, a0, b0, *I);
}

extern "C" int midas_quad_trap2d(int nx, int ny, const double *f, double a, double b, double c, double d, double *I)
{
    MIDAS_CHECK(f && I && nx > 0 && ny > 0, -1, "quad_trap2d: bad arguments");
    *I = trap_2dc(f, a, b, c, d, nx, ny);
    return 0;
}

extern "C" int midas_quad_gauss_legendre(int n, double a, double b, double *x, double *w)
{
    MIDAS_CHECK(x && w && n > 0, -1, "quad_gauss_legendre: bad arguments");
    gauss_legendre(n, a, b, x, w);
    return 0;
}

extern "C" int midas_scattering_quantities_grid(const double *S, long lds, int nrx_tot, const double *rx_th,
                                                const double *rx_ph, int nrx, int rx_type, const double *rx_w,
                                                int ntr, const double *tr_th, const double *tr_ph, int tr_type,
                                                double k0, double dm, int quirks, const double *C_ext,
                                                const double *C_abs, int cext_intfields, double *Q, double *Q_av)
{
    MIDAS_CHECK(S && rx_th && rx_ph && tr_th && tr_ph && C_ext && C_abs && Q && Q_av, -1,
                "scattering_quantities_grid: bad arguments");
    MIDAS_CHECK(nrx > 0 && ntr > 0 && nrx_tot >= nrx && lds >= nrx_tot && k0 > 0 && dm > 0, -1,
                "scattering_quantities_grid: nrx %d, ntr %d, nrx_tot %d, lds %ld", nrx, ntr, nrx_tot, lds);
    MIDAS_CHECK(rx_type >= MIDAS_INT_SD && rx_type <= MIDAS_INT_SM && tr_type >= MIDAS_INT_UN &&
                    tr_type <= MIDAS_INT_SM,
                -1, "scattering_quantities_grid: receiver / transmitter rule %d / %d", rx_type, tr_type);
    MIDAS_CHECK(rx_type != MIDAS_INT_LB || rx_w, -1, "scattering_quantities_grid: lb receivers need weights");
    const double pi = host_pi(quirks), ap = dm / 2.0, X = k0 * dm / 2.0;
    const bool rx1d = rx_type == MIDAS_INT_SD || rx_type == MIDAS_INT_LB;
    auto sq = [](const double *z) {
        const double m = std::hypot(z[0], z[1]);
        return m * m;
    };
    auto s_at = [&](int r, int c) { return S + 2 * ((long)r + (long)c * lds); };

    // the transmitter grid (theta fastest) and its ranges
    const std::vector<double> tth = unique_first(tr_th, ntr), tph = unique_first(tr_ph, ntr);
    const int ntt = (int)tth.size(), ntp = (int)tph.size();
    MIDAS_CHECK(ntt * ntp == ntr, -1, "scattering_quantities_grid: %d transmitters are not a %d x %d grid", ntr,
                ntt, ntp);
    const double trth_min = tth[0] * pi / 180., trth_max = tth[ntt - 1] * pi / 180.;
    const double trph_min = tph[0] * pi / 180., trph_max = tph[ntp - 1] * pi / 180.;
    // the receiver grid for the 2-D rules
    std::vector<double> rth, rph;
    int nrt = 0, nrp = 0;
    double rth_min = 0, rth_max = 0, rph_min = 0, rph_max = 0, sn_r = 0;
    if (!rx1d) {
        rth = unique_first(rx_th, nrx);
        rph = unique_first(rx_ph, nrx);
        nrt = (int)rth.size();
        nrp = (int)rph.size();
        MIDAS_CHECK(nrt * nrp == nrx, -1, "scattering_quantities_grid: %d receivers are not a %d x %d grid", nrx, nrt,
                    nrp);
        rth_min = rth[0] * pi / 180.;
        rth_max = rth[nrt - 1] * pi / 180.;
        rph_min = rph[0] * pi / 180.;
        rph_max = rph[nrp - 1] * pi / 180.;
        sn_r = std::max(rph_max - rph_min, 1.0) * (std::cos(rth_min) - std::cos(rth_max));
    }
    // per transmitter (Compute_Scattering_Quantities.f90:176-388)
    std::vector<double> fxy(rx1d ? 0 : (size_t)nrx), fxyg(rx1d ? 0 : (size_t)nrx), fy(nrp), fyg(nrp);
    std::vector<double> xth, wth, xph, wph;
    if (rx_type == MIDAS_INT_GL) {
        xth.resize(nrt);
        wth.resize(nrt);
        xph.resize(nrp);
        wph.resize(nrp);
        gauss_legendre(nrt, rth_min, rth_max, xth.data(), wth.data());
        if (nrp > 1) gauss_legendre(nrp, rph_min, rph_max, xph.data(), wph.data());
        else wph[0] = 1.;
        std::reverse(wth.begin(), wth.end());
        std::reverse(wph.begin(), wph.end());
    }
    for (int t = 0; t < ntr; ++t) {
        const double th_i = tr_th[t], ph_i = tr_ph[t];
        const double kix = -std::sin(th_i * pi / 180.) * std::cos(ph_i * pi / 180.);
        const double kiy = -std::sin(th_i * pi / 180.) * std::sin(ph_i * pi / 180.);
        const double kiz = -std::cos(th_i * pi / 180.);
        const int ifw = txrx_index(nrx_tot, rx_th, rx_ph, th_i, ph_i);
        MIDAS_CHECK(ifw >= 0, -2, "scattering_quantities_grid: forward direction of transmitter %d not found", t);
        const double q_ext = (2 / (X * X)) * (std::fabs(s_at(ifw, 4 * t + 3)[1]) + std::fabs(s_at(ifw, 4 * t)[1]));
        const double q_ext_intf = C_ext[2 * t] / (pi * ap * ap);
        double q_sca = 0.0, g = 0.0;
        auto s_it = [&](int r) {
            return ((sq(s_at(r, 4 * t + 3)) + sq(s_at(r, 4 * t))) + sq(s_at(r, 4 * t + 1))) + sq(s_at(r, 4 * t + 2));
        };
        auto cg_at = [&](double th_s, double ph_s) {
            const double ksx = -std::sin(th_s * pi / 180.) * std::cos(ph_s * pi / 180.);
            const double ksy = -std::sin(th_s * pi / 180.) * std::sin(ph_s * pi / 180.);
            const double ksz = -std::cos(th_s * pi / 180.);
            return kix * ksx + kiy * ksy + kiz * ksz;
        };
        if (rx1d) {
            double sf = 0.0, sg = 0.0;
            for (int r = 0; r < nrx; ++r) {
                const double f = rx_type == MIDAS_INT_LB ? rx_w[r] * s_it(r) : s_it(r);
                sf = sf + f;
                sg = sg + f * cg_at(rx_th[r], rx_ph[r]);
            }
            if (rx_type == MIDAS_INT_LB) {
                const double c_sca = 1. / (2 * k0 * k0) * sf;
                q_sca = (4 * pi) / (2 * pi * X * X) * sf;
                g = 1. / (2. * k0 * k0 * c_sca) * sg;
            } else {
                const double q_sd = (4 * pi / nrx) * sf;
                const double c_sca = 1. / (2 * k0 * k0) * q_sd;
                q_sca = 1 / (2 * pi * X * X) * q_sd;
                g = 1. / (2. * k0 * k0 * c_sca) * (4 * pi / nrx) * sg;
            }
        } else {
            // integrands over the receiver grid: |S|^2 sin(theta), times cos(k_i, k_s) for g
            for (int ii = 0; ii < nrt; ++ii) {
                const double s = std::sin(rth[ii] * pi / 180.);
                for (int jj = 0; jj < nrp; ++jj) {
                    const int r = jj * nrt + ii;
                    const double v = s_it(r);
                    fxy[ii + (long)jj * nrt] = v * s;
                    fxyg[ii + (long)jj * nrt] = v * s * cg_at(rth[ii], rph[jj]);
                }
            }
            double qi = 0.0, qg = 0.0;
            if (rx_type == MIDAS_INT_GL) {
                for (int ii = 0; ii < nrt; ++ii)
                    for (int jj = 0; jj < nrp; ++jj) qi = qi + wph[jj] * wth[ii] * fxy[ii + (long)jj * nrt];
                for (int ii = 0; ii < nrt; ++ii)
                    for (int jj = 0; jj < nrp; ++jj) qg = qg + wph[jj] * wth[ii] * fxyg[ii + (long)jj * nrt];
            } else if (rx_type == MIDAS_INT_UN) {
                for (int jj = 1; jj <= nrp; ++jj) TRY(adaptive_quad(nrt, nrp, nrt, nrp, fxy.data(), false, jj, rth_min, rth_max, fy[jj - 1]));
                if (nrp > 1) TRY(adaptive_quad(nrt, nrp, nrp, 1, fy.data(), true, 1, rph_min, rph_max, qi));
                else qi = fy[0];
                for (int jj = 1; jj <= nrp; ++jj) TRY(adaptive_quad(nrt, nrp, nrt, nrp, fxyg.data(), false, jj, rth_min, rth_max, fyg[jj - 1]));
                if (nrp > 1) TRY(adaptive_quad(nrt, nrp, nrp, 1, fyg.data(), true, 1, rph_min, rph_max, qg));
                else qg = fyg[0];
            } else if (rx_type == MIDAS_INT_TR) {
                qi = trap_2dc(fxy.data(), rth_min, rth_max, rph_min, rph_max, nrt, nrp);
                qg = trap_2dc(fxyg.data(), rth_min, rth_max, rph_min, rph_max, nrt, nrp);
            }
            // 'sm': no rule in the reference, Q_sca and g stay 0
            if (rx_type != MIDAS_INT_SM) {
                const double c_sca = 1. / (2 * k0 * k0) * qi;
                q_sca = 2. / (sn_r * X * X) * qi;
                g = 1. / (2. * k0 * k0 * c_sca) * qg;
            }
        }
        const int ibk = txrx_index(nrx_tot, rx_th, rx_ph, 180. - th_i, std::fmod(ph_i + 180., 360.));
        MIDAS_CHECK(ibk >= 0, -2, "scattering_quantities_grid: backward direction of transmitter %d not found", t);
        const double q_bks = (1. / (2. * pi * X * X)) *
                             (((sq(s_at(ibk, 4 * t)) + sq(s_at(ibk, 4 * t + 3))) + sq(s_at(ibk, 4 * t + 1))) +
                              sq(s_at(ibk, 4 * t + 2)));
        double *q = Q + 5L * t;
        q[0] = cext_intfields ? q_ext_intf : q_ext;
        q[1] = C_abs[2 * t] / (pi * ap * ap);   // Q_abs: always the internal-field value (:396)
        q[2] = q_sca;
        q[3] = q_bks;
        q[4] = g;
    }
    // averages over the transmitter grid (:432-519)
    if (ntr == 1) {
        for (int j = 0; j < 5; ++j) Q_av[j] = Q[j];
        return 0;
    }
    const double sn_t = std::max(trph_max - trph_min, 1.0) * (std::cos(trth_min) - std::cos(trth_max));
    std::vector<double> G(5L * ntr);   // (theta, phi) grid per quantity, weighted by sin(theta)
    for (int ii = 0; ii < ntt; ++ii) {
        const double s = std::sin(tr_th[ii] * pi / 180.);
        for (int jj = 0; jj < ntp; ++jj)
            for (int j = 0; j < 5; ++j) G[j * (long)ntr + ii + (long)jj * ntt] = Q[5L * (jj * ntt + ii) + j] * s;
    }
    double integ[5] = {0, 0, 0, 0, 0};
    if (tr_type == MIDAS_INT_GL) {
        std::vector<double> x1(ntt), w1(ntt), x2(ntp), w2(ntp);
        gauss_legendre(ntt, trth_min, trth_max, x1.data(), w1.data());
        if (ntp > 1) gauss_legendre(ntp, trph_min, trph_max, x2.data(), w2.data());
        else w2[0] = 1.;
        std::reverse(w1.begin(), w1.end());
        std::reverse(w2.begin(), w2.end());
        for (int ii = 0; ii < ntt; ++ii)
            for (int jj = 0; jj < ntp; ++jj)
                for (int j = 0; j < 5; ++j) integ[j] = integ[j] + w2[jj] * w1[ii] * G[j * (long)ntr + ii + (long)jj * ntt];
    } else if (tr_type == MIDAS_INT_UN) {
        // order of the reference: ext, abs, sca, bks, g -> Q indices 0, 1, 2, 3, 4
        std::vector<double> fyq(ntp);
        for (int j = 0; j < 5; ++j) {
            for (int jj = 1; jj <= ntp; ++jj)
                TRY(adaptive_quad(ntt, ntp, ntt, ntp, G.data() + j * (long)ntr, false, jj, trth_min, trth_max,
                                  fyq[jj - 1]));
            if (ntp > 1) TRY(adaptive_quad(ntt, ntp, ntp, 1, fyq.data(), true, 1, trph_min, trph_max, integ[j]));
            else integ[j] = fyq[0];
        }
    } else if (tr_type == MIDAS_INT_TR) {
        // the reference passes the lower phi bound twice (:497-501)
        for (int j = 0; j < 5; ++j)
            integ[j] = trap_2dc(G.data() + j * (long)ntr, trth_min, trth_max, trph_min, trph_min, ntt, ntp);
    }
    for (int j = 0; j < 5; ++j) Q_av[j] = 1. / sn_t * integ[j];
    return 0;
}
