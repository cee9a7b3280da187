/*
 * a12_oracle.c -- CPU restatement of the ATLAS12/SYNTHE hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  See
 * a12_oracle.h.  Built with -O2 -ffp-contract=off so FP32/FP64 expression
 * types follow the reference source one operation at a time.
 *
 * Citations are cconroy20/atlas12 file:line.
 */
#include "a12_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define KW_MAX 128

/* REAL(8)**INTEGER with a run-time exponent as the compiled reference
 * evaluates it: repeated squaring, the bits of the exponent taken from the
 * lowest (a constant exponent is expanded into left-to-right products
 * instead; the two agree up to the cube). */
static double powi_f64(double x, int m)
{
    unsigned n = m < 0 ? -m : m;
    double y = (n & 1u) ? x : 1.0;
    while (n >>= 1) {
        x = x * x;
        if (n & 1u) y = y * x;
    }
    return m < 0 ? 1.0 / y : y;
}

/* mod_constants, src/atlas12_modules.f90:315-355 */
static const double C_SQRTPI = 1.7724538509055159;
static const double C_CLIGHT_NMS = 2.99792458e17;
static const double C_HPLANCK = 6.62607015e-27;
static const double C_CLIGHT = 2.99792458e10;

/* ===================================================================== */
/* Numerics                                                               */
/* ===================================================================== */

/* DERIV, src/atlas12_modules.f90:5228-5266 (tangent half-angle average). */
void or_deriv(const double *x, const double *f, double *dfdx, int n)
{
    dfdx[0] = (f[1] - f[0]) / (x[1] - x[0]);
    dfdx[n - 1] = (f[n - 1] - f[n - 2]) / (x[n - 1] - x[n - 2]);
    if (n == 2) return;
    double s = fabs(x[1] - x[0]) / (x[1] - x[0]);
    for (int j = 1; j < n - 1; ++j) {
        double scale = fmax(fmax(fabs(f[j - 1]), fabs(f[j])), fabs(f[j + 1])) / fabs(x[j]);
        if (scale == 0.0) scale = 1.0;
        double dr = (f[j + 1] - f[j]) / (x[j + 1] - x[j]) / scale;
        double dl = (f[j] - f[j - 1]) / (x[j] - x[j - 1]) / scale;
        double tr = dr / (s * sqrt(1.0 + dr * dr) + 1.0);
        double tl = dl / (s * sqrt(1.0 + dl * dl) + 1.0);
        dfdx[j] = (tr + tl) / (1.0 - tr * tl) * scale;
    }
}

/* PARCOE, src/atlas12_modules.f90:5358-5434 (blended piecewise parabolas;
 * linear at intervals 1, 2, 3 and N-1). */
void or_parcoe(const double *f, const double *x, double *a, double *b,
               double *c, int n)
{
    c[0] = 0.0;
    b[0] = (f[1] - f[0]) / (x[1] - x[0]);
    a[0] = f[0] - x[0] * b[0];
    c[n - 1] = 0.0;
    b[n - 1] = (f[n - 1] - f[n - 2]) / (x[n - 1] - x[n - 2]);
    a[n - 1] = f[n - 1] - x[n - 1] * b[n - 1];
    if (n == 2) return;
    if (n == 3) {
        a[1] = a[2]; b[1] = b[2]; c[1] = c[2];
        return;
    }
    for (int j = 1; j < n - 1; ++j) {
        double d = (f[j] - f[j - 1]) / (x[j] - x[j - 1]);
        c[j] = f[j + 1] / ((x[j + 1] - x[j]) * (x[j + 1] - x[j - 1]))
             - f[j] / ((x[j] - x[j - 1]) * (x[j + 1] - x[j]))
             + f[j - 1] / ((x[j] - x[j - 1]) * (x[j + 1] - x[j - 1]));
        b[j] = d - (x[j] + x[j - 1]) * c[j];
        a[j] = f[j - 1] - x[j - 1] * d + x[j] * x[j - 1] * c[j];
    }
    c[1] = 0.0;
    b[1] = (f[2] - f[1]) / (x[2] - x[1]);
    a[1] = f[1] - x[1] * b[1];
    c[2] = 0.0;
    b[2] = (f[3] - f[2]) / (x[3] - x[2]);
    a[2] = f[2] - x[2] * b[2];
    for (int j = 1; j < n - 1; ++j) {
        if (c[j] != 0.0) {
            double wt = fabs(c[j + 1]) / (fabs(c[j + 1]) + fabs(c[j]));
            a[j] = a[j + 1] + wt * (a[j] - a[j + 1]);
            b[j] = b[j + 1] + wt * (b[j] - b[j + 1]);
            c[j] = c[j + 1] + wt * (c[j] - c[j + 1]);
        }
    }
    a[n - 2] = a[n - 1];
    b[n - 2] = b[n - 1];
    c[n - 2] = c[n - 1];
}

/* INTEG with IQUAD = 0, src/atlas12_modules.f90:5293-5324. */
void or_integ(const double *x, const double *f, double *fint, int n,
              double start)
{
    double a[KW_MAX], b[KW_MAX], c[KW_MAX];
    or_parcoe(f, x, a, b, c, n);
    fint[0] = start;
    for (int i = 0; i < n - 1; ++i) {
        double xlo = x[i], xhi = x[i + 1];
        fint[i + 1] = fint[i] + (a[i] + b[i] * 0.5 * (xhi + xlo)
                      + c[i] / 3.0 * ((xhi + xlo) * xhi + xlo * xlo)) * (xhi - xlo);
    }
}

/* MAP1, src/atlas12_modules.f90:5570-5692.  Returns LL-1 as a Fortran
 * (1-based) interval index, exactly as the reference. */
int or_map1(const double *xold, const double *fold, int nold,
            const double *xnew, double *fnew, int nnew)
{
    /* Work in Fortran 1-based indices to keep the bracket logic literal. */
    const double *XO = xold - 1, *FO = fold - 1;
    double A = 0, B = 0, C = 0, ABAC = 0, BBAC = 0, CBAC = 0;
    double AFOR = 0, BFOR = 0, CFOR = 0, D, WT;
    int L = 2, LL = 0;
    for (int k = 0; k < nnew; ++k) {
        double xk = xnew[k];
        int past = 0;
        while (xk >= XO[L]) {
            L = L + 1;
            if (L > nold) { past = 1; break; }
        }
        if (!past && L == LL) {
            fnew[k] = A + (B + C * xk) * xk;
            continue;
        }
        if (L <= 3 || past) {
            L = (nold < L) ? nold : L;
            if (!past && L == LL) {
                fnew[k] = A + (B + C * xk) * xk;
                continue;
            }
            C = 0.0;
            B = (FO[L] - FO[L - 1]) / (XO[L] - XO[L - 1]);
            A = FO[L] - XO[L] * B;
            LL = L;
            fnew[k] = A + B * xk;
            continue;
        }
        int L1 = L - 1;
        if (L == LL + 1 && L != 3 && L != 4) {
            CBAC = CFOR; BBAC = BFOR; ABAC = AFOR;
        } else {
            int L2 = L - 2;
            D = (FO[L1] - FO[L2]) / (XO[L1] - XO[L2]);
            CBAC = FO[L] / ((XO[L] - XO[L1]) * (XO[L] - XO[L2]))
                 + (FO[L2] / (XO[L] - XO[L2]) - FO[L1] / (XO[L] - XO[L1]))
                 / (XO[L1] - XO[L2]);
            BBAC = D - (XO[L1] + XO[L2]) * CBAC;
            ABAC = FO[L2] - XO[L2] * D + XO[L1] * XO[L2] * CBAC;
        }
        if (L >= nold) {
            A = ABAC; B = BBAC; C = CBAC;
            LL = L;
            fnew[k] = A + (B + C * xk) * xk;
            continue;
        }
        D = (FO[L] - FO[L1]) / (XO[L] - XO[L1]);
        CFOR = FO[L + 1] / ((XO[L + 1] - XO[L]) * (XO[L + 1] - XO[L1]))
             + (FO[L1] / (XO[L + 1] - XO[L1]) - FO[L] / (XO[L + 1] - XO[L]))
             / (XO[L] - XO[L1]);
        BFOR = D - (XO[L] + XO[L1]) * CFOR;
        AFOR = FO[L1] - XO[L1] * D + XO[L] * XO[L1] * CFOR;
        WT = 0.0;
        if (fabs(CFOR) != 0.0) WT = fabs(CFOR) / (fabs(CFOR) + fabs(CBAC));
        A = AFOR + WT * (ABAC - AFOR);
        B = BFOR + WT * (BBAC - BFOR);
        C = CFOR + WT * (CBAC - CFOR);
        LL = L;
        fnew[k] = A + (B + C * xk) * xk;
    }
    return LL - 1;
}

/* SOLVIT, src/atlas12_modules.f90:5739-5847: in-place LU with partial
 * pivoting (swaps restricted to columns I+1..N, reciprocal pivot on the
 * diagonal, 1e-30 pivot floor).  a is column-major n x n. */
int or_solvit(double *a, int n, double *b)
{
#define A_(i, j) a[((j) - 1) * n + ((i) - 1)]
    const double TOL = 1.0e-30;
    int ipiv[KW_MAX];
    int hits = 0;
    for (int I = 1; I <= n - 1; ++I) {
        int M = I;
        for (int K = I + 1; K <= n; ++K)
            if (fabs(A_(K, I)) > fabs(A_(M, I))) M = K;
        ipiv[I] = M;
        if (M != I) {
            for (int K = I + 1; K <= n; ++K) {
                double T = A_(I, K);
                A_(I, K) = A_(M, K);
                A_(M, K) = T;
            }
        }
        if (fabs(A_(M, I)) < TOL) {
            A_(M, I) = copysign(TOL, A_(M, I));
            ++hits;
        }
        double piv = 1.0 / A_(M, I);
        A_(M, I) = A_(I, I);
        A_(I, I) = piv;
        for (int K = I + 1; K <= n; ++K) A_(K, I) = A_(K, I) * piv;
        for (int J = I + 1; J <= n; ++J) {
            double C = A_(I, J);
            if (C == 0.0) continue;
            for (int K = I + 1; K <= n; ++K) A_(K, J) = A_(K, J) - A_(K, I) * C;
        }
    }
    if (fabs(A_(n, n)) < TOL) {
        A_(n, n) = copysign(TOL, A_(n, n));
        ++hits;
    }
    A_(n, n) = 1.0 / A_(n, n);
    double *B = b - 1;
    for (int I = 1; I <= n - 1; ++I) {
        int M = ipiv[I];
        if (M != I) { double T = B[M]; B[M] = B[I]; B[I] = T; }
        double C = B[I];
        for (int K = I + 1; K <= n; ++K) B[K] = B[K] - A_(K, I) * C;
    }
    for (int J = n; J >= 2; --J) {
        B[J] = B[J] * A_(J, J);
        double C = B[J];
        for (int K = 1; K <= J - 1; ++K) B[K] = B[K] - A_(K, J) * C;
    }
    B[1] = B[1] * A_(1, 1);
#undef A_
    return hits;
}

/* EXPI, src/atlas12_modules.f90:21103-21225: E_1 by the Cody & Thacher
 * (1968) rational approximations, E_n by upward recurrence.  The
 * coefficients are the published ones (they have to match). */
double or_expi(int n, double x)
{
    static const double A0 = -44178.5471728217, A1 = 57721.7247139444,
        A2 = 9938.31388962037, A3 = 1842.11088668000, A4 = 101.093806161906,
        A5 = 5.03416184097568;
    static const double B0 = 76537.3323337614, B1 = 32597.1881290275,
        B2 = 6106.10794245759, B3 = 635.419418378382, B4 = 37.2298352833327;
    static const double C0 = 4.65627107975096e-7, C1 = 0.999979577051595,
        C2 = 9.04161556946329, C3 = 24.3784088791317, C4 = 23.0192559391333,
        C5 = 6.90522522784444, C6 = 0.430967839469389;
    static const double D1 = 10.0411643829054, D2 = 32.4264210695138,
        D3 = 41.2807841891424, D4 = 20.4494785013794, D5 = 3.31909213593302,
        D6 = 0.103400130404874;
    static const double E0 = -0.999999999998447, E1C = -26.6271060431811,
        E2 = -241.055827097015, E3 = -895.927957772937, E4 = -1298.85688746484,
        E5 = -545.374158883133, E6 = -5.66575206533869;
    static const double F1 = 28.6271060422192, F2 = 292.310039388533,
        F3 = 1332.78537748257, F4 = 2777.61949509163, F5 = 2404.01713225909,
        F6 = 631.657483280800;
    double ex = exp(-x), e1;
    if (x > 4.0) {
        e1 = (ex + ex * (E0 + (E1C + (E2 + (E3 + (E4 + (E5 + E6 / x) / x) / x) / x) / x) / x)
              / (x + F1 + (F2 + (F3 + (F4 + (F5 + F6 / x) / x) / x) / x) / x)) / x;
    } else if (x > 1.0) {
        e1 = ex * (C6 + (C5 + (C4 + (C3 + (C2 + (C1 + C0 * x) * x) * x) * x) * x) * x)
             / (D6 + (D5 + (D4 + (D3 + (D2 + (D1 + x) * x) * x) * x) * x) * x);
    } else if (x > 0.0) {
        e1 = (A0 + (A1 + (A2 + (A3 + (A4 + A5 * x) * x) * x) * x) * x)
             / (B0 + (B1 + (B2 + (B3 + (B4 + x) * x) * x) * x) * x) - log(x);
    } else {
        e1 = 0.0;
    }
    double r = e1;
    for (int i = 1; i <= n - 1; ++i) r = (ex - x * r) / (double)i;
    return r;
}

/* ===================================================================== */
/* Voigt: Weideman (1994) N=32 + Humlicek (1982) asymptote               */
/* src/synthe_module.f90:463-537, coefficients :91-126.                  */
/* ===================================================================== */

/* Weideman's coefficients for N = 32, L = sqrt(N/sqrt 2), Horner order
 * (index 0 multiplies Z^31).  These are the published constants of the
 * method; tests/ regenerate them from Weideman's FFT construction. */
static const double WEI_L = 4.75682846001088411;
static const double WEI_A[32] = {
    -1.30317978630500875e-12, 3.74088129316536249e-12, 8.03036789996388945e-12,
    -2.15436320778387687e-11, -5.54423594816646238e-11, 1.16582510935237737e-10,
    4.15374309183345315e-10, -5.23102048119632885e-10, -3.20801509172336887e-09,
    8.12488945684665156e-10, 2.37975567798974168e-08, 2.29304390650999664e-08,
    -1.48130789151209774e-07, -4.18407637021697758e-07, 4.25583313757500854e-07,
    4.40153173157854990e-06, 6.82103194400198485e-06, -2.14096192017107501e-05,
    -1.30754492546153456e-04, -2.45329802700214317e-04, 3.92591360700703107e-04,
    4.51954110534921738e-03, 1.90061557848454077e-02, 5.73044035298372195e-02,
    1.40607162268937685e-01, 2.95444510715087316e-01, 5.46013972063934094e-01,
    9.01925489364799882e-01, 1.34554416923454490e+00, 1.82566962963248147e+00,
    2.26353729990026764e+00, 2.57225340812456960e+00};

double or_voigt_f64(double v, double a)
{
    double x = fabs(v), y = a, s = x + y;
    if (s > 15.0) {
        double dr = x * x - y * y - 0.5;
        double di = 2.0 * x * y;
        double d2 = dr * dr + di * di;
        return (-y * dr + x * di) / (C_SQRTPI * d2);
    }
    double dr = WEI_L + y, di = -x;
    double nr = WEI_L - y, ni = x;
    double d2 = dr * dr + di * di;
    double zr = (nr * dr + ni * di) / d2;
    double zi = (ni * dr - nr * di) / d2;
    double pr = WEI_A[0], pim = 0.0;
    for (int k = 1; k < 32; ++k) {
        double tmp = pr * zr - pim * zi + WEI_A[k];
        pim = pr * zi + pim * zr;
        pr = tmp;
    }
    double sqr = dr * dr - di * di;
    double sqi = 2.0 * dr * di;
    double sqd = sqr * sqr + sqi * sqi;
    double t1 = 2.0 * (pr * sqr + pim * sqi) / sqd;
    double t2 = dr / (C_SQRTPI * d2);
    return t1 + t2;
}

float or_voigt_ref(float v, float a)
{
    return (float)or_voigt_f64((double)fabsf(v), (double)a);
}

/* ===================================================================== */
/* JOSH, src/atlas12_modules.f90:18398-18776                              */
/* ===================================================================== */

/* CH_J, CK_J, XTAU8: src/atlas12_modules.f90:18409-18451 (the fixed
 * 51-point quadrature of the method). */
static const double CH_J[51] = {
    7.15528131e-07, 1.49142693e-06, 1.52106577e-06, 2.98150826e-06,
    5.33941056e-06, 9.13329677e-06, 1.61715943e-05, 2.97035986e-05,
    5.33166603e-05, 9.11154202e-05, 1.61084638e-04, 2.95118050e-04,
    5.27450291e-04, 8.67939554e-04, 1.61498412e-03, 2.50720908e-03,
    3.20994272e-03, 5.61912498e-03, 8.60872678e-03, 1.04706492e-02,
    1.33110350e-02, 1.62635669e-02, 1.90288834e-02, 2.18877215e-02,
    2.36015432e-02, 2.10819542e-02, 1.80345085e-02, 1.64786074e-02,
    1.49382707e-02, 1.19676525e-02, 9.90213640e-03, 8.17766134e-03,
    6.11252524e-03, 4.84035723e-03, 3.06078210e-03, 2.40512565e-03,
    2.01712688e-03, 1.33288081e-03, 7.83530239e-04, 4.31428343e-04,
    1.76504589e-04, 4.75738016e-05, 1.65963702e-05, 5.41117970e-06,
    2.08043571e-06, 7.11612643e-07, 8.08788982e-08, 1.95130507e-08,
    4.33638281e-09, 8.87765583e-10, 3.90236420e-11};
static const double CK_J[51] = {
    3.57771910e-07, 7.45730404e-07, 7.60575176e-07, 1.49091113e-06,
    2.67016185e-06, 4.56793896e-06, 8.08956065e-06, 1.48632944e-05,
    2.66928291e-05, 4.56529851e-05, 8.08134864e-05, 1.48363324e-04,
    2.66052346e-04, 4.39771306e-04, 8.25088180e-04, 1.29440730e-03,
    1.67680858e-03, 2.98973685e-03, 4.68314718e-03, 5.84855257e-03,
    7.64854718e-03, 9.63155832e-03, 1.16419578e-02, 1.38551742e-02,
    1.54840983e-02, 1.42877987e-02, 1.25930300e-02, 1.17983138e-02,
    1.09717194e-02, 8.98320694e-03, 7.59950886e-03, 6.38808031e-03,
    4.86854184e-03, 3.91568616e-03, 2.51398841e-03, 2.00142385e-03,
    1.70069211e-03, 1.14058319e-03, 6.80292083e-04, 3.80097074e-04,
    1.57705377e-04, 4.31706540e-05, 1.51795348e-05, 4.98576401e-06,
    1.92979223e-06, 6.63957223e-07, 7.65236692e-08, 1.84933668e-08,
    4.12596224e-09, 8.47334369e-10, 3.81791959e-11};
static const double XTAU8[51] = {
    0.0, 0.0000032, 0.0000056, 0.00001, 0.000018, 0.000032, 0.000056, 0.0001,
    0.00018, 0.00032, 0.00056, 0.001, 0.0018, 0.0032, 0.0056, 0.01, 0.016,
    0.025, 0.042, 0.065, 0.096, 0.139, 0.196, 0.273, 0.375, 0.5, 0.63, 0.78,
    0.95, 1.15, 1.35, 1.6, 1.85, 2.15, 2.45, 2.75, 3.15, 3.65, 4.25, 5.0, 6.0,
    7.0, 8.0, 9.0, 10.0, 11.5, 13.0, 14.5, 16.0, 18.0, 20.0};

double or_josh(or_josh_io *io, int ifscat, int ifsurf)
{
    const int n = io->nrhox, NX = OR_NXTAU;
    double abtot[KW_MAX] = {0}, alpha[KW_MAX], snubar[KW_MAX];
    double taunu_l[KW_MAX], snu_l[KW_MAX], jnu_l[KW_MAX], hnu_l[KW_MAX], jmins_l[KW_MAX];
    double *taunu = io->taunu ? io->taunu : taunu_l;
    double *snu = io->snu ? io->snu : snu_l;
    double *jnu = io->jnu ? io->jnu : jnu_l;
    double *hnu = io->hnu ? io->hnu : hnu_l;
    double *jmins = io->jmins ? io->jmins : jmins_l;
    double xs[OR_NXTAU], xsbar[OR_NXTAU], xalpha[OR_NXTAU];
    double xsbar8[OR_NXTAU], xalpha8[OR_NXTAU], xs8[OR_NXTAU];
    double xh[OR_NXTAU], xjs[OR_NXTAU];
    double amat[OR_NXTAU * OR_NXTAU];
    int maxj = 0;

    io->vef_sweeps = 0;
    io->vef_unconverged = 0;
    io->pivot_floor_hits = 0;
    for (int j = 0; j < n; ++j) {
        abtot[j] = io->acont[j] + io->aline[j] + io->sigmac[j] + io->sigmal[j];
        alpha[j] = (io->sigmac[j] + io->sigmal[j]) / abtot[j];
        snubar[j] = (io->acont[j] * io->scont[j] + io->aline[j] * io->sline[j])
                    / (io->acont[j] + io->aline[j]);
    }
    or_integ(io->rhox, abtot, taunu, n, abtot[0] * io->rhox[0]);

    int mode = 0; /* 60 surface flux; 0 full */
    if (ifscat == 0) {
        for (int j = 0; j < n; ++j) snu[j] = snubar[j];
        maxj = or_map1(taunu, snu, n, XTAU8, xs8, NX);
        for (int l = 0; l < NX; ++l) xs[l] = xs8[l];
        if (ifsurf == 1) {
            mode = 60;
            goto dispatch;
        }
        for (int j = 0; j < n; ++j) alpha[j] = 0.0;
    }
    if (taunu[0] > XTAU8[NX - 1]) maxj = 1;
    if (maxj != 1) {
        maxj = or_map1(taunu, snubar, n, XTAU8, xsbar8, NX);
        maxj = or_map1(taunu, alpha, n, XTAU8, xalpha8, NX);
        for (int l = 0; l < NX; ++l) {
            xalpha8[l] = fmax(xalpha8[l], 0.0);
            xsbar8[l] = fmax(xsbar8[l], 1.0e-38);
            xalpha[l] = xalpha8[l];
            xsbar[l] = xsbar8[l];
            if (XTAU8[l] < taunu[0]) {
                xsbar8[l] = fmax(snubar[0], 1.0e-38);
                xalpha8[l] = fmax(alpha[0], 0.0);
                xsbar[l] = xsbar8[l];
                xalpha[l] = xalpha8[l];
            }
        }
        double amax = xalpha[0];
        for (int l = 1; l < NX; ++l) amax = fmax(amax, xalpha[l]);
        if (amax <= 0.0) {
            for (int l = 0; l < NX; ++l) xs[l] = fmax(xsbar[l], 1.0e-37);
        } else {
            for (int l = 0; l < NX; ++l) {
                for (int m = 0; m < NX; ++m)
                    amat[m * NX + l] = -xalpha[l] * io->coefj[m * NX + l];
                amat[l * NX + l] = amat[l * NX + l] + 1.0;
                xs[l] = (1.0 - xalpha[l]) * xsbar[l];
            }
            io->pivot_floor_hits = or_solvit(amat, NX, xs);
            for (int l = 0; l < NX; ++l) xs[l] = fmax(xs[l], 1.0e-37);
        }
        if (ifsurf == 1) {
            mode = 60;
            goto dispatch;
        }
        for (int m = 0; m < NX; ++m) xs8[m] = xs[m];
        (void)or_map1(XTAU8, xs8, NX, taunu, snu, maxj);
    }
    /* Deep atmosphere: variable Eddington factor, :18581-18639 */
    if (maxj != n) {
        int maxj1 = maxj + 1;
        if (maxj == 1) maxj1 = 1;
        for (int j = maxj1; j <= n; ++j) snu[j - 1] = snubar[j - 1];
        int m = maxj - 1 > 1 ? maxj - 1 : 1;
        int nm1 = n - m + 1;
        int nmj = n - maxj + 1;
        double error = 0.0;
        int l;
        for (l = 1; l <= 50; ++l) {
            error = 0.0;
            int ifneg = 0;
            for (int j = m; j <= n; ++j) {
                if (snu[j - 1] <= 0.0) {
                    ifneg = 1;
                    for (int jj = m; jj <= n; ++jj) {
                        snubar[jj - 1] = io->bnu[jj - 1];
                        snu[jj - 1] = io->bnu[jj - 1];
                    }
                    break;
                }
            }
            or_deriv(taunu + m - 1, snu + m - 1, hnu + m - 1, nm1);
            for (int j = m; j <= n; ++j) {
                if (hnu[j - 1] <= 0.0) {
                    ifneg = 1;
                    for (int jj = m; jj <= n; ++jj) {
                        snubar[jj - 1] = io->bnu[jj - 1];
                        snu[jj - 1] = io->bnu[jj - 1];
                    }
                    or_deriv(taunu + m - 1, snu + m - 1, hnu + m - 1, nm1);
                    break;
                }
            }
            for (int j = m; j <= n; ++j) hnu[j - 1] = hnu[j - 1] / 3.0;
            or_deriv(taunu + maxj - 1, hnu + maxj - 1, jmins + maxj - 1, nmj);
            for (int j = maxj1; j <= n; ++j) {
                if (ifneg == 1) jmins[j - 1] = 0.0;
                jnu[j - 1] = jmins[j - 1] + snu[j - 1];
                double snew = (1.0 - alpha[j - 1]) * snubar[j - 1] + alpha[j - 1] * jnu[j - 1];
                error = fabs(snew - snu[j - 1]) / snew + error;
                snu[j - 1] = snew;
            }
            if (error < 0.00001) break;
        }
        io->vef_sweeps = l > 50 ? 50 : l;
        if (error >= 0.00001) io->vef_unconverged = 1;
    }
    if (maxj == 1) {
        io->knu1 = jnu[0] / 3.0;
        io->maxj = maxj;
        return hnu[0];
    }
    /* J, H, K from the 51-point operators, :18654-18676 */
    {
        double xjs8[OR_NXTAU], xh8[OR_NXTAU];
        for (int l = 0; l < NX; ++l) {
            xjs[l] = -xs[l];
            for (int m = 0; m < NX; ++m) xjs[l] = xjs[l] + io->coefj[m * NX + l] * xs[m];
            xjs8[l] = xjs[l];
            xh[l] = 0.0;
            for (int m = 0; m < NX; ++m) xh[l] = xh[l] + io->coefh[m * NX + l] * xs[m];
            xh8[l] = xh[l];
        }
        (void)or_map1(XTAU8, xjs8, NX, taunu, jmins, maxj);
        (void)or_map1(XTAU8, xh8, NX, taunu, hnu, maxj);
        double xk = 0.0;
        for (int m = 0; m < NX; ++m) xk = xk + CK_J[m] * xs[m];
        io->knu1 = xk;
        for (int j = 0; j < maxj; ++j) {
            snu[j] = fmax(snu[j], 1.0e-38);
            jnu[j] = fmax(jmins[j] + snu[j], 1.0e-38);
        }
        io->maxj = maxj;
        return hnu[0];
    }
dispatch:
    /* output mode 60: surface flux, :18686-18692 */
    {
        double h1 = 0.0;
        for (int m = 0; m < NX; ++m) h1 = h1 + CH_J[m] * xs[m];
        hnu[0] = h1;
        io->maxj = maxj;
        (void)mode;
        return h1;
    }
}

/* ===================================================================== */
/* SYNTHE line opacity                                                    */
/* ===================================================================== */

/* Continuum at the line grid, src/synthe.f90:481-516: log-quadratic
 * Lagrange interpolation over edge intervals, floor at log10 = -15, then
 * 10**x in FP32. */
void or_synthe_continuum(const or_synthe_atm *atm, int j, float *out,
                         double *out_f64)
{
    const int ne = atm->nedge;
    double half[4096], del[4096];
    const double *w = atm->wledge;
    for (int e = 1; e < ne; ++e) {
        half[e - 1] = (w[e - 1] + w[e]) * 0.5;
        del[e - 1] = (w[e] - w[e - 1]) * (w[e] - w[e - 1]) * 0.5;
    }
    const double *ab = atm->continall + (size_t)j * atm->nf;
    int ie = 0; /* 0-based iedge */
    for (int i = 0; i < atm->length; ++i) {
        double wave8 = atm->wbegin * powi_f64(atm->ratio, i);
        while (wave8 >= w[ie + 1] && ie < ne - 2) ie++;
        double x = ((wave8 - half[ie]) * (wave8 - w[ie + 1]) * ab[3 * ie]
                    + (w[ie] - wave8) * (wave8 - w[ie + 1]) * 2.0 * ab[3 * ie + 1]
                    + (wave8 - w[ie]) * (wave8 - half[ie]) * ab[3 * ie + 2]) / del[ie];
        float xf = (float)x;
        if (xf < -15.0f) xf = -15.0f;
        out[i] = powf(10.0f, xf);
        if (out_f64) {
            double xd = x < -15.0 ? -15.0 : x;
            out_f64[i] = pow(10.0, xd);
        }
    }
}

void or_synthe_line_opacity_range(const or_synthe_atm *atm, const or_line_t *lines,
                                  int nlines, int prec, int j0, int j1, float *out32,
                                  double *out64, int64_t *evals, int32_t *nstep_out)
{
    const int L = atm->length, mw6 = atm->mw6;
    float *cont = (float *)malloc(sizeof(float) * L);
    double *cont64 = (double *)malloc(sizeof(double) * L);
    float *buf = (float *)malloc(sizeof(float) * L);
    double *buf64 = (double *)malloc(sizeof(double) * L);
    float *prof = (float *)malloc(sizeof(float) * (OR_MAXPROF + 2));
    double *prof64 = (double *)malloc(sizeof(double) * (OR_MAXPROF + 2));
    double *xnfdop = (double *)malloc(sizeof(double) * mw6);
    const double resolu = atm->resolu;
    int64_t nev = 0;

    for (int j = j0; j < j1; ++j) {
        or_synthe_continuum(atm, j, cont, cont64);
        for (int i = 0; i < L; ++i) { buf[i] = 0.0f; buf64[i] = 0.0; }
        /* per-depth prep, src/synthe.f90:518-551 */
        const double *qx = atm->xnfpel + (size_t)j * mw6;
        const double *qd = atm->dopple + (size_t)j * mw6;
        for (int i = 0; i < mw6; ++i)
            xnfdop[i] = qd[i] > 0.0 ? qx[i] / atm->rho[j] / qd[i] : 0.0;
        /* t, xne, xnfh, xnfhe, xnfh2 pass through REAL() (FP32) in the
         * reference before this expression (src/synthe.f90:392-410). */
        double t32 = (double)(float)atm->t[j];
        double xne32 = (double)(float)atm->xne[j];
        double h32 = (double)(float)atm->xnfh[j];
        double he32 = (double)(float)atm->xnfhe1[j];
        double h232 = (double)(float)atm->xnfh2[j];
        float txnxn = (float)((h32 + 0.42 * he32 + 0.85 * h232) * pow(t32 / 10000.0, 0.3));
        double txnxn64 = (atm->xnfh[j] + 0.42 * atm->xnfhe1[j] + 0.85 * atm->xnfh2[j])
                         * pow(atm->t[j] / 10000.0, 0.3);
        const double hckt = atm->hckt[j];

        for (int il = 0; il < nlines; ++il) {
            const or_line_t *ln = &lines[il];
            const int nb = ln->nbuff;           /* 1-based grid index */
            const int nel = ln->nelion - 1;
            int nstep = -1;
            int cidx = nb < 1 ? 1 : (nb > L ? L : nb);
            if (prec == OR_PREC_REF) {
                float kap0 = ln->cgf * (float)xnfdop[nel];
                float kapmin = cont[cidx - 1] * 1.0e-3f;
                if (!(kap0 < kapmin)) {
                    kap0 = kap0 * (float)exp(-(double)ln->elo * hckt);
                    if (!(kap0 < kapmin)) {
                        float gw = ln->gamwf * txnxn;
                        float adamp = (float)(((double)ln->gamrf + (double)ln->gamsf * xne32
                                               + (double)gw) / qd[nel]);
                        int n10 = (int)(10.0 * qd[nel] * resolu);
                        float dnel = (float)qd[nel];
                        if (nb >= 1 && nb <= L) {
                            buf[nb - 1] += kap0 * or_voigt_ref(0.0f, adamp);
                            ++nev;
                        }
                        float dv = 1.0f / dnel / (float)resolu;
                        int n;
                        for (n = 1; n <= n10; ++n) {
                            prof[n] = kap0 * or_voigt_ref((float)n * dv, adamp);
                            if (prof[n] < kapmin) break;
                        }
                        nstep = n;
                        if (nstep > n10) {
                            float xw = n10 >= 1 ? prof[n10] * ((float)n10 * (float)n10) : 0.0f;
                            float sq = sqrtf(xw / kapmin);
                            int maxstep = sq >= 999999.0f ? OR_MAXPROF : (int)sq + 1;
                            for (int k = n10 + 1; k <= maxstep; ++k)
                                prof[k] = xw / ((float)k * (float)k);
                            nstep = maxstep;
                        }
                        if (!(nb + nstep < 1 || nb - nstep > L)) {
                            /* red then blue wing, per-point clipping :653-682 */
                            for (int k = 1; k <= nstep; ++k) {
                                int p = nb + k;
                                if (p >= 1 && p <= L) { buf[p - 1] += prof[k]; ++nev; }
                            }
                            for (int k = 1; k <= nstep; ++k) {
                                int p = nb - k;
                                if (p >= 1 && p <= L) { buf[p - 1] += prof[k]; ++nev; }
                            }
                        }
                    }
                }
            } else {
                double kap0 = (double)ln->cgf * xnfdop[nel];
                double kapmin = cont64[cidx - 1] * 1.0e-3;
                if (!(kap0 < kapmin)) {
                    kap0 = kap0 * exp(-(double)ln->elo * hckt);
                    if (!(kap0 < kapmin)) {
                        double adamp = ((double)ln->gamrf + (double)ln->gamsf * atm->xne[j]
                                        + (double)ln->gamwf * txnxn64) / qd[nel];
                        int n10 = (int)(10.0 * qd[nel] * resolu);
                        if (nb >= 1 && nb <= L) {
                            buf64[nb - 1] += kap0 * or_voigt_f64(0.0, adamp);
                            ++nev;
                        }
                        double dv = 1.0 / qd[nel] / resolu;
                        int n;
                        for (n = 1; n <= n10; ++n) {
                            prof64[n] = kap0 * or_voigt_f64((double)n * dv, adamp);
                            if (prof64[n] < kapmin) break;
                        }
                        nstep = n;
                        if (nstep > n10) {
                            double xw = n10 >= 1 ? prof64[n10] * ((double)n10 * (double)n10) : 0.0;
                            double sq = sqrt(xw / kapmin);
                            int maxstep = sq >= 999999.0 ? OR_MAXPROF : (int)sq + 1;
                            for (int k = n10 + 1; k <= maxstep; ++k)
                                prof64[k] = xw / ((double)k * (double)k);
                            nstep = maxstep;
                        }
                        if (!(nb + nstep < 1 || nb - nstep > L)) {
                            for (int k = 1; k <= nstep; ++k) {
                                int p = nb + k;
                                if (p >= 1 && p <= L) { buf64[p - 1] += prof64[k]; ++nev; }
                            }
                            for (int k = 1; k <= nstep; ++k) {
                                int p = nb - k;
                                if (p >= 1 && p <= L) { buf64[p - 1] += prof64[k]; ++nev; }
                            }
                        }
                    }
                }
            }
            if (nstep_out) nstep_out[(size_t)j * nlines + il] = nstep < 0 ? -1 : nstep;
        }
        if (prec == OR_PREC_REF) {
            if (out32) memcpy(out32 + (size_t)j * L, buf, sizeof(float) * L);
        } else {
            if (out64) memcpy(out64 + (size_t)j * L, buf64, sizeof(double) * L);
        }
    }
    if (evals) *evals = nev;
    free(cont); free(cont64); free(buf); free(buf64); free(prof); free(prof64);
    free(xnfdop);
}

void or_synthe_line_opacity(const or_synthe_atm *atm, const or_line_t *lines,
                            int nlines, int prec, float *out32, double *out64,
                            int64_t *evals, int32_t *nstep_out)
{
    or_synthe_line_opacity_range(atm, lines, nlines, prec, 0, atm->nrhox, out32, out64,
                                 evals, nstep_out);
}

/* ===================================================================== */
/* SYNTHE transfer per wavelength point                                   */
/* src/synthe.f90:698-718, setup_opacity_sv :797-833,                     */
/* process_wavelength_point :848-902.  IFSCAT = 1, IFSURF = 1 as set by  */
/* READIN(20) (src/atlas12_modules.f90:1293, :6689).                      */
/* ===================================================================== */

static int rt_edge(const or_synthe_rt *rt, double wave)
{
    /* iedge_sv advances monotonically in the reference; the bracket is the
     * first interval whose upper edge exceeds wave. */
    int ie = 0;
    while (ie + 1 < rt->nedge && wave >= rt->wledge[ie + 1]) ie++;
    return ie;
}

/* The transfer of one wavelength point; taunu (nullable, nrhox): TAUNU of
 * the second JOSH call, the table synthe.f90 writes to .linform (:889-890). */
static void synthe_point(const or_synthe_rt *rt, int iwave, const float *opac_col,
                         const double *opac_col64, double *h_line, double *h_cont,
                         double *taunu)
{
    const int n = rt->nrhox;
    double acont[KW_MAX], sigmac[KW_MAX], aline[KW_MAX], sigmal[KW_MAX];
    double scont[KW_MAX], sline[KW_MAX], bnu[KW_MAX], ehvkt[KW_MAX], stim[KW_MAX];
    double asyn[KW_MAX];
    double wave = rt->wbegin * powi_f64(rt->ratio, iwave);
    double freq8 = C_CLIGHT_NMS / wave;
    for (int j = 0; j < n; ++j) {
        if (opac_col64)
            asyn[j] = opac_col64[j] * (1.0 - exp(-freq8 * rt->hkt_syn[j]));
        else
            asyn[j] = (double)(float)((double)opac_col[j] * (1.0 - exp(-freq8 * rt->hkt_syn[j])));
    }
    int ie = rt_edge(rt, wave);
    const double *w = rt->wledge;
    double half = (w[ie] + w[ie + 1]) * 0.5;
    double del = (w[ie + 1] - w[ie]) * (w[ie + 1] - w[ie]) * 0.5;
    double c1 = (wave - half) * (wave - w[ie + 1]) / del;
    double c2 = (w[ie] - wave) * (wave - w[ie + 1]) * 2.0 / del;
    double c3 = (wave - w[ie]) * (wave - half) / del;
    for (int j = 0; j < n; ++j) {
        const double *ca = rt->contabs + (size_t)j * rt->nf + 3 * ie;
        const double *cs = rt->contscat + (size_t)j * rt->nf + 3 * ie;
        acont[j] = pow(10.0, c1 * ca[0] + c2 * ca[1] + c3 * ca[2]);
        sigmac[j] = pow(10.0, c1 * cs[0] + c2 * cs[1] + c3 * cs[2]);
    }
    const double bnu_prefac = 2.0 * C_HPLANCK / (C_CLIGHT * C_CLIGHT) * 1.0e45;
    double freq = C_CLIGHT_NMS / wave;
    double f15 = freq / 1.0e15;
    for (int j = 0; j < n; ++j) {
        ehvkt[j] = exp(-freq * rt->hkt_atl[j]);
        stim[j] = 1.0 - ehvkt[j];
        bnu[j] = bnu_prefac * (f15 * f15 * f15) * ehvkt[j] / stim[j];
        aline[j] = 0.0;
        sigmal[j] = 0.0;
        sline[j] = bnu[j];
        scont[j] = bnu[j];
    }
    or_josh_io io;
    memset(&io, 0, sizeof io);
    io.nrhox = n;
    io.acont = acont; io.aline = aline; io.sigmac = sigmac; io.sigmal = sigmal;
    io.scont = scont; io.sline = sline; io.rhox = rt->rhox; io.bnu = bnu;
    io.coefj = rt->coefj; io.coefh = rt->coefh;
    double surf = or_josh(&io, 1, 1);
    for (int j = 0; j < n; ++j) {
        aline[j] = asyn[j] * (1.0 - 0.0);
        sline[j] = bnu[j] * stim[j] / (1.0 - ehvkt[j]);
        sigmal[j] = asyn[j] * 0.0;
    }
    io.taunu = taunu;
    double hl = or_josh(&io, 1, 1);
    double resid = hl / surf;
    *h_line = resid * surf;
    *h_cont = surf;
}

void or_synthe_transfer_point(const or_synthe_rt *rt, int iwave,
                              const float *opac_col, const double *opac_col64,
                              double *h_line, double *h_cont)
{
    synthe_point(rt, iwave, opac_col, opac_col64, h_line, h_cont, NULL);
}

void or_synthe_taunu(const or_synthe_rt *rt, const float *opac, const double *opac64,
                     double *taunu, int i0, int i1)
{
    const int n = rt->nrhox, L = rt->length;
    float col[KW_MAX];
    double col64[KW_MAX], hl, hc;
    for (int i = i0; i < i1; ++i) {
        for (int j = 0; j < n; ++j) {
            if (opac64) col64[j] = opac64[(size_t)j * L + i];
            else col[j] = opac[(size_t)j * L + i];
        }
        synthe_point(rt, i, col, opac64 ? col64 : NULL, &hl, &hc, taunu + (size_t)(i - i0) * n);
    }
}

void or_synthe_transfer(const or_synthe_rt *rt, const float *opac,
                        const double *opac64, double *h_line, double *h_cont,
                        int i0, int i1)
{
    const int n = rt->nrhox, L = rt->length;
    float col[KW_MAX];
    double col64[KW_MAX];
    for (int i = i0; i < i1; ++i) {
        for (int j = 0; j < n; ++j) {
            if (opac64) col64[j] = opac64[(size_t)j * L + i];
            else col[j] = opac[(size_t)j * L + i];
        }
        or_synthe_transfer_point(rt, i, col, opac64 ? col64 : NULL,
                                 &h_line[i - i0], &h_cont[i - i0]);
    }
}
