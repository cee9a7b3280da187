// a12_cplx_core.h -- SYNTHE's complex-profile lines (compute_line_opacity,
// src/synthe_module.f90:3167-3702, with hydrogen_line_profile :1032-1980)
// as host+device functions.  Included by a12_cplx.hip, whose kernels run
// them for every (line, depth) and (depth, point), and by the host build
// that checks them against the reference's own buffers.
//
// The reference walks each line's red and blue wings point by point away
// from the centre, adding kappa into an FP32 buffer until kappa drops below
// the continuum times the cutoff.  Whether a point is reached depends only
// on kappa at the points before it, not on the buffer, so the work splits
// into:
//   plan_line:   per (line, depth), the set-up at the top of each CASE
//                (kappa0 after the Boltzmann factor, widths, merge points);
//   conv_build:  per (H line, depth), the Stehle (x) Voigt convolution
//                hydrogen_line_profile caches between calls;
//   red_step / blue_step: kappa at one point of a wing and whether it is
//                added and whether the walk ends there -- the walk's
//                extents follow from a scan over these;
//   point_kappa: the value a point receives, summed per point in line
//                order (the reference's order of additions into BUFFER).
// The wavelength of each wing point is the reference's running product
// (wave = wave * ratio, or / ratio) from the wing's first point, so the
// chain is reproduced exactly (see chain_*).  Expressions keep the
// reference's types (REAL(4) where it declares REAL(4)) and its left-to-right
// evaluation; the translation unit compiles with FP contraction off.
#pragma once

#include <math.h>
#include <stdint.h>

#include "a12gpu.h"

#ifndef A12_HD
#define A12_HD __host__ __device__
#endif

namespace cplx {

#include "a12_cplx_data.h"

// ---- mod_constants (src/atlas12_modules.f90:315-402) -----------------------
constexpr double PI = 3.141592653589793, FOURPI = 4.0 * PI, SQRTPI = 1.7724538509055159;
constexpr double CLIGHT = 2.99792458e10, CLIGHT_NMS = 2.99792458e17, CLIGHT_ANGS = 2.99792458e18;
constexpr double CLIGHT_KMS = 2.99792458e5;
constexpr double RYDBERG_INF = 109737.31568157, RYDBERG_H = 109677.583403, ELIM_HI = 109678.764;
constexpr double FREQ_RYDH = RYDBERG_H * CLIGHT;
constexpr double RYD_ANG = 1.0e8 / RYDBERG_H;
constexpr double LN10 = 2.302585092994046;  // LOG(10.0D0), correctly rounded
constexpr float CLIGHT4 = (float)CLIGHT;     // REAL(CLIGHT,4)
// :392 H_SERIES_LIMITS (cm^-1), CONTH of compute_line_opacity and column 1
// of CONTX
static constexpr double H_SERIES_LIMITS[15] = {
    109678.764, 27419.659, 12186.462, 6854.871, 4387.113, 3046.604, 2238.320, 1713.713,
    1354.044, 1096.776, 906.426, 761.650, 648.980, 559.579, 487.456};

// Weideman (1994) N = 32 coefficients and L of voigt_profile (:70-95)
constexpr double WEI_L = 4.75682846001088411;
static constexpr double WEI_A[32] = {
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

constexpr int NDA = 60;       // NSTARK_DALPHA
constexpr int NTEMP = 10;     // NSTARK_TEMPS
constexpr int NCONV_MAX = 600, NWING_MAX = 200;
constexpr int MAXFINE = 14;

// REAL(8)**INTEGER with a run-time exponent: repeated squaring from the
// lowest bit of the exponent.
A12_HD inline double powi_rt(double x, int m)
{
    unsigned k = m < 0 ? -m : m;
    double y = (k & 1u) ? x : 1.0;
    while (k >>= 1) {
        x = x * x;
        if (k & 1u) y = y * x;
    }
    return m < 0 ? 1.0 / y : y;
}

// voigt_profile (:463-537): REAL(4) argument and result, REAL(8) inside.
A12_HD inline float voigt_profile(float v, float a)
{
    const double x = (double)fabsf(v), y = (double)a, s = x + y;
    if (s > 15.0) {
        const double dr = x * x - y * y - 0.5;
        const double di = 2.0 * x * y;
        const double d2 = dr * dr + di * di;
        return (float)((-y * dr + x * di) / (SQRTPI * d2));
    }
    const double dr = WEI_L + y, di = -x;
    const double nr = WEI_L - y, ni = x;
    const double d2 = dr * dr + di * di;
    const double zr = (nr * dr + ni * di) / d2;
    const double zi = (ni * dr - nr * di) / d2;
    double pr = WEI_A[0], pim = 0.0;
    for (int k = 1; k < 32; ++k) {
        const double t = pr * zr - pim * zi + WEI_A[k];
        pim = pr * zi + pim * zr;
        pr = t;
    }
    const double sqr = dr * dr - di * di;
    const double sqi = 2.0 * dr * di;
    const double sqd = sqr * sqr + sqi * sqi;
    const double term1 = 2.0 * (pr * sqr + pim * sqi) / sqd;
    const double term2 = dr / (SQRTPI * d2);
    return (float)(term1 + term2);
}

// vac_to_air (:546-558)
A12_HD inline double vac_to_air(double w)
{
    const double waven = 1.0e7 / w;
    return w / (1.0000834254 + 2406147.0 / (1.30e10 - waven * waven) +
                15998.0 / (3.89e9 - waven * waven));
}

// EHYD(n) of compute_line_opacity (:3302-3313): observed levels n <= 8,
// Rydberg formula above
A12_HD inline double ehyd(int n)
{
    switch (n) {
    case 1: return 0.0;
    case 2: return 82259.105;
    case 3: return 97492.302;
    case 4: return 102823.893;
    case 5: return 105291.651;
    case 6: return 106632.160;
    case 7: return 107440.444;
    case 8: return 107965.051;
    default: return ELIM_HI - RYDBERG_H / ((double)n * (double)n);
    }
}
// ALPHAHYD(n) (:3317-3319): wavelength (nm) of n -> n+1
A12_HD inline double alphahyd(int n) { return 1.0e7 / (ehyd(n + 1) - ehyd(n)); }

// CONTX(ncon, nelionx) (:3225-3292); 0 where the column holds no edge
A12_HD inline double contx(int ncon, int nelionx)
{
    if (ncon < 1 || ncon > 26) return 0.0;
    const int i = ncon - 1;
    switch (nelionx) {
    case 1: return i < 10 ? H_SERIES_LIMITS[i] : 0.0;
    case 2: return i < 6 ? CONTX_HE1[i] : 0.0;
    case 3: return i < 6 ? CONTX_HE2[i] : 0.0;
    case 4: return i < 10 ? CONTX_C1[i] : 0.0;
    case 6: return i < 4 ? CONTX_MG1[i] : 0.0;
    case 8: return i < 2 ? CONTX_AL1[i] : 0.0;
    case 10: return i < 10 ? CONTX_SI1[i] : 0.0;
    default: return 0.0;
    }
}

// hydrogen_oscillator_strength (:679-713) for the lower level 1, the only
// one hydrogen_line_profile asks for (its SAVE cache then always matches).
A12_HD inline float hyd_osc1(int m)
{
    const int n = 1;
    if (m <= n) return 0.0f;
    const float xn = (float)n;
    const float ginf = 0.2027f / powf(xn, 0.71f);
    const float gca = 0.124f / xn;
    const float fkn = xn * 1.9603f;
    const float wtc = 0.45f - 2.4f / (xn * xn * xn) * (xn - 1.0f);
    const float xm = (float)m, xmn = (float)(m - n);
    const float q = xm / (xmn * (xm + (float)n));
    const float fk = fkn * (q * q * q);
    const float xmn12 = powf(xmn, 1.2f);
    const float wt = (xmn12 - 1.0f) / (xmn12 + wtc);
    return fk * (1.0f - wt * ginf - (0.222f + gca / xm) * (1.0f - wt));
}

// expint1 (:731-753)
A12_HD inline float expint1(float x)
{
    if (x <= 0.0f) return 0.0f;
    if (x <= 1.0f)
        return -logf(x) - 0.57721566f +
               x * (0.99999193f + x * (-0.24991055f + x * (0.05519968f + x * (-0.00976004f + x * 0.00107857f))));
    if (x <= 30.0f)
        return (x * (x + 2.334733f) + 0.25062f) / (x * (x + 3.330657f) + 1.681534f) / x * expf(-x);
    return 0.0f;
}

// stark_quasistatic_profile (:771-950); p >= 0 (the reference stops on p < 0)
A12_HD inline float stark_qs(float b, float p, int n, int m)
{
    const float b2 = b * b;
    const float sb = sqrtf(b);
    if (b > 500.0f) return (1.5f / sb + 27.0f / b2) / b2;
    const int mmn = m - n;
    int indx = 7;
    if (n <= 3 && mmn <= 2) indx = 2 * (n - 1) + mmn;
    int im = (int)(5.0f * p) + 1;
    if (im > 4) im = 4;
    const int ip = im + 1;
    const float wtpp = 5.0f * (p - QS_PP[im - 1]);
    const float wtpm = 1.0f - wtpp;
    if (b > 25.12f) {
        const float c = QS_CC[indx - 1][ip - 1] * wtpp + QS_CC[indx - 1][im - 1] * wtpm;
        const float d = QS_DD[indx - 1][ip - 1] * wtpp + QS_DD[indx - 1][im - 1] * wtpm;
        const float corr = 1.0f + d / (c + b * sb);
        return (1.5f / sb + 27.0f / b2) / b2 * corr;
    }
    int jp = 2;
    for (int j = 2; j <= 15; ++j) {
        jp = j;
        if (b <= QS_BETA[j - 1]) break;
    }
    const int jm = jp - 1;
    const float wtbp = (b - QS_BETA[jm - 1]) / (QS_BETA[jp - 1] - QS_BETA[jm - 1]);
    const float wtbm = 1.0f - wtbp;
    const float *P = QS_PROB[indx - 1];
    const float vim_jm = P[(jm - 1) * 5 + im - 1], vip_jm = P[(jm - 1) * 5 + ip - 1];
    const float vim_jp = P[(jp - 1) * 5 + im - 1], vip_jp = P[(jp - 1) * 5 + ip - 1];
    const float cbp = vip_jp * wtpp + vim_jp * wtpm;
    const float cbm = vip_jm * wtpp + vim_jm * wtpm;
    const float corr = 1.0f + cbp * wtbp + cbm * wtbm;
    float pr1 = 0.0f, pr2 = 0.0f;
    float wt = 0.5f * (10.0f - b);
    wt = wt < 1.0f ? wt : 1.0f;
    wt = wt > 0.0f ? wt : 0.0f;
    if (b <= 10.0f) pr1 = 8.0f / (83.0f + (2.0f + 0.95f * b2) * b);
    if (b >= 8.0f) pr2 = (1.5f / sb + 27.0f / b2) / b2;
    return (pr1 * wt + pr2 * (1.0f - wt)) * corr;
}

// ---- inputs ----------------------------------------------------------------
// The window (src/synthe.f90:316-324) and the call's arguments.
struct Grid {
    int length, ixwlbeg, ifvac;
    double wbegin, ratio, ratiolg, wlend, cutoff, dopratio;
};

// Per depth: the module values compute_line_opacity and
// hydrogen_line_profile read (T, XNE rounded through REAL() as the driver
// stores them, src/synthe.f90:374-404; XNFPH(j,1:2) = REAL(qxnfpel(1:2))),
// EMERGE / EMERGEH (:3320-3360) and hydrogen_line_profile's depth arrays
// (:1216-1259).
struct Depth {
    double t, xne, hckt, rho, xnfph1, xnfph2, emerge, emergeh;
    float txnxn;
    float pp, fo, y1b, y1s, t3nhe, t3nh2, c1d, c2d, gcon1, gcon2;
};

// xf_* values of one depth (double, as run_xnfpelsyn leaves them) -> Depth.
A12_HD inline void depth_setup(double xf_t, double xf_xne, double xf_hckt, double xf_rho,
                               double xf_xnfh, double xf_xnfhe1, double xf_xnfh2, double qxnfpel1,
                               double qxnfpel2, Depth &D)
{
    const double T = (double)(float)xf_t, XNE = (double)(float)xf_xne;
    const double XNFH = (double)(float)xf_xnfh, XNFHE1 = (double)(float)xf_xnfhe1,
                 XNFH2 = (double)(float)xf_xnfh2;
    D.t = T;
    D.xne = XNE;
    D.hckt = xf_hckt;
    D.rho = (double)(float)xf_rho;
    D.xnfph1 = (double)(float)qxnfpel1;
    D.xnfph2 = (double)(float)qxnfpel2;
    // src/synthe.f90:548
    D.txnxn = (float)((XNFH + 0.42 * XNFHE1 + 0.85 * XNFH2) * pow(T / 10000.0, 0.3));
    // :3352-3358 (nmerge REAL, squared as REAL(8))
    const float inglis = (float)(1600.0 / pow(XNE, (double)(2.0f / 15.0f)));
    const float nmerge = inglis - 1.5f;
    D.emerge = RYDBERG_INF / ((double)nmerge * (double)nmerge);
    D.emergeh = RYDBERG_H / ((double)nmerge * (double)nmerge);
    // :1216-1259
    const float xne16 = (float)pow(XNE, 1.0 / 6.0);
    D.pp = (float)((double)(xne16 * 0.08989f) / sqrt(T));
    D.fo = xne16 * xne16 * xne16 * xne16 * 1.25E-9f;
    D.y1b = (float)(2.0 / (1.0 + (double)0.012f / T * sqrt(XNE / T)));
    const float t4 = (float)(T / 10000.0);
    const float t43 = powf(t4, 0.3f);
    D.y1s = t43 / xne16;
    D.t3nhe = (float)((double)t43 * XNFHE1);
    D.t3nh2 = (float)((double)t43 * XNFH2);
    D.c1d = (float)((double)(D.fo * 78940.0f) / T);
    D.c2d = (float)((double)(D.fo * D.fo / 5.96E-23f) / XNE);
    D.gcon1 = (float)((double)0.2f + (double)(0.09f * sqrtf(t4)) / (1.0 + XNE / (double)1.E13f));
    D.gcon2 = (float)((double)0.2f / (1.0 + XNE / (double)1.E15f));
}

// hydrogen_line_profile's per-transition constants (:1264-1364).
struct HLine {
    float xn, xm, xn2, xm2, xmn2, xm2mn2, gnm, xknm, y1num, y1wht, c1con, c2con;
    float radamp, resont, vdw, stark;
    double freqnm, dbeta, wavenm;
    int n, m, ifins;
    float finest[MAXFINE], finswt[MAXFINE];
};

A12_HD inline void hline_setup(int n, int m, HLine &H)
{
    H.n = n;
    H.m = m;
    const int mmn = m - n;
    H.xn = (float)n;
    H.xm = (float)m;
    H.xn2 = H.xn * H.xn;
    H.xm2 = H.xm * H.xm;
    H.xmn2 = H.xm2 * H.xn2;
    H.xm2mn2 = H.xm2 - H.xn2;
    H.gnm = H.xm2mn2 / H.xmn2;
    if (mmn <= 3 && n <= 4) H.xknm = XKNMTB[(mmn - 1) * 4 + (n - 1)];
    else H.xknm = 5.5E-5f / H.gnm * H.xmn2 / (1.0f + 0.13f / (float)mmn);
    H.y1num = 320.0f;
    if (m == 2) H.y1num = 550.0f;
    if (m == 3) H.y1num = 380.0f;
    H.y1wht = 1.E13f;
    if (mmn <= 3) H.y1wht = 1.E14f;
    if (mmn <= 2 && n <= 2) {
        if (n == 1 && mmn == 1) H.y1wht = 1.E18f;
        else if (n == 1 && mmn == 2) H.y1wht = 1.E17f;
        else if (n == 2 && mmn == 1) H.y1wht = 1.E16f;
        else if (n == 2 && mmn == 2) H.y1wht = 1.E14f;
    }
    H.freqnm = FREQ_RYDH * (double)H.gnm;
    H.dbeta = CLIGHT_ANGS / (H.freqnm * H.freqnm) / (double)H.xknm;
    H.wavenm = CLIGHT_ANGS / H.freqnm;
    H.c1con = (float)((double)H.xknm / H.wavenm * (double)H.gnm * (double)H.xm2mn2);
    const double kw = (double)H.xknm / H.wavenm;
    H.c2con = (float)(kw * kw);
    float radamp = ASUM[n - 1] + ASUM[m - 1];
    if (n == 1) radamp = ASUMLYMAN[m - 1];
    H.radamp = (float)((double)radamp / FOURPI / H.freqnm);
    float resont = hyd_osc1(m) / H.xm / (1.0f - 1.0f / H.xm2);
    if (n != 1) resont = resont + hyd_osc1(n) / H.xn / (1.0f - 1.0f / H.xn2);
    H.resont = resont * 3.579E-24f / H.gnm;
    H.vdw = 4.45E-26f / H.gnm * powf(H.xm2 * (7.0f * H.xm2 + 5.0f), 0.4f);
    H.stark = (float)((double)1.6678E-18f * H.freqnm * (double)H.xknm);
    for (int i = 0; i < MAXFINE; ++i) H.finest[i] = H.finswt[i] = 0.0f;
    if (n > 4 || m > 10) {
        H.ifins = 1;
        H.finest[0] = 0.0f;
        H.finswt[0] = 1.0f;
    } else if (mmn == 1) {
        H.ifins = LNGHAL[n - 1];
        const int ipos = ISTAL[n - 1];
        for (int i = 0; i < H.ifins; ++i) {
            H.finest[i] = STALPH[ipos - 1 + i] * 1.E7f;
            H.finswt[i] = STWTAL[ipos - 1 + i] / H.xn2 / 3.0f;
        }
    } else {
        H.ifins = LNCOMP[n - 1];
        for (int i = 0; i < H.ifins; ++i) {
            H.finest[i] = STCOMP[(n - 1) * 5 + i] * 1.E7f;
            H.finswt[i] = STCPWT[(n - 1) * 5 + i] / H.xn2;
        }
    }
}

// One Stehle series (STEHLE_DATA(s), src/atlas12_modules.f90:1929-1945).
struct StarkSeries {
    int loaded, nmin, nmax, nd;
    const double *dens, *temp, *lg, *kalpha, *prof;  // prof (ntrans, nd, 10, 60)
    const double *dg;  // 10**lg (the DALPHA_GRID and the bounds the profile tests)
    const int *maxd;
};
struct Stark {
    StarkSeries s[4];
};
// A view of the four series of the boundary's Stehle struct (host or
// device pointers alike); dg: (4, 60) = 10**log_dalpha per series.
inline void stark_view(const a12_kapp_stark *st, const double *dg, Stark &S)
{
    for (int i = 0; i < 4; ++i) {
        StarkSeries &s = S.s[i];
        s.loaded = st[i].loaded;
        s.nmin = st[i].n_upper_min;
        s.nmax = st[i].n_upper_max;
        s.nd = st[i].n_dens;
        s.dens = st[i].dens;
        s.temp = st[i].temp;
        s.lg = st[i].log_dalpha;
        s.kalpha = st[i].k_alpha;
        s.prof = st[i].profiles;
        s.maxd = st[i].max_dens_idx;
        s.dg = dg + i * NDA;
    }
}
// 10**log_dalpha of the four series (host libm, as the reference forms them).
inline void stark_pow10(const a12_kapp_stark *st, double *dg)
{
    for (int i = 0; i < 4; ++i)
        for (int ia = 0; ia < NDA; ++ia) dg[i * NDA + ia] = st[i].loaded ? pow(10.0, st[i].log_dalpha[ia]) : 0.0;
}

// The Stehle path is taken for n <= 4 inside the series' upper range at or
// below its highest density (:1378-1408).
A12_HD inline bool use_stehle(const Stark &S, int n, int m, double xne)
{
    if (n < 1 || n > 4) return false;
    const StarkSeries &s = S.s[n - 1];
    if (!s.loaded || m < s.nmin || m > s.nmax) return false;
    return xne <= s.dens[s.nd - 1];
}

// The cached convolution of one (transition, depth) (:1422-1640).
struct HConv {
    double prof[NDA], lprof[NDA];  // lprof: LOG(prof) where prof > 0
    double F0, lambda0, dop_dalpha;
    float voigt_a;
    int pad;
};

// First i in [0, n-1) with x <= g[i+1]; n-2 when there is none (the
// reference's DO ... EXIT search, clamped to the last interval).
A12_HD inline int search_le(const double *g, int n, double x)
{
    int lo = 0, hi = n - 2;  // answer in [0, n-2]
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (x <= g[mid + 1]) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// interp_stark_at (:2007-2037)
A12_HD inline double interp_stark_at(double da_pt, const double *lg, const double *dg, const double *sp)
{
    if (da_pt >= dg[NDA - 1]) return sp[NDA - 1];
    const double log_da = log10(da_pt);
    const int idx = search_le(lg, NDA, log_da);
    const double frac = (log_da - lg[idx]) / (lg[idx + 1] - lg[idx]);
    if (sp[idx] > 0.0 && sp[idx + 1] > 0.0)
        return exp(log(sp[idx]) * (1.0 - frac) + log(sp[idx + 1]) * frac);
    return sp[idx] * (1.0 - frac) + sp[idx + 1] * frac;
}

// The widths of one call (:1370-1378): Stark, van der Waals, radiation,
// resonance, their Lorentzian sum.
struct HWidths {
    float hwstk, hwvdw, hwrad, hwres, hwlor;
};
A12_HD inline HWidths h_widths(const HLine &H, const Depth &D)
{
    HWidths w;
    w.hwstk = H.stark * D.fo;
    w.hwvdw = H.vdw * D.t3nhe + 2.0f * H.vdw * D.t3nh2;
    w.hwrad = H.radamp;
    w.hwres = (float)((double)H.resont * D.xnfph1 * 2.0);
    w.hwlor = w.hwres + w.hwvdw + w.hwrad;
    return w;
}

// Convolution grid of one (transition, depth): the interpolated Stehle
// profile, the Doppler/Lorentz widths in delta-alpha units, and the grid
// sizes (:1424-1566).  conv_params sets the scalars, conv_fill one entry
// of the interpolated profile, conv_sizes the grid's extent (after every
// conv_fill); conv_grid_point() gives the grid's entries.
struct ConvSetup {
    double sp[NDA];
    const double *prof4[4];  // the four bracketing table rows
    const double *lg, *dg;
    double F0, lambda0, frac_d, frac_t, dop_dalpha, dx_wing, da_trans, da_wing_max;
    float voigt_a;
    int n_log_use, n_wing, ic;
};

A12_HD inline void conv_params(const HLine &H, const Depth &D, const Stark &S, float hwdop, ConvSetup &C)
{
    const int n = H.n, m = H.m;
    const StarkSeries &s = S.s[n - 1];
    C.lg = s.lg;
    C.dg = s.dg;
    const double xn8 = n, xm8 = m;
    const double nm = xn8 * xm8;
    C.lambda0 = RYD_ANG * (nm * nm) / ((xm8 - xn8) * (xm8 + xn8));
    const int itrans = m - s.nmin + 1;
    const double dtop = s.dens[s.maxd[itrans - 1] - 1];
    const double necap = D.xne < dtop ? D.xne : dtop;
    C.F0 = 1.25e-9 * pow(necap, 2.0 / 3.0);
    const int nd = s.nd;
    const double log_ne = log10(necap);
    int id1 = nd;
    for (int i = 1; i <= nd - 1; ++i)
        if (necap <= s.dens[i]) {
            id1 = i;
            break;
        }
    id1 = id1 > nd - 1 ? nd - 1 : id1;
    id1 = id1 < 1 ? 1 : id1;
    const int id2 = id1 + 1;
    double frac_d = (log_ne - log10(s.dens[id1 - 1])) / (log10(s.dens[id2 - 1]) - log10(s.dens[id1 - 1]));
    frac_d = frac_d < 1.0 ? frac_d : 1.0;
    C.frac_d = frac_d > 0.0 ? frac_d : 0.0;
    int it1 = NTEMP;
    for (int i = 1; i <= NTEMP - 1; ++i)
        if (D.t <= s.temp[i]) {
            it1 = i;
            break;
        }
    it1 = it1 > NTEMP - 1 ? NTEMP - 1 : it1;
    it1 = it1 < 1 ? 1 : it1;
    const int it2 = it1 + 1;
    double frac_t = (D.t - s.temp[it1 - 1]) / (s.temp[it2 - 1] - s.temp[it1 - 1]);
    frac_t = frac_t < 1.0 ? frac_t : 1.0;
    C.frac_t = frac_t > 0.0 ? frac_t : 0.0;
    const double *pt = s.prof + (size_t)(itrans - 1) * nd * NTEMP * NDA;
    C.prof4[0] = pt + ((size_t)(id1 - 1) * NTEMP + (it1 - 1)) * NDA;
    C.prof4[1] = pt + ((size_t)(id1 - 1) * NTEMP + (it2 - 1)) * NDA;
    C.prof4[2] = pt + ((size_t)(id2 - 1) * NTEMP + (it1 - 1)) * NDA;
    C.prof4[3] = pt + ((size_t)(id2 - 1) * NTEMP + (it2 - 1)) * NDA;
    const HWidths w = h_widths(H, D);
    C.dop_dalpha = C.lambda0 * C.lambda0 / CLIGHT_ANGS * H.freqnm * (double)hwdop / C.F0;
    float hwlor_conv = w.hwlor;
    if (n == 1 && m == 2) hwlor_conv = w.hwres * 4.0f + w.hwvdw + w.hwrad;
    const double lor_dalpha = C.lambda0 * C.lambda0 / CLIGHT_ANGS * H.freqnm * (double)hwlor_conv / C.F0;
    C.voigt_a = (float)(lor_dalpha / C.dop_dalpha);
    C.dx_wing = C.dop_dalpha / 3.0;
}

// The bilinear (log Ne, T) interpolation of the table at delta-alpha ia.
A12_HD inline void conv_fill(ConvSetup &C, int ia)
{
    const double fd = C.frac_d, ft = C.frac_t;
    const double v00 = C.prof4[0][ia], v01 = C.prof4[1][ia], v10 = C.prof4[2][ia], v11 = C.prof4[3][ia];
    if (v00 > 0 && v01 > 0 && v10 > 0 && v11 > 0)
        C.sp[ia] = exp(log(v00) * (1.0 - fd) * (1.0 - ft) + log(v01) * (1.0 - fd) * ft +
                       log(v10) * fd * (1.0 - ft) + log(v11) * fd * ft);
    else
        C.sp[ia] = v00 * (1.0 - fd) * (1.0 - ft) + v01 * (1.0 - fd) * ft + v10 * fd * (1.0 - ft) +
                   v11 * fd * ft;
}

A12_HD inline void conv_sizes(ConvSetup &C)
{
    const double d_log = (C.lg[NDA - 1] - C.lg[0]) / (double)(NDA - 1);
    int nlu = NDA;
    for (int ia = 1; ia <= NDA; ++ia)
        if (C.dg[ia - 1] * d_log * LN10 > C.dx_wing) {
            nlu = ia - 1;
            break;
        }
    nlu = nlu > 1 ? nlu : 1;
    C.n_log_use = nlu;
    C.da_trans = C.dg[nlu - 1];
    int n_wing = NWING_MAX;
    C.da_wing_max = C.da_trans + (double)n_wing * C.dx_wing;
    if (C.da_wing_max > C.dg[NDA - 1]) {
        C.da_wing_max = C.dg[NDA - 1];
        n_wing = (int)((C.da_wing_max - C.da_trans) / C.dx_wing);
        if (n_wing < 0) n_wing = 0;
    }
    if (2 * n_wing + 2 * nlu + 1 > NCONV_MAX) {
        n_wing = (NCONV_MAX - 2 * nlu - 1) / 2;
        C.da_wing_max = C.da_trans + (double)n_wing * C.dx_wing;
    }
    C.n_wing = n_wing;
    C.ic = 2 * n_wing + 2 * nlu + 1;
}

// Entry c (0-based) of the symmetric grid [-wing][-log][0][+log][+wing]
// (:1570-1605): its delta-alpha and profile value.
A12_HD inline void conv_grid_point(const ConvSetup &C, int c, double &da, double &pr)
{
    const int nw = C.n_wing, nl = C.n_log_use;
    if (c < nw) {
        const double da_pt = C.da_trans + (double)(nw - c) * C.dx_wing;
        da = -da_pt;
        pr = interp_stark_at(da_pt, C.lg, C.dg, C.sp);
    } else if (c < nw + nl) {
        const int ia = nl - (c - nw);  // 1-based
        da = -C.dg[ia - 1];
        pr = C.sp[ia - 1];
    } else if (c == nw + nl) {
        da = 0.0;
        pr = C.sp[0];
    } else if (c <= nw + 2 * nl) {
        const int ia = c - (nw + nl);
        da = C.dg[ia - 1];
        pr = C.sp[ia - 1];
    } else {
        const int iw = c - (nw + 2 * nl);
        const double da_pt = C.da_trans + (double)iw * C.dx_wing;
        da = da_pt;
        pr = interp_stark_at(da_pt, C.lg, C.dg, C.sp);
    }
}

// Trapezoid weight of entry c from the grid's delta-alphas (:1608-1612).
A12_HD inline double conv_step(const double *da, int c, int ic)
{
    if (c == 0) return da[1] - da[0];
    if (c == ic - 1) return da[ic - 1] - da[ic - 2];
    return 0.5 * (da[c + 1] - da[c - 1]);
}

// Output ia of the convolution (:1615-1630), summed over the grid in order.
A12_HD inline double conv_output(const ConvSetup &C, const double *da, const double *pr,
                                 const double *step, int ia)
{
    const double da_i = C.dg[ia];
    if (da_i > C.da_wing_max && C.da_wing_max > 0.0)
        return C.sp[ia] + (double)C.voigt_a * C.dop_dalpha / (PI * (da_i * da_i));
    double gsum = 0.0;
    for (int c = 0; c < C.ic; ++c) {
        const float v = (float)((da_i - da[c]) / C.dop_dalpha);
        const double k = (double)voigt_profile(v, C.voigt_a);
        gsum = gsum + pr[c] * k * step[c];
    }
    return gsum / (C.dop_dalpha * SQRTPI);
}

// lya_h2_blue_cutoff (:1916-1940).  The table index is held inside the
// table at the window's upper end, where the reference reads one past it.
A12_HD inline float lya_h2_blue(double freq, float xnfph_h)
{
    const float spc = 200.0f * CLIGHT4;
    const float freq22 = (82259.10f - 22000.0f) * CLIGHT4;
    if (freq < (double)(50000.0f * CLIGHT4)) return 0.0f;
    float cval;
    if (freq < (double)freq22) {
        cval = (float)((double)((CUTOFFH2[1] - CUTOFFH2[0]) / spc) * (freq - (double)freq22) + (double)CUTOFFH2[0]);
    } else {
        const int ic = (int)((freq - (double)freq22) / (double)spc);
        const float cfreq = (float)ic * spc + freq22;
        const int i1 = ic < 90 ? ic : 90, i2 = ic + 1 < 90 ? ic + 1 : 90;
        cval = (float)((double)((CUTOFFH2[i2] - CUTOFFH2[i1]) / spc) * (freq - (double)cfreq) + (double)CUTOFFH2[i1]);
    }
    return powf(10.0f, cval - 14.0f) * xnfph_h * 2.0f / CLIGHT4;
}

// lya_h2plus_red_cutoff (:1956-1978), index held as above.
A12_HD inline float lya_h2plus_red(double freq, float xnfph_h2plus)
{
    const float spc = 100.0f * CLIGHT4;
    const float freq15 = (82259.10f - 15000.0f) * CLIGHT4;
    float cval;
    if (freq < (double)freq15) {
        cval = (float)((double)((CUTOFFH2PLUS[1] - CUTOFFH2PLUS[0]) / spc) * (freq - (double)freq15) +
                       (double)CUTOFFH2PLUS[0]);
    } else {
        const int ic = (int)((freq - (double)freq15) / (double)spc);
        const float cfreq = (float)ic * spc + freq15;
        const int i1 = ic < 110 ? ic : 110, i2 = ic + 1 < 110 ? ic + 1 : 110;
        cval = (float)((double)((CUTOFFH2PLUS[i2] - CUTOFFH2PLUS[i1]) / spc) * (freq - (double)cfreq) +
                       (double)CUTOFFH2PLUS[i1]);
    }
    return powf(10.0f, cval - 14.0f) / CLIGHT4 * xnfph_h2plus;
}

// hydrogen_line_profile (:1032-1980) at offset delw (nm) from the line,
// with hwdop = DOPPH(j).  cv: the convolution the reference's cache holds
// for this (n, m, j) on the Stehle path (unused on the analytic path).
A12_HD inline float h_profile(const HLine &H, const Depth &D, const Stark &S, bool stehle,
                              const HConv *cv, double delw, float hwdop)
{
    const int n = H.n, m = H.m;
    const double wl4 = H.wavenm + delw * 10.0;
    const double freq = CLIGHT_ANGS / wl4;
    const double del = fabs(freq - H.freqnm);
    const HWidths w = h_widths(H, D);
    const double dop = H.freqnm * (double)hwdop;
    const double f4000 = (double)((82259.10f - 4000.0f) * CLIGHT4);
    const double f20000 = (double)((82259.10f - 20000.0f) * CLIGHT4);
    if (stehle) {
        const StarkSeries &s = S.s[n - 1];
        const double dalpha = cv->lambda0 * cv->lambda0 / CLIGHT_ANGS * del / cv->F0;
        double I;
        if (dalpha <= s.dg[0]) {
            I = cv->prof[0];
        } else if (dalpha >= s.dg[NDA - 1]) {
            I = s.kalpha[m - s.nmin] / pow(dalpha, 2.5) +
                (double)cv->voigt_a * cv->dop_dalpha / (PI * (dalpha * dalpha));
        } else {
            const double lda = log10(dalpha);
            const int i1 = search_le(s.lg, NDA, lda);
            const double fa = (lda - s.lg[i1]) / (s.lg[i1 + 1] - s.lg[i1]);
            if (cv->prof[i1] > 0.0 && cv->prof[i1 + 1] > 0.0)
                I = exp(cv->lprof[i1] * (1.0 - fa) + cv->lprof[i1 + 1] * fa);
            else
                I = cv->prof[i1] * (1.0 - fa) + cv->prof[i1 + 1] * fa;
        }
        const double phi = I * (cv->lambda0 * cv->lambda0) / (CLIGHT_ANGS * cv->F0);
        float result = (float)(phi * SQRTPI * dop);
        if (n == 1 && m == 2) {
            if (freq <= f4000) result = result + lya_h2_blue(freq, (float)D.xnfph1) * (float)(SQRTPI * dop);
            if (freq >= f20000 && freq <= f4000)
                result = result + lya_h2plus_red(freq, (float)D.xnfph2) * (float)(SQRTPI * dop);
            if (freq < (double)(50000.0f * CLIGHT4)) result = 0.0f;
        }
        return result;
    }
    // Peterson/Kurucz analytic path (:1720-1912)
    const float hwdop_ = hwdop;
    int nwid = 1;
    if (hwdop_ < w.hwstk || hwdop_ < w.hwlor) {
        nwid = 2;
        if (w.hwlor < w.hwstk) nwid = 3;
    }
    float mx = hwdop_;
    mx = w.hwlor > mx ? w.hwlor : mx;
    mx = w.hwstk > mx ? w.hwstk : mx;
    const double hfwid = H.freqnm * (double)mx;
    float result = 0.0f;
    const bool core = del <= hfwid;
    if (!core || nwid == 1) {
        for (int i = 0; i < H.ifins; ++i) {
            const float d = (float)(fabs(freq - H.freqnm - (double)H.finest[i]) / dop);
            if (d <= 7.0f) result = result + expf(-d * d) * H.finswt[i];
        }
        if (core) return result;
    }
    if (!core || nwid == 2) {
        if (n == 1 && m == 2) {
            const float hwres4 = w.hwres * 4.0f;
            const float hwlor = hwres4 + w.hwvdw + w.hwrad;
            const double hhw = H.freqnm * (double)hwlor;
            float hprofres;
            if (freq > f4000)
                hprofres = (float)((double)hwres4 * H.freqnm / PI / (del * del + hhw * hhw) * SQRTPI * dop);
            else
                hprofres = lya_h2_blue(freq, (float)D.xnfph1) * (float)(SQRTPI * dop);
            float hprofrad = 0.0f;
            if (freq > (double)2.4190611E15f && freq < (double)0.77f * FREQ_RYDH)
                hprofrad = (float)((double)w.hwrad * H.freqnm / PI / (del * del + hhw * hhw) * SQRTPI * dop);
            float hprofvdw = (float)((double)w.hwvdw * H.freqnm / PI / (del * del + hhw * hhw) * SQRTPI * dop);
            if (freq < (double)1.8E15f) hprofvdw = 0.0f;
            result = result + hprofres + hprofrad + hprofvdw;
            if (core) return result;
        } else {
            const double hhw = H.freqnm * (double)w.hwlor;
            float top = (float)hhw;
            if (n == 1 && m <= 5) {
                const double rad = H.freqnm * (double)w.hwrad;
                if (m == 3 && freq > (double)0.885f * FREQ_RYDH && freq < (double)0.890f * FREQ_RYDH)
                    top = (float)(hhw - rad);
                if (m == 4 && freq > (double)0.936f * FREQ_RYDH && freq < (double)0.938f * FREQ_RYDH)
                    top = (float)(hhw - rad);
                if (m == 5 && freq > (double)0.959f * FREQ_RYDH && freq < (double)0.961f * FREQ_RYDH)
                    top = (float)(hhw - rad);
            }
            result = result + (float)((double)top / PI / (del * del + hhw * hhw) * SQRTPI * dop);
            if (core) return result;
        }
    }
    const float wty1 = (float)(1.0 / (1.0 + D.xne / (double)H.y1wht));
    const float y1scal = H.y1num * D.y1s * wty1 + D.y1b * (1.0f - wty1);
    const float c1 = D.c1d * H.c1con * y1scal;
    const float c2 = D.c2d * H.c2con;
    const float g1 = 6.77f * sqrtf(c1);
    float lg = 0.2114f + logf(sqrtf(c2) / c1);
    lg = 0.0f > lg ? 0.0f : lg;
    const float gnot = g1 * lg * (1.0f - D.gcon1 - D.gcon2);
    const float beta = (float)(fabs(del) / (double)D.fo * H.dbeta);
    const float y1 = c1 * beta;
    const float y2 = c2 * (beta * beta);
    float gam = gnot;
    if (!(y2 <= 1.E-4f && y1 <= 1.E-5f)) {
        const float t90 = 90.0f * y1;
        const float my1 = 80.0f < y1 ? 80.0f : y1;
        gam = g1 * (0.5f * expf(-my1) + expint1(y1) - 0.5f * expint1(y2)) *
              (1.0f - D.gcon1 / (1.0f + t90 * t90 * t90) - D.gcon2 / (1.0f + 2000.0f * y1));
        if (gam <= 1.E-20f) gam = 0.0f;
    }
    float prqs = stark_qs(beta, D.pp, n, m);
    if (m == 2) {
        prqs = prqs * 0.5f;
        if (freq >= f20000 && freq <= f4000) {
            result = result + lya_h2plus_red(freq, (float)D.xnfph2) * (float)(SQRTPI * dop);
        } else if (freq > f4000) {
            const float b4000 = (float)((double)(4000.0f * CLIGHT4 / D.fo) * H.dbeta);
            const float p4000 = (float)((double)(stark_qs(b4000, D.pp, n, m) * 0.5f / D.fo) * H.dbeta);
            const float c4000 = (float)((double)(powf(10.0f, -11.07f - 14.0f) / CLIGHT4) * D.xnfph2);
            result = result + (float)((double)(c4000 / p4000 * prqs / D.fo) * H.dbeta * SQRTPI * dop);
        }
    }
    float f = 0.0f;
    if (gam > 0.0f) f = (float)((double)gam / PI / (double)(gam * gam + beta * beta));
    const float p1 = (0.9f * y1) * (0.9f * y1);
    const float fns = (p1 + 0.03f * sqrtf(y1)) / (p1 + 1.0f);
    return result + (float)((double)((prqs * (1.0f + fns) + f) / D.fo) * H.dbeta * SQRTPI * dop);
}

// ---- per (line, depth) ------------------------------------------------------
enum Kind : int {
    K_SKIP = 0,   // no contribution at this depth
    K_METAL = 1,  // itype 0, 3, <= -3 (:3402-3458)
    K_HISO = 2,   // H/D alpha or NCON = 0 (:3485-3514)
    K_HBETA = 3,  // H beta with a merge (:3516-3550)
    K_HMERGE = 4, // H with the merged continuum (:3552-3621)
    K_AUT = 5,    // autoionising (:3623-3670)
    K_CONT = 6    // merged-continuum edge, itype >= 4 (:3672-3720)
};

struct Plan {
    int kind;
    int red;      // the red wing is walked
    int minred;   // first red point
    int ib0;      // MIN(length+1, nbuff): the blue walk covers ib0-1 .. 1
    int blue_if_red_done;  // blue runs when the red walk reached its end (no early stop)
    int blue_if_red_stop;  // blue runs when the red walk stopped early
    float kappa0;
    float adamp, dopwl;    // metal
    float hwdop;           // hydrogen: DOPPH(j)
    float frelin;          // autoionising
    double wl, wcon, wtail, alpha_lim;
    // merged-continuum edge
    int nb1, nb2, nb3;
    float dnb, ckappa;
};

A12_HD inline int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

// XNFPEL of the driver (src/synthe.f90:536-546) from qxnfpel: cleared at
// 1e25, divided by RHO.
A12_HD inline double xnfpel_of(double q, double rho) { return (q < 1.0e25 ? q : 0.0) / rho; }

// The set-up of one line at one depth: everything before the wing loops.
// xnfdop and dopple are the depth's REAL(8) arrays of the driver
// (src/synthe.f90:536-546), xnfpel_nel its XNFPEL(NELION); cont the depth's
// FP32 continuum.
A12_HD inline void plan_line(const a12_cplx_line &ln, const Grid &G, const Depth &D,
                             const double *xnfdop, const double *dopple, double xnfpel_nel,
                             const float *cont, Plan &P)
{
    const int L = G.length;
    P.kind = K_SKIP;
    P.red = 0;
    P.minred = ln.nbuff > 1 ? ln.nbuff : 1;
    P.ib0 = ln.nbuff < L + 1 ? ln.nbuff : L + 1;
    P.blue_if_red_done = P.blue_if_red_stop = 1;
    P.kappa0 = 0.0f;
    P.adamp = P.dopwl = P.hwdop = P.frelin = 0.0f;
    P.wl = ln.wlvac * G.dopratio;
    P.wcon = P.wtail = P.alpha_lim = 0.0;
    P.nb1 = P.nb2 = P.nb3 = 0;
    P.dnb = P.ckappa = 0.0f;
    const int itype = ln.itype;
    const double wl = P.wl;
    const float kapmin = (float)((double)cont[clampi(ln.nbuff, 1, L) - 1] * G.cutoff);
    if (itype == 2) return;
    if (itype == 0 || itype == 3 || itype <= -3) {
        float kappa0 = (float)((double)ln.gf * xnfdop[ln.nelion - 1]);
        if (kappa0 < kapmin) return;
        const float bolt = (float)exp((double)(-ln.elo) * D.hckt);
        kappa0 = kappa0 * bolt;
        if (kappa0 < kapmin) return;
        P.kind = K_METAL;
        P.kappa0 = kappa0;
        if (ln.ncon > 0) {
            const double c = contx(ln.ncon, ln.nelionx);
            P.wcon = 1.0e7 / (c - D.emerge) * G.dopratio;
            P.wtail = 1.0e7 / (c - D.emerge - 500.0) * G.dopratio;
        }
        const double dop = dopple[ln.nelion - 1];
        P.adamp = (float)(((double)ln.gammar + (double)ln.gammas * D.xne + (double)(ln.gammaw * D.txnxn)) / dop);
        P.dopwl = (float)dop * (float)wl;
        P.red = wl <= G.wlend;
        if (P.red && (P.minred == 1 || wl < G.wbegin)) P.blue_if_red_done = P.blue_if_red_stop = 0;
        return;
    }
    if (itype == -1 || itype == -2) {
        const int nblo = ln.nblo, nbup = ln.nbup;
        float kappa0 = (float)((double)ln.gf * xnfdop[0]);
        float kmin = kapmin;
        if (nbup == 2) kmin = (float)((double)cont[L - 1] * G.cutoff);
        if (kappa0 < kmin) return;
        const float bolth = (float)exp((double)(-ln.elo) * D.hckt);
        P.hwdop = (float)dopple[0];
        if (itype == -2) P.hwdop = P.hwdop / 1.4142f;
        kappa0 = kappa0 * bolth;
        if (kappa0 < kmin) return;
        P.kappa0 = kappa0;
        P.alpha_lim = alphahyd(nblo);
        if (ln.ncon == 0 || nbup == nblo + 1) {
            P.kind = K_HISO;
            P.red = wl <= G.wlend;
            if (P.red && (P.minred == 1 || wl < G.wbegin)) P.blue_if_red_done = P.blue_if_red_stop = 0;
        } else if (nbup == nblo + 2) {
            P.kind = K_HBETA;
            P.red = wl <= G.wlend;
            if (P.red) {
                // a red walk that ends early jumps straight to the blue wing
                P.blue_if_red_stop = 1;
                P.blue_if_red_done = !(P.minred == 1 || wl < G.wbegin || ln.nbuff < 1);
            } else {
                P.blue_if_red_done = P.blue_if_red_stop = ln.nbuff >= 1;
            }
        } else {
            P.kind = K_HMERGE;
            const double conth = H_SERIES_LIMITS[ln.ncon - 1];
            const double wshift = 1.0e7 / (conth - RYDBERG_H / (81.0 * 81.0));
            double wmerge = 1.0e7 / (conth - D.emergeh);
            if (wmerge < 0.0) wmerge = wshift + wshift;
            double wcon = wshift > wmerge ? wshift : wmerge;
            double wtail = 1.0e7 / (1.0e7 / wcon - 500.0);
            wcon = wshift + wshift < wcon ? wshift + wshift : wcon;
            if (wtail < 0.0) wtail = wcon + wcon;
            wtail = wcon + wcon < wtail ? wcon + wcon : wtail;
            if (G.ifvac == 0) {
                wcon = vac_to_air(wcon);
                wtail = vac_to_air(wtail);
            }
            wcon = wcon * G.dopratio;
            P.wcon = wcon;
            P.wtail = wtail;
            if (G.ifvac == 0) P.wl = vac_to_air(1.0e7 / (ehyd(nbup) - ehyd(nblo))) * G.dopratio;
            if (P.wl <= G.wlend) {
                if (wcon > G.wlend) {
                    P.kind = K_SKIP;
                    return;
                }
                P.red = G.wbegin <= P.alpha_lim;
                bool noblue = P.wl < G.wbegin;
                if (P.red && P.minred == 1) noblue = true;
                if (noblue) P.blue_if_red_done = P.blue_if_red_stop = 0;
            }
        }
        return;
    }
    if (itype == 1) {
        float kappa0 = ln.gammaw * ln.gf * (float)xnfpel_nel;
        if (kappa0 < kapmin) return;
        kappa0 = kappa0 * (float)exp((double)(-ln.elo) * D.hckt);
        if (kappa0 < kapmin) return;
        P.kind = K_AUT;
        P.kappa0 = kappa0;
        P.frelin = (float)(CLIGHT_NMS / wl);
        P.red = wl <= G.wlend;
        if (P.red && (ln.nbuff == 1 || wl < G.wbegin)) P.blue_if_red_done = P.blue_if_red_stop = 0;
        return;
    }
    // itype >= 4: merged continuum, NLAST = itype
    const double it2 = (double)itype * (double)itype;
    double wshift = 1.0e7 / (1.0e7 / wl - RYDBERG_INF / it2);
    double wmerge = 1.0e7 / (1.0e7 / wl - D.emerge);
    if (ln.nelion == 1) {
        wshift = 1.0e7 / (1.0e7 / wl - RYDBERG_H / it2);
        wmerge = 1.0e7 / (1.0e7 / wl - D.emergeh);
    }
    if (wmerge < 0.0) wmerge = wshift + wshift;
    wmerge = wmerge > wshift ? wmerge : wshift;
    wmerge = wshift + wshift < wmerge ? wshift + wshift : wmerge;
    double wtail = 1.0e7 / (1.0e7 / wmerge - 500.0);
    if (wtail < 0.0) wtail = wmerge + wmerge;
    wtail = wmerge + wmerge < wtail ? wmerge + wmerge : wtail;
    if (G.ifvac == 0) {
        wmerge = vac_to_air(wmerge) * G.dopratio;
        wtail = vac_to_air(wtail) * G.dopratio;
    }
    int ixwl = (int)(log(wl) / G.ratiolg);
    const float edgeblue = expf((float)ixwl * (float)G.ratiolg);
    if (edgeblue > (float)wl) ixwl = ixwl - 1;
    int nb1 = ixwl + 1 - G.ixwlbeg + 1;
    ixwl = (int)(log(wmerge) / G.ratiolg + 0.5);
    const int nb2 = ixwl - G.ixwlbeg + 1;
    ixwl = (int)(log(wtail) / G.ratiolg + 0.5);
    const int nb3 = ixwl - G.ixwlbeg + 1;
    if (nb1 > L || nb3 < 1) return;
    P.kind = K_CONT;
    P.dnb = (float)(nb3 - nb2);
    P.nb1 = nb1 > 1 ? nb1 : 1;
    P.nb2 = nb2;
    P.nb3 = nb3;
    P.ckappa = ln.gf * (float)xnfpel_nel * (float)exp((double)(-ln.elo) * D.hckt);
}

// Whether the line calls hydrogen_line_profile at this depth (its
// per-transition cache then holds this line's convolution): wr0 / wb0 are
// the first red / blue wing wavelengths.
A12_HD inline bool calls_profile(const Plan &P, const Grid &G, double wr0, double wb0)
{
    if (P.kind != K_HISO && P.kind != K_HBETA && P.kind != K_HMERGE) return false;
    int blue = P.blue_if_red_done;  // a red walk that evaluated nothing ran to its end ...
    if (P.red) {
        if (P.kind == K_HISO && P.minred <= G.length) return true;
        if (P.kind == K_HBETA && P.minred <= G.length) {
            if (wr0 <= P.alpha_lim) return true;
            blue = P.blue_if_red_stop;  // ... except at the alpha limit, which ends it at once
        }
        if (P.kind == K_HMERGE && wr0 >= P.wcon) return true;  // the pre-check at minred
    }
    if (!blue || P.ib0 - 1 < 1) return false;
    return P.kind != K_HMERGE || wb0 >= P.wcon;
}

// ---- wing chains ------------------------------------------------------------
// The red walk's wavelength at walk index k is wr(k): wr(0) = wbegin *
// ratio**(minred-1), wr(k+1) = wr(k) * ratio; the blue walk's wb(0) =
// wbegin * ratio**(ib0-1) / ratio, wb(k+1) = wb(k) / ratio.  The
// autoionising branch walks REAL(4) frequencies instead: fr(0) =
// REAL(CLIGHT_NMS / wr(0)), fr(k+1) = fr(k) / REAL(ratio); fb(0) =
// REAL(CLIGHT_NMS / (wbegin * ratio**(ib0-1))) * REAL(ratio), fb(k+1) =
// fb(k) * REAL(ratio).  Chains are kept at every CK-th index and re-run from
// there.
constexpr int CK = 32;

A12_HD inline double chain_start(const Grid &G, int aut, int red, int minred, int ib0)
{
    if (red) {
        const double w = G.wbegin * powi_rt(G.ratio, minred - 1);
        return aut ? (double)(float)(CLIGHT_NMS / w) : w;
    }
    const double w = G.wbegin * powi_rt(G.ratio, ib0 - 1);
    return aut ? (double)((float)(CLIGHT_NMS / w) * (float)G.ratio) : w / G.ratio;
}
A12_HD inline double chain_next(const Grid &G, int aut, int red, double v)
{
    if (aut) return red ? (double)((float)v / (float)G.ratio) : (double)((float)v * (float)G.ratio);
    return red ? v * G.ratio : v / G.ratio;
}

// ---- one wing point -----------------------------------------------------------
struct Step {
    float kap;
    int add, stop;
};

// The untapered kappa of a metal / hydrogen / autoionising point at chain
// value v (wavelength, or frequency for autoionising lines).
A12_HD inline float point_raw(const Plan &P, const a12_cplx_line &ln, const HLine *H, const Depth &D,
                              const Stark &S, bool stehle, const HConv *cv, double v)
{
    switch (P.kind) {
    case K_METAL: {
        const float vv = fabsf((float)(v - P.wl)) / P.dopwl;
        return P.kappa0 * voigt_profile(vv, P.adamp);
    }
    case K_HISO:
    case K_HBETA:
    case K_HMERGE:
        return P.kappa0 * h_profile(*H, D, S, stehle, cv, v - P.wl, P.hwdop);
    case K_AUT: {
        const float f = (float)v;
        const float eps = 2.0f * (f - P.frelin) / ln.gammar;
        return P.kappa0 * (ln.gammas * eps + ln.gammaw) / (eps * eps + 1.0f) / ln.gammaw;
    }
    default:
        return 0.0f;
    }
}
A12_HD inline float taper(const Plan &P, double w, float kap)
{
    if (w < P.wtail) kap = kap * (float)((w - P.wcon) / (P.wtail - P.wcon));
    return kap;
}
// The value added at a wing point (tapered where the branch tapers).
A12_HD inline float point_kappa(const Plan &P, const a12_cplx_line &ln, const HLine *H, const Depth &D,
                                const Stark &S, bool stehle, const HConv *cv, double v)
{
    const float k = point_raw(P, ln, H, D, S, stehle, cv, v);
    return (P.kind == K_METAL || P.kind == K_HMERGE) ? taper(P, v, k) : k;
}
// The merged-continuum edge's value at point p (:3712-3716).
A12_HD inline float cont_kappa(const Plan &P, int p)
{
    float tail = 1.0f;
    if (p > P.nb2) tail = (float)(P.nb3 - p) / P.dnb;
    return P.ckappa * tail;
}

// Red walk index k (point minred + k <= length), chain value v; thr =
// continuum(point) * cutoff in REAL(8), thr0 the merge pre-check's
// threshold at MIN(minred, length).
A12_HD inline Step red_step(const Plan &P, const a12_cplx_line &ln, const HLine *H, const Depth &D,
                            const Stark &S, bool stehle, const HConv *cv, int k, double v, double thr,
                            double thr0)
{
    Step s{0.0f, 0, 0};
    switch (P.kind) {
    case K_METAL:
        if (v >= P.wcon) {
            s.kap = taper(P, v, point_raw(P, ln, H, D, S, stehle, cv, v));
            if ((double)s.kap < thr) s.stop = 1;
            else s.add = 1;
        }
        return s;
    case K_HISO:
    case K_AUT:
        s.kap = point_raw(P, ln, H, D, S, stehle, cv, v);
        s.add = 1;
        s.stop = (double)s.kap < thr;
        return s;
    case K_HBETA:
        if (v > P.alpha_lim) {
            s.stop = 1;
            return s;
        }
        s.kap = point_raw(P, ln, H, D, S, stehle, cv, v);
        s.add = 1;
        s.stop = (double)s.kap < thr;
        return s;
    case K_HMERGE: {
        if (k == 0) {
            if (v < P.wcon) {
                s.stop = 1;
                return s;
            }
            const float raw = point_raw(P, ln, H, D, S, stehle, cv, v);
            if ((double)raw < thr0) {
                s.stop = 1;
                return s;
            }
            s.kap = taper(P, v, raw);
            s.add = 1;
            return s;
        }
        if (v >= P.wcon) {
            s.kap = taper(P, v, point_raw(P, ln, H, D, S, stehle, cv, v));
            s.add = 1;
            s.stop = (double)s.kap < thr;
        }
        return s;
    }
    default:
        s.stop = 1;
        return s;
    }
}

// Blue walk index k (point ib0-1-k >= 1).  A metal or merged line's walk
// ends at the first point below wcon: the blue wavelengths fall
// monotonically, so the reference adds nothing after it either.
A12_HD inline Step blue_step(const Plan &P, const a12_cplx_line &ln, const HLine *H, const Depth &D,
                             const Stark &S, bool stehle, const HConv *cv, double v, double thr)
{
    Step s{0.0f, 0, 0};
    if ((P.kind == K_METAL || P.kind == K_HMERGE) && v < P.wcon) {
        s.stop = 1;
        return s;
    }
    s.kap = point_kappa(P, ln, H, D, S, stehle, cv, v);
    s.add = 1;
    s.stop = (double)s.kap < thr;
    return s;
}

// The points one (line, depth) adds to: red [rlo, rhi], blue [blo, bhi]
// (1-based, empty when lo > hi), or the edge's [nb1, min(nb3, length)].
struct Extent {
    int rlo, rhi, blo, bhi;
};

// The whole convolution of one (transition, depth) in one pass, the
// reference's sequence (:1422-1640); da/pr/step: NCONV_MAX scratch each.
A12_HD inline void conv_finish(const ConvSetup &C, HConv &cv)
{
    for (int ia = 0; ia < NDA; ++ia) cv.lprof[ia] = cv.prof[ia] > 0.0 ? log(cv.prof[ia]) : 0.0;
    cv.F0 = C.F0;
    cv.lambda0 = C.lambda0;
    cv.dop_dalpha = C.dop_dalpha;
    cv.voigt_a = C.voigt_a;
    cv.pad = 0;
}
A12_HD inline void conv_build(const HLine &H, const Depth &D, const Stark &S, float hwdop, HConv &cv,
                              double *da, double *pr, double *step)
{
    ConvSetup C;
    conv_params(H, D, S, hwdop, C);
    for (int ia = 0; ia < NDA; ++ia) conv_fill(C, ia);
    conv_sizes(C);
    for (int c = 0; c < C.ic; ++c) conv_grid_point(C, c, da[c], pr[c]);
    for (int c = 0; c < C.ic; ++c) step[c] = conv_step(da, c, C.ic);
    for (int ia = 0; ia < NDA; ++ia) cv.prof[ia] = conv_output(C, da, pr, step, ia);
    conv_finish(C, cv);
}

// The extents of one (line, depth) by walking both wings point by point
// (the reference's loops with the running wavelength products).
A12_HD inline Extent walk_extent(const Plan &P, const a12_cplx_line &ln, const HLine *H, const Depth &D,
                                 const Stark &S, bool stehle, const HConv *cv, const Grid &G,
                                 const float *cont)
{
    const int L = G.length;
    Extent e{1, 0, 1, 0};
    if (P.kind == K_SKIP) return e;
    if (P.kind == K_CONT) {
        e.rlo = P.nb1;
        e.rhi = P.nb3 < L ? P.nb3 : L;
        return e;
    }
    const int aut = P.kind == K_AUT;
    int stopped = 0;
    if (P.red) {
        double v = chain_start(G, aut, 1, P.minred, P.ib0);
        const int pc = P.minred < L ? P.minred : L;
        const double thr0 = (double)cont[pc - 1] * G.cutoff;
        for (int k = 0; P.minred + k <= L; ++k) {
            const int p = P.minred + k;
            const Step s = red_step(P, ln, H, D, S, stehle, cv, k, v, (double)cont[p - 1] * G.cutoff, thr0);
            if (s.add) {
                if (e.rlo > e.rhi) e.rlo = p;
                e.rhi = p;
            }
            if (s.stop) {
                stopped = 1;
                break;
            }
            v = chain_next(G, aut, 1, v);
        }
    }
    if (!(stopped ? P.blue_if_red_stop : P.blue_if_red_done)) return e;
    double v = chain_start(G, aut, 0, P.minred, P.ib0);
    for (int k = 0; P.ib0 - 1 - k >= 1; ++k) {
        const int p = P.ib0 - 1 - k;
        const Step s = blue_step(P, ln, H, D, S, stehle, cv, v, (double)cont[p - 1] * G.cutoff);
        if (s.add) {
            if (e.blo > e.bhi) e.bhi = p;
            e.blo = p;
        }
        if (s.stop) break;
        v = chain_next(G, aut, 0, v);
    }
    return e;
}

// Chain value at walk index k by running the chain from index 0.
A12_HD inline double chain_at(const Grid &G, int aut, int red, int minred, int ib0, int k)
{
    double v = chain_start(G, aut, red, minred, ib0);
    for (int i = 0; i < k; ++i) v = chain_next(G, aut, red, v);
    return v;
}

}  // namespace cplx
