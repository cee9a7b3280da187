// a12_kapp_core.h -- KAPP, the continuum opacity of one (frequency, depth)
// cell, as host+device functions over flat arrays.  Included by
// a12_kapp.hip, whose kernels run these for every cell of a frequency grid.
//
// The reference (src/atlas12_modules.f90:11202-17980) runs KAPP once per
// frequency with loops over the depths and caches per-temperature factors
// between calls (ITEMP/TB_T_LAST).  Here the work splits three ways:
//   kp_freq:  everything that depends on the frequency only (cross sections,
//             table rows), one row per frequency (KF_* slots);
//   kp_state: the per-depth factors of the state (the caches), one column
//             per depth (KS_* planes);
//   kp_cell:  the rest, per (frequency, depth), in the reference's order of
//             operations.
// Expressions keep the reference's evaluation order (left to right, its
// parentheses, REAL(4) where it declares REAL(4)); the translation unit
// compiles with FP contraction off.
#pragma once

#ifndef A12_HD
#define A12_HD __host__ __device__
#endif

namespace kapp {

#include "a12_kapp_data.h"

// ---- mod_constants (src/atlas12_modules.f90:305-464) -----------------------
constexpr double PI = 3.141592653589793;
constexpr double KBOL = 1.380649e-16, HPLANCK = 6.62607015e-27, CLIGHT = 2.99792458e10;
constexpr double CLIGHT_NMS = 2.99792458e17, CLIGHT_ANGS = 2.99792458e18;
constexpr double EMASS = 9.1093837139e-28, SIGMA_THOMSON = 6.6524587051e-25;
constexpr double EV_ERG = 1.602176634e-12;
constexpr double HCK = HPLANCK * CLIGHT / KBOL;
constexpr double HOVERK = HPLANCK / KBOL;
constexpr double KBOL_EV = KBOL / EV_ERG;
constexpr double LN10 = 2.30258509299405;
constexpr double THETA_COEFF = 0.4342944819032518 / KBOL_EV;
constexpr double BNU_PREFAC = 2.0 * HPLANCK / (CLIGHT * CLIGHT) * 1.0e45;
constexpr double COEFF_FF = 3.69196e8;
constexpr double RYDBERG_H = 109677.583403, RYDBERG_HE = 109722.277609;
constexpr double ELIM_HI = 109678.764;
constexpr double FREQ_RYDH = RYDBERG_H * CLIGHT;
constexpr double RYD_ANG = 1.0e8 / RYDBERG_H;
constexpr double HMINUS_EA = 0.75419;
constexpr double BETA_COEFF_HM88 = 8.798905203085208e+14;
constexpr double LOG_BETA_MIN = -2.0, LOG_BETA_STEP = 2.4825302042523617e-02;

// ---- level constants written inline in the sources -------------------------
// COULFF's log10 Z^2 for Z = 1..6 (:12008)
constexpr double COULFF_Z2LOG[6] = {0.0, 0.60206, 0.95424, 1.20412, 1.39794, 1.55630};
// HOP's explicit levels: statistical weight 2 n^2 and excitation energy
// (cm^-1) as written per level (:11406-11505)
constexpr double HOP_G[16] = {0, 2.0, 8.0, 18.0, 32.0, 50.0, 72.0, 98.0, 128.0, 162.0, 200.0, 242.0,
                              288.0, 338.0, 392.0, 450.0};
constexpr double HOP_E[16] = {0, 0, 82259.105, 97492.302, 102823.893, 105291.651, 106632.160, 107440.444,
                              107965.051, 108324.719, 108581.992, 108772.336, 108917.117, 109029.789,
                              109119.188, 109191.313};
// HE2OP's levels n = 9..1 (excitation energies, cm^-1) and the thresholds
// in WAVENO that switch each on (:13590-13648); E(9) is the limit term
constexpr double HE2_E[9] = {433490.46, 432051.19, 429951.90, 426717.413, 421353.135,
                             411477.925, 390142.359, 329182.321, 438908.85};
constexpr double HE2_LIM[9] = {5418.390, 6857.660, 8956.950, 12191.437, 17555.715, 27430.925, 48766.491,
                               109726.529, 438908.850};
// HE1OP's excited-core series: level energies below the He II 1s and 2s/2p
// limits, in the order the reference tests their edges (:13047-13081)
constexpr double HE1_EDGE1[4] = {171135.000, 169087.0, 166277.546, 159856.069};
constexpr double HE1_EDGE2[5] = {186209.471, 186101.0, 185564.0, 184864.0, 183236.0};

// ---- frequency row (kp_freq) ------------------------------------------------
enum {
    KF_XK = 0,          // XKARZAS(FREQ, 1, n, n), n = 1..80 at KF_XK + n - 1
    KF_HE1T = 80,       // HE1OP TRANS(1..10)
    KF_HE1N = 90,       // HE1OP TRANSN(4..27) at KF_HE1N + n - 4
    KF_HE1IMIN = 114,   // HE1OP IMIN
    KF_FFTT = 115,      // HMINOP FFTT(1..11)
    KF_HMBF = 126,      // HMINOP bound-free cross section (1e-18 cm^2)
    KF_H2MKT = 127,     // H2MINOP KT(1..7)
    KF_HEMKT = 134,     // HEMIOP KT(1..12)
    KF_HRAY = 146,      // HRAYOP XSECT
    KF_HERA = 147,      // HERAOP SIG
    KF_H2RA = 148,      // H2RAOP SIG
    KF_H2PFR = 149,     // H2PLOP FR
    KF_H2PES = 150,     // H2PLOP ES
    KF_CH = 151,        // CHOP CROSSCHT(1..15)
    KF_CHN = 166,       // CHOP N (energy bin)
    KF_OH = 167,        // OHOP CROSSOHT(1..15)
    KF_OHN = 182,       // OHOP N
    KF_FELO = 183,      // FELO_SIGMA(ion, i, WAVENO) at KF_FELO + 60 (ion - 1) + i - 1
    KF_HLN = 303,       // HLINOP lower level N (0: no line opacity at this frequency)
    KF_HLMF = 304,      // HLINOP MFREQ
    KF_HI = 305,        // MBF_HIGH_ION XSECT of edge e at KF_HI + e (0 below the edge), 60 edges
    KF_N = 365
};

// ---- per-depth planes (kp_state), each n values ----------------------------
enum {
    KS_WOCC = 0,                  // occupation_prob(n, XNE), n = 0..80 (81 planes)
    KS_HOPE = KS_WOCC + 81,       // HOP explicit levels: EXP(-E_n HCKT), n = 0..15 (16)
    KS_HIE = KS_HOPE + 16,        // EXP(-(ELIM_HI - RYDBERG_H/n^2) HCKT), n = 0..80 (81)
    KS_HE1B = KS_HIE + 81,        // HE1OP BOLT(J, 1..10) at +n-1 (10)
    KS_HE1BN = KS_HE1B + 10,      // HE1OP BOLTN(J, 4..27) at +n-4 (24)
    KS_HE1EXLIM = KS_HE1BN + 24,  // EXLIM
    KS_HE1BOLTEX,                 // BOLTEX
    KS_HE1FREET,                  // FREET
    KS_HMTHETA,                   // HMINOP THETA
    KS_HMXHMIN,                   // HMINOP XHMIN
    KS_XNH2M,                     // H2MINOP XNH2M
    KS_XNH2,                      // H2COLLOP/H2RAOP XNH2
    KS_CIADEN,                    // H2COLLOP CIADEN(J, 1..4) (4)
    KS_CIAIT = KS_CIADEN + 4,     // CIAIT (4, as doubles)
    KS_CIAWT = KS_CIAIT + 4,      // CIAWT (4)
    KS_HE2E = KS_CIAWT + 4,       // HE2OP EXP(-E HCKT) of its 8 levels and the limit (9)
    KS_HLBOLT = KS_HE2E + 9,      // HLINOP BOLT(J, 1..4) at +n-1 (4)
    KS_F0 = KS_HLBOLT + 4,        // STARK F0
    KS_FELOE,                     // FELO EXP(-FELO_E HCKT): 2 x 60 planes
    KS_HIEXP = KS_FELOE + 120,    // MBF_HIGH_ION exp(E_exc / TKEV) per edge (60)
    KS_TB = KS_HIEXP + 60,        // XNFP(J, species) * TB_BOLTZ(level, J), one plane per TOPbase level
};

// The tables of a12_kapp_tables on one side (host or device), plus the
// init-time logarithms the sources form on their first call.
struct KTab {
    const float *k_freqn, *k_xn, *k_xl, *k_ekarzas;
    const double *gff;                        // (146, 81)
    int cia_npair, cia_ntmax, cia_nnu;
    double cia_nu0, cia_dnu;
    const int *cia_nt, *cia_use;
    const double *cia_temp, *cia_tab;
    const double *cross_ch, *cross_oh;
    int hi_nedge;
    const double *hi_edge;                    // (nedge, 7)
    int felo_nlev[2], felo_idx[2];
    const double *felo_nuth, *felo_g, *felo_e, *felo_x, *felo_logsig, *felo_seaton;
    const int *felo_ifancy, *felo_npts;
    int tb_nsp, tb_nlev_tot;
    const int *tb_idx, *tb_nlev, *tb_g, *tb_np, *tb_off;   // tb_off: first point of each level
    const int *tb_sp;                                        // species of each level
    const double *tb_thr, *tb_ebind, *tb_eexc, *tb_logemin, *tb_logestride, *tb_ephot, *tb_sigma;
    int st_loaded[4], st_numin[4], st_numax[4], st_ndens[4];
    const double *st_dens[4], *st_temp[4], *st_ldalpha[4], *st_kalpha[4], *st_prof[4];
    const int *st_maxdens[4];
    // HMINOP/H2MINOP/HEMIOP first-call logarithms (host libm, as the reference)
    const double *hm_wfflog;                  // (22)
    const double *hm_fflog;                   // (22, 11): FFLOG(IWAVE, ITHETA) at [iw*11 + it]
    const double *h2m_logw, *h2m_logt, *h2m_logk;   // (10), (7), (7, 10): [it*10 + iw]
    const double *hem_logw, *hem_logt, *hem_logk;   // (17), (12), (12, 17)
    double saha_prefac, lg_ryh_over_k, lg_k_over_h;
};

// The depth state (a12_kapp_state) with XNF/XNFP as (mion, n), depth fastest.
struct KState {
    int n, mion;
    const double *t, *tkev, *tlog, *hkt, *hckt, *xne, *rho, *bmin, *bhyd, *eqh2, *xnf, *xnfp;
};

A12_HD inline double xnf(const KState &S, int idx, int j) { return S.xnf[(size_t)(idx - 1) * S.n + j]; }
A12_HD inline double xnfp(const KState &S, int idx, int j) { return S.xnfp[(size_t)(idx - 1) * S.n + j]; }
A12_HD inline double bhyd(const KState &S, int m, int j) { return S.bhyd[(size_t)(m - 1) * S.n + j]; }

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

// ---- XKARZAS (src/atlas12_modules.f90:11635-11728), REAL(4) interpolation
A12_HD inline double xkarzas(const KTab &T, double freq, double zeff2, int n, int l)
{
    const int NPTS = 29;
    const float freqlg = (float)log10(freq / zeff2);
    float x;
    if (l >= n || n > 6) {
        if (n > 15) {
            const float rn2 = (float)n * (float)n;
            float f_last = (float)log10(FREQ_RYDH / (double)rn2);
            if (freqlg < f_last) return 0.0;
            float fprev = 0.0f, fi = 0.0f;
            int i;
            for (i = 2; i <= NPTS - 1; ++i) {
                fprev = fi;
                fi = (float)log10((double)(T.k_ekarzas[i - 1] + 1.0f / rn2) * FREQ_RYDH);
                if (freqlg > fi) break;
            }
            float fa, fb;  // FREQN15(I), FREQN15(I-1)
            if (i > NPTS - 1) {
                i = NPTS;
                fa = f_last;
                fb = fi;
            } else {
                fa = fi;
                fb = fprev;
            }
            const float *xn = T.k_xn + 14 * NPTS;
            x = (freqlg - fa) / (fb - fa) * (xn[i - 2] - xn[i - 1]) + xn[i - 1];
        } else {
            const float *fr = T.k_freqn + (n - 1) * NPTS;
            if (freqlg < fr[NPTS - 1]) return 0.0;
            int i;
            for (i = 2; i <= NPTS; ++i)
                if (freqlg > fr[i - 1]) break;
            if (i > NPTS) i = NPTS;
            const float *xn = T.k_xn + (n - 1) * NPTS;
            x = (freqlg - fr[i - 1]) / (fr[i - 2] - fr[i - 1]) * (xn[i - 2] - xn[i - 1]) + xn[i - 1];
        }
    } else {
        const float *fr = T.k_freqn + (n - 1) * NPTS;
        if (freqlg < fr[NPTS - 1]) return 0.0;
        int i;
        for (i = 2; i <= NPTS; ++i)
            if (freqlg > fr[i - 1]) break;
        if (i > NPTS) i = NPTS;
        const float *xl = T.k_xl + ((n - 1) * 6 + l) * NPTS;
        x = (freqlg - fr[i - 1]) / (fr[i - 2] - fr[i - 1]) * (xl[i - 2] - xl[i - 1]) + xl[i - 1];
    }
    return exp((double)x * LN10) / zeff2;
}

// ---- holtsmark_Q, occupation_prob (:22492-22567) ----------------------------
A12_HD inline double occupation_prob(int n, double xne, const double *qh)
{
    if (n <= 1 || xne <= 0.0) return 1.0;
    const double dn = (double)n;
    const double beta = BETA_COEFF_HM88 / (dn * dn * dn * dn * dn * pow(xne, 2.0 / 3.0));
    if (beta <= 0.01) return 0.0;
    if (beta >= 50.0) return 1.0;
    const double idx_f = (log10(beta) - LOG_BETA_MIN) / LOG_BETA_STEP;
    int idx = (int)idx_f;
    idx = idx < 0 ? 0 : (idx > 148 ? 148 : idx);
    const double frac = idx_f - (double)idx;
    return qh[idx] + frac * (qh[idx + 1] - qh[idx]);
}

// ---- MAP1 at one point (:5545-5640) -----------------------------------------
A12_HD inline double map1_point(const double *xo, const double *fo, int nold, double x)
{
    const double *XO = xo - 1, *FO = fo - 1;
    int L = 2;
    bool past = false;
    while (x >= XO[L]) {
        L = L + 1;
        if (L > nold) { past = true; break; }
    }
    if (L <= 3 || past) {
        L = nold < L ? nold : L;
        const double B = (FO[L] - FO[L - 1]) / (XO[L] - XO[L - 1]);
        const double A = FO[L] - XO[L] * B;
        return A + B * x;
    }
    const int L1 = L - 1, L2 = L - 2;
    double D = (FO[L1] - FO[L2]) / (XO[L1] - XO[L2]);
    const double CBAC = FO[L] / ((XO[L] - XO[L1]) * (XO[L] - XO[L2]))
                      + (FO[L2] / (XO[L] - XO[L2]) - FO[L1] / (XO[L] - XO[L1])) / (XO[L1] - XO[L2]);
    const double BBAC = D - (XO[L1] + XO[L2]) * CBAC;
    const double ABAC = FO[L2] - XO[L2] * D + XO[L1] * XO[L2] * CBAC;
    if (L >= nold) return ABAC + (BBAC + CBAC * x) * x;
    D = (FO[L] - FO[L1]) / (XO[L] - XO[L1]);
    const double CFOR = FO[L + 1] / ((XO[L + 1] - XO[L]) * (XO[L + 1] - XO[L1]))
                      + (FO[L1] / (XO[L + 1] - XO[L1]) - FO[L] / (XO[L + 1] - XO[L])) / (XO[L] - XO[L1]);
    const double BFOR = D - (XO[L] + XO[L1]) * CFOR;
    const double AFOR = FO[L1] - XO[L1] * D + XO[L] * XO[L1] * CFOR;
    double WT = 0.0;
    if (fabs(CFOR) != 0.0) WT = fabs(CFOR) / (fabs(CFOR) + fabs(CBAC));
    const double A = AFOR + WT * (ABAC - AFOR);
    const double B = BFOR + WT * (BBAC - BFOR);
    const double C = CFOR + WT * (CBAC - CFOR);
    return A + (B + C * x) * x;
}

// ---- LINTER (:5859-5880) ------------------------------------------------------
A12_HD inline double linter(const double *xo, const double *yo, int nold, double x)
{
    int i;
    for (i = 2; i <= nold; ++i)
        if (x < xo[i - 1]) break;
    if (i > nold) i = nold;
    return yo[i - 2] + (yo[i - 1] - yo[i - 2]) / (xo[i - 1] - xo[i - 2]) * (x - xo[i - 2]);
}

// ---- COULFF (:11995-12196): van Hoof's thermally averaged g_ff ---------------
A12_HD inline double coulff(const KTab &T, double tlog, double freqlg, int nz)
{
    const double lgam2 = COULFF_Z2LOG[nz - 1] + T.lg_ryh_over_k - tlog / LN10;
    const double lu = (freqlg - tlog) / LN10 - T.lg_k_over_h;
    const double xi = (lgam2 - (-6.0)) / 0.2 + 1.0;
    const double xj = (lu - (-16.0)) / 0.2 + 1.0;
    int i0 = (int)xi, j0 = (int)xj;
    i0 = i0 < 1 ? 1 : (i0 > 80 ? 80 : i0);
    j0 = j0 < 1 ? 1 : (j0 > 145 ? 145 : j0);
    double fi = xi - (double)i0, fj = xj - (double)j0;
    fi = fi < 0.0 ? 0.0 : (fi > 1.0 ? 1.0 : fi);
    fj = fj < 0.0 ? 0.0 : (fj > 1.0 ? 1.0 : fj);
    const double *g = T.gff;
    const double lg00 = g[(j0 - 1) * 81 + (i0 - 1)], lg01 = g[j0 * 81 + (i0 - 1)];
    const double lg10 = g[(j0 - 1) * 81 + i0], lg11 = g[j0 * 81 + i0];
    const double lg = (1.0 - fi) * ((1.0 - fj) * lg00 + fj * lg01) + fi * ((1.0 - fj) * lg10 + fj * lg11);
    return pow(10.0, lg);
}

// ---- HE1OP's cross sections (:13160-13527) ----------------------------------
A12_HD inline double crosshe(double freq)
{
    if (freq < 5.945209e15) return 0.0;
    const double wave = CLIGHT_ANGS / freq;
    int i;
    if (wave > 50.0) {
        i = (int)(93.0 - (wave - 50.0) / 5.0);
        i = i < 2 ? 2 : (i > 92 ? 92 : i);
        return ((wave - (92 - i) * 5.0 - 50.0) / 5.0 * (HE_X505[i - 2] - HE_X505[i - 1]) + HE_X505[i - 1]) * 1.e-18;
    } else if (wave > 20.0) {
        i = (int)(17.0 - (wave - 20.0) / 2.0);
        i = i < 2 ? 2 : (i > 16 ? 16 : i);
        return ((wave - (16 - i) * 2.0 - 20.0) / 2.0 * (HE_X50[i - 2] - HE_X50[i - 1]) + HE_X50[i - 1]) * 1.e-18;
    } else if (wave > 10.0) {
        i = (int)(12.0 - (wave - 10.0) / 1.0);
        i = i < 2 ? 2 : (i > 11 ? 11 : i);
        return ((wave - (11 - i) * 1.0 - 10.0) / 1.0 * (HE_X20[i - 2] - HE_X20[i - 1]) + HE_X20[i - 1]) * 1.e-18;
    }
    i = (int)(22.0 - wave / 0.5);
    i = i < 2 ? 2 : (i > 21 ? 21 : i);
    return ((wave - (21 - i) * 0.5) / 0.5 * (HE_X10[i - 2] - HE_X10[i - 1]) + HE_X10[i - 1]) * 1.e-18;
}

// tabulated log10 cross section against log10 frequency, as HE12S1S/HE12S3S/
// HE12P1P/HE12P3P interpolate it below their resonance regime
A12_HD inline double he_tab(const double *fq, const double *xs, double freq)
{
    const double freqlg = log10(freq);
    int i;
    for (i = 2; i <= 16; ++i)
        if (freqlg > fq[i - 1]) break;
    if (i > 16) i = 16;
    const double x = (freqlg - fq[i - 1]) / (fq[i - 2] - fq[i - 1]) * (xs[i - 2] - xs[i - 1]) + xs[i - 1];
    return exp(x * LN10);
}

A12_HD inline double he12s1s(double freq)
{
    if (freq < 32033.214 * CLIGHT) return 0.0;
    if (freq > 2.4 * RYDBERG_HE * CLIGHT) {
        const double wavno = freq / CLIGHT;
        const double ek = (wavno - 32033.214) / RYDBERG_HE;
        const double eps = 2.0 * (ek - 2.612316) / 0.00322;
        return 0.008175 * pow(484940.0 / wavno, 2.71) * 8.067e-18 * ((eps + 76.21) * (eps + 76.21)) / (1.0 + eps * eps);
    }
    return he_tab(HE_FREQ1S, HE_X1S, freq);
}

A12_HD inline double he12s3s(double freq)
{
    if (freq < 38454.691 * CLIGHT) return 0.0;
    if (freq > 2.4 * RYDBERG_HE * CLIGHT) {
        const double wavno = freq / CLIGHT;
        const double ek = (wavno - 38454.691) / RYDBERG_HE;
        const double eps = 2.0 * (ek - 2.47898) / 0.000780;
        return 0.01521 * pow(470310.0 / wavno, 3.12) * 8.067e-18 * ((eps - 122.4) * (eps - 122.4)) / (1.0 + eps * eps);
    }
    return he_tab(HE_FREQ3S, HE_X3S, freq);
}

A12_HD inline double he12p1p(double freq)
{
    if (freq < 27175.76 * CLIGHT) return 0.0;
    if (freq > 2.4 * RYDBERG_HE * CLIGHT) {
        const double wavno = freq / CLIGHT;
        const double ek = (wavno - 27175.76) / RYDBERG_HE;
        const double e1s = 2.0 * (ek - 2.446534) / 0.01037;
        const double e1d = 2.0 * (ek - 2.59427) / 0.00538;
        return 0.0009487 * pow(466750.0 / wavno, 3.69) * 8.067e-18 *
               ((e1s - 29.30) * (e1s - 29.30) / (1.0 + e1s * e1s) + (e1d + 172.4) * (e1d + 172.4) / (1.0 + e1d * e1d));
    }
    return he_tab(HE_FREQ1P, HE_X1P, freq);
}

A12_HD inline double he12p3p(double freq)
{
    if (freq < 29223.753 * CLIGHT) return 0.0;
    return he_tab(HE_FREQ3P, HE_X3P, freq);
}

// ---- FELO_SIGMA (:14627-14683) -------------------------------------------------
A12_HD inline double felo_sigma(const KTab &T, int ion, int i, double wave_cm)
{
    const int k0 = ion * 60 + i;
    const double nu = wave_cm * 2.99792458e10;
    const double nu_th = T.felo_nuth[k0];
    if (T.felo_ifancy[k0] == 2) {
        if (nu < nu_th) return 0.0;
        const double *s = T.felo_seaton + 4 * k0;
        const double ratio = nu_th / nu;
        double sig = s[0] * (s[2] + (1.0 - s[2]) * pow(ratio, s[3])) * pow(ratio, s[1]);
        return sig * 1.0e-18;
    }
    if (nu <= 0.0) return 0.0;
    const double x = log10(nu / nu_th);
    const int nfit = T.felo_npts[k0];
    if (nfit < 2) return 0.0;
    const double *fx = T.felo_x + 200 * k0, *ls = T.felo_logsig + 200 * k0;
    const double xmin = fx[0], xmax = fx[nfit - 1];
    if (x < (xmin > 0.0 ? xmin : 0.0)) return 0.0;
    if (x > xmax) return 0.0;
    int k;
    for (k = 1; k <= nfit - 1; ++k)
        if (x <= fx[k]) break;
    const double frac = (x - fx[k - 1]) / (fx[k] - fx[k - 1]);
    const double logs = ls[k - 1] + frac * (ls[k] - ls[k - 1]);
    return pow(10.0, logs) * 1.0e-18;
}

// ---- hydrogen_f_value (:17760-17829) ---------------------------------------
A12_HD inline double hydrogen_f_value(int n, int m)
{
    const double xn = (double)n, xm = (double)m;
    if (m <= n) return 0.0;
    const int idx = m - n;
    double gf;
    if (n == 1 && idx <= 29) gf = GF_LY[idx - 1];
    else if (n == 2 && idx <= 28) gf = GF_BA[idx - 1];
    else if (n == 3 && idx <= 27) gf = GF_PA[idx - 1];
    else if (n == 4 && idx <= 3) gf = GF_BR[idx - 1];
    else {
        const double q = 1.0 / (xn * xn) - 1.0 / (xm * xm);
        gf = 1.9603 * (xn * xn) / ((xm * xm * xm) * ((q * q) * q));
        gf = gf * 0.80;
    }
    return gf / (2.0 * (xn * xn));
}

// ---- STARK (:17528-17640): VCS-like fallback profile ---------------------------
A12_HD inline double exint(double x)
{
    return -log(x) - 0.57516 +
           (0.97996 + (-0.21654 + (0.033572 + (-0.0029222 + 1.05439e-4 * x) * x) * x) * x) * x;
}

A12_HD inline double stark(int n, int m, double freq, double f0, double hkt, double xne)
{
    const double A0 = 0.0265384;
    const double xn = (double)n, xm = (double)m;
    const double x = xn / xm, xx = x * x;
    const double nn = xn * xn, mm = xm * xm;
    const int mminn = m - n;
    double knm, fnm;
    if (mminn <= 5) knm = ST_KNM[(n - 1) * 5 + mminn - 1];
    else knm = 5.5e-5 * ((nn * mm) * (nn * mm)) / (mm - nn);
    if (mminn <= 10) fnm = ST_F[(n - 1) * 10 + mminn - 1];
    else {
        const double q = (20.0 * xn + 100.0) / ((xn + 10.0) * xm * (1.0 - xx));
        fnm = ST_F[(n - 1) * 10 + 9] * ((q * q) * q);
    }
    const double freqnm = FREQ_RYDH * (1.0 / nn - 1.0 / mm);
    const double del = fabs(freq - freqnm);
    const double dbeta = CLIGHT_ANGS / (freqnm * freqnm) / f0 / knm;
    const double beta = dbeta * del;
    const double y1 = mm * del * hkt / 2.0;
    const double y2 = (PI * PI / 2.0 / A0 / CLIGHT) * (del * del) / xne;
    const double qstat = 1.5 + 0.5 * (y1 * y1 - 1.384) / (y1 * y1 + 1.384);
    double impact = 0.0;
    if (y1 < 8.0 && y1 < y2) {
        double exy2 = 0.0;
        if (y2 <= 8.0) exy2 = exint(y2);
        impact = 1.438 * sqrt(y1 * (1.0 - xx)) * (0.4 * exp(-y1) + exint(y1) - 0.5 * exy2);
    }
    double prof, ratio;
    if (beta <= 20.0) {
        prof = 8.0 / (80.0 + (beta * beta) * beta);
        ratio = qstat + impact;
    } else {
        prof = 1.5 / beta / beta / sqrt(beta);
        const double dioi = 6.28 * 1.48e-25 * (2.0 * mm * FREQ_RYDH / del) * xne *
                            (sqrt(2.0 * mm * FREQ_RYDH / del) * (1.3 * qstat + 0.30 * impact) - 3.9 * FREQ_RYDH * hkt);
        const double c = 1.0 + dioi;
        ratio = qstat * (c < 1.25 ? c : 1.25) + impact;
    }
    return A0 * fnm * prof * dbeta * ratio;
}

// ---- STARK_MMM (:17854-18106): Stehle tables, STARK outside them -------------
A12_HD inline double stark_mmm(const KTab &T, int n, int m, double freq, double tj, double xne,
                               double f0_stark, double hkt)
{
    const double A0 = 0.0265384;
    if (n < 1 || n > 4 || m <= n) return stark(n, m, freq, f0_stark, hkt, xne);
    const int s = n - 1;
    if (!T.st_loaded[s]) return stark(n, m, freq, f0_stark, hkt, xne);
    if (m < T.st_numin[s] || m > T.st_numax[s]) return stark(n, m, freq, f0_stark, hkt, xne);
    const int itrans = m - T.st_numin[s] + 1;
    const int nd = T.st_ndens[s], nt = 10;
    const double *dens = T.st_dens[s], *temp = T.st_temp[s], *lda = T.st_ldalpha[s];
    if (xne > dens[nd - 1]) return stark(n, m, freq, f0_stark, hkt, xne);
    const double dmax = dens[T.st_maxdens[s][itrans - 1] - 1];
    const double ne_eff = xne < dmax ? xne : dmax;
    const double xn = (double)n, xm = (double)m;
    const double lambda0 = RYD_ANG * ((xn * xm) * (xn * xm)) / ((xm - xn) * (xm + xn));
    const double freq_nm = CLIGHT_ANGS / lambda0;
    const double del_freq = fabs(freq - freq_nm);
    const double f0 = 1.25e-9 * pow(ne_eff, 2.0 / 3.0);
    if (f0 <= 0.0) return stark(n, m, freq, f0_stark, hkt, xne);
    const double dalpha = (lambda0 * lambda0) / CLIGHT_ANGS * del_freq / f0;
    const double log_ne = log10(ne_eff);
    if (tj < temp[0] * 0.5 || tj > temp[nt - 1] * 2.0) return stark(n, m, freq, f0_stark, hkt, xne);
    int id1;
    for (id1 = 1; id1 <= nd - 1; ++id1)
        if (ne_eff <= dens[id1]) break;
    id1 = id1 < 1 ? 1 : (id1 > nd - 1 ? nd - 1 : id1);
    const int id2 = id1 + 1;
    double frac_d = (log_ne - log10(dens[id1 - 1])) / (log10(dens[id2 - 1]) - log10(dens[id1 - 1]));
    frac_d = frac_d < 0.0 ? 0.0 : (frac_d > 1.0 ? 1.0 : frac_d);
    int it1;
    for (it1 = 1; it1 <= nt - 1; ++it1)
        if (tj <= temp[it1]) break;
    it1 = it1 < 1 ? 1 : (it1 > nt - 1 ? nt - 1 : it1);
    const int it2 = it1 + 1;
    double frac_t = (log(tj) - log(temp[it1 - 1])) / (log(temp[it2 - 1]) - log(temp[it1 - 1]));
    frac_t = frac_t < 0.0 ? 0.0 : (frac_t > 1.0 ? 1.0 : frac_t);
    int ia1, ia2;
    double frac_a;
    if (dalpha <= 0.0) {
        ia1 = ia2 = 1;
        frac_a = 0.0;
    } else {
        const double log_dalpha = log10(dalpha);
        if (log_dalpha <= lda[0]) {
            ia1 = ia2 = 1;
            frac_a = 0.0;
        } else if (log_dalpha >= lda[59]) {
            const double f_nm = hydrogen_f_value(n, m);
            const double i_dalpha = T.st_kalpha[s][itrans - 1] / pow(dalpha, 2.5);
            return A0 * f_nm * i_dalpha * (lambda0 * lambda0) / (CLIGHT_ANGS * f0);
        } else {
            for (ia1 = 1; ia1 <= 59; ++ia1)
                if (log_dalpha <= lda[ia1]) break;
            ia1 = ia1 < 1 ? 1 : (ia1 > 59 ? 59 : ia1);
            ia2 = ia1 + 1;
            frac_a = (log_dalpha - lda[ia1 - 1]) / (lda[ia2 - 1] - lda[ia1 - 1]);
            frac_a = frac_a < 0.0 ? 0.0 : (frac_a > 1.0 ? 1.0 : frac_a);
        }
    }
    const double *p = T.st_prof[s] + (size_t)(itrans - 1) * nd * 600;
#define PV(ia, it, id) p[((size_t)((id) - 1) * 10 + ((it) - 1)) * 60 + ((ia) - 1)]
    double v000 = PV(ia1, it1, id1), v001 = PV(ia2, it1, id1);
    double v010 = PV(ia1, it2, id1), v011 = PV(ia2, it2, id1);
    double v100 = PV(ia1, it1, id2), v101 = PV(ia2, it1, id2);
    double v110 = PV(ia1, it2, id2), v111 = PV(ia2, it2, id2);
#undef PV
    double i_dalpha;
    if (v000 <= 0.0 || v001 <= 0.0 || v010 <= 0.0 || v011 <= 0.0 || v100 <= 0.0 || v101 <= 0.0 ||
        v110 <= 0.0 || v111 <= 0.0) {
        const double v00 = v000 + frac_a * (v001 - v000), v01 = v010 + frac_a * (v011 - v010);
        const double v10 = v100 + frac_a * (v101 - v100), v11 = v110 + frac_a * (v111 - v110);
        const double v0 = v00 + frac_t * (v01 - v00), v1 = v10 + frac_t * (v11 - v10);
        i_dalpha = v0 + frac_d * (v1 - v0);
    } else {
        v000 = log(v000); v001 = log(v001); v010 = log(v010); v011 = log(v011);
        v100 = log(v100); v101 = log(v101); v110 = log(v110); v111 = log(v111);
        const double v00 = v000 + frac_a * (v001 - v000), v01 = v010 + frac_a * (v011 - v010);
        const double v10 = v100 + frac_a * (v101 - v100), v11 = v110 + frac_a * (v111 - v110);
        const double v0 = v00 + frac_t * (v01 - v00), v1 = v10 + frac_t * (v11 - v10);
        i_dalpha = exp(v0 + frac_d * (v1 - v0));
    }
    i_dalpha = i_dalpha > 0.0 ? i_dalpha : 0.0;
    const double f_nm = hydrogen_f_value(n, m);
    return A0 * f_nm * i_dalpha * (lambda0 * lambda0) / (CLIGHT_ANGS * f0);
}

// ---- the frequency row -----------------------------------------------------------
A12_HD inline void kp_freq(const KTab &T, double wave, double *row)
{
    for (int k = 0; k < KF_N; ++k) row[k] = 0.0;
    const double freq = CLIGHT_NMS / wave;
    const double waveno = 1.0e7 / wave;
    const double freqlg = log(freq);
    // HOP (:11359): the Karzas & Latter cross sections of levels 1..80
    for (int n = 1; n <= 80; ++n) row[KF_XK + n - 1] = xkarzas(T, freq, 1.0, n, n);
    // HE1OP (:12939-13087)
    {
        int imin = 0;
        for (int n = 1; n <= 10; ++n)
            if (HE1_FREQ[n - 1] <= freq) { imin = n; break; }
        double tr[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (imin > 0) {
            if (imin <= 1) tr[0] = crosshe(freq);
            if (imin <= 2) tr[1] = he12s3s(freq);
            if (imin <= 3) tr[2] = he12s1s(freq);
            if (imin <= 4) tr[3] = he12p3p(freq);
            if (imin <= 5) tr[4] = he12p1p(freq);
            if (imin <= 6) tr[5] = xkarzas(T, freq, 1.236439, 3, 0);
            if (imin <= 7) tr[6] = xkarzas(T, freq, 1.102898, 3, 0);
            if (imin <= 8) tr[7] = xkarzas(T, freq, 1.045499, 3, 1);
            if (imin <= 9) tr[8] = xkarzas(T, freq, 1.001427, 3, 2);
            if (imin <= 10) tr[9] = xkarzas(T, freq, 0.9926, 3, 1);
            const double ryd = RYDBERG_HE * CLIGHT;
            // the series converging on He II 1s (ELIM 527490.06) and 2s/2p
            // (ELIM 588451.59): each level from its own edge frequency on
            for (int q = 0; q < 4; ++q) {
                const double fhe = (527490.06 - HE1_EDGE1[q]) * CLIGHT;
                if (!(freq >= fhe)) break;
                tr[4 - q] = tr[4 - q] + xkarzas(T, freq, fhe / ryd, 1, 0);
            }
            for (int q = 0; q < 5; ++q) {
                const double fhe = (588451.59 - HE1_EDGE2[q]) * CLIGHT;
                if (!(freq >= fhe)) break;
                tr[9 - q] = tr[9 - q] + xkarzas(T, freq, fhe / ryd, 1, 0);
            }
        }
        for (int n = 0; n < 10; ++n) row[KF_HE1T + n] = tr[n];
        row[KF_HE1IMIN] = (double)imin;
        if (freq >= 1.25408e16)
            for (int n = 4; n <= 27; ++n) {
                const double dn = (double)n;
                row[KF_HE1N + n - 4] = xkarzas(T, freq, 4.0 - 3.0 / (dn * dn), 1, 0);
            }
    }
    // HMINOP (:12514-12533): FFTT per THETAFF point, bound-free by MAP1
    {
        const double wavelog = log(wave);
        for (int it = 0; it < 11; ++it) {
            double ylog[22];
            for (int iw = 0; iw < 22; ++iw) ylog[iw] = T.hm_fflog[iw * 11 + it];
            const double fftlog = linter(T.hm_wfflog, ylog, 22, wavelog);
            row[KF_FFTT + it] = exp(fftlog) / HM_THETAFF[it] * THETA_COEFF * KBOL;
        }
        row[KF_HMBF] = freq > 1.82365e14 ? map1_point(HM_WBF, HM_BF, 116, wave) : 0.0;
    }
    // H2MINOP (:12677-12687) and HEMIOP (:13808-13818): log k at this wavelength
    {
        double wl = wave;
        if (wl > H2M_W[9]) wl = H2M_W[9];
        int iw = 1;
        for (int i = 1; i <= 9; ++i)
            if (wl >= H2M_W[i - 1]) iw = i;
        const double xw = (log(wl) - T.h2m_logw[iw - 1]) / (T.h2m_logw[iw] - T.h2m_logw[iw - 1]);
        for (int j = 0; j < 7; ++j)
            row[KF_H2MKT + j] = T.h2m_logk[j * 10 + iw - 1] * (1.0 - xw) + T.h2m_logk[j * 10 + iw] * xw;
        wl = wave;
        if (wl > HEM_W[16]) wl = HEM_W[16];
        iw = 1;
        for (int i = 1; i <= 16; ++i)
            if (wl >= HEM_W[i - 1]) iw = i;
        const double xh = (log(wl) - T.hem_logw[iw - 1]) / (T.hem_logw[iw] - T.hem_logw[iw - 1]);
        for (int j = 0; j < 12; ++j)
            row[KF_HEMKT + j] = T.hem_logk[j * 17 + iw - 1] * (1.0 - xh) + T.hem_logk[j * 17 + iw] * xh;
    }
    // HRAYOP (:12826-12895)
    {
        const double D1 = 0.01 * FREQ_RYDH, DAB = 0.005 * FREQ_RYDH, DBC = 0.002 * FREQ_RYDH,
                     DCD = 0.001 * FREQ_RYDH;
        double xs = 0.0, g;
        int i;
        if (freq < D1) {
            const double r = freq / D1;
            xs = SIGMA_THOMSON * (GAV_M[0] * GAV_M[0]) * (((r * r) * r) * r);
        } else if (freq <= 0.74 * FREQ_RYDH) {
            i = (int)(freq / D1);
            i = i + 1 < 74 ? i + 1 : 74;
            g = GAV_M[i - 2] + (GAV_M[i - 1] - GAV_M[i - 2]) / D1 * (freq - (double)(i - 1) * D1);
            xs = SIGMA_THOMSON * (g * g);
        } else if (freq < 0.77 * FREQ_RYDH) {
            xs = 0.0;
        } else if (freq <= 0.885 * FREQ_RYDH) {
            i = (int)((freq - 0.755 * FREQ_RYDH) / DAB) + 1;
            i = i + 1 < 27 ? i + 1 : 27;
            g = GAV_AB[i - 2] + (GAV_AB[i - 1] - GAV_AB[i - 2]) / DAB *
                                    (freq - (0.755 * FREQ_RYDH + (double)(i - 1 - 1) * DAB));
            xs = SIGMA_THOMSON * (g * g);
        } else if (freq < 0.890 * FREQ_RYDH) {
            xs = 0.0;
        } else if (freq <= 0.936 * FREQ_RYDH) {
            i = (int)((freq - 0.890 * FREQ_RYDH) / DBC) + 1;
            i = i + 1 < 24 ? i + 1 : 24;
            g = GAV_BC[i - 2] + (GAV_BC[i - 1] - GAV_BC[i - 2]) / DBC *
                                    (freq - (0.890 * FREQ_RYDH + (double)(i - 1 - 1) * DBC));
            xs = SIGMA_THOMSON * (g * g);
        } else if (freq < 0.938 * FREQ_RYDH) {
            xs = 0.0;
        } else if (freq <= 0.959 * FREQ_RYDH) {
            i = (int)((freq - 0.938 * FREQ_RYDH) / DCD) + 1;
            i = i + 1 < 22 ? i + 1 : 22;
            g = GAV_CD[i - 2] + (GAV_CD[i - 1] - GAV_CD[i - 2]) / DCD *
                                    (freq - (0.938 * FREQ_RYDH + (double)(i - 1 - 1) * DCD));
            xs = SIGMA_THOMSON * (g * g);
        } else if (freq < 0.961 * FREQ_RYDH) {
            xs = 0.0;
        } else if (freq <= FREQ_RYDH) {
            xs = SIGMA_THOMSON * GAV_LYC[0];
        } else {
            xs = SIGMA_THOMSON * map1_point(GAV_FLYC, GAV_LYC, 64, freq / FREQ_RYDH);
        }
        row[KF_HRAY] = xs;
    }
    // HERAOP (:13918-13941), H2RAOP (:17289-17338)
    {
        const double w = CLIGHT_ANGS / (freq < 5.15e15 ? freq : 5.15e15);
        const double ww = w * w;
        const double q = 1.0 + (2.44e5 + 5.94e10 / (ww - 2.90e5)) / ww;
        row[KF_HERA] = 5.484e-14 / ww / ww * (q * q);
        const double w2 = CLIGHT_ANGS / (freq < 2.922e15 ? freq : 2.922e15);
        const double ww2 = w2 * w2;
        row[KF_H2RA] = (8.14e-13 + 1.28e-6 / ww2 + 1.61 / (ww2 * ww2)) / (ww2 * ww2);
    }
    // H2PLOP (:12224-12252)
    {
        const double fl = freqlg;
        row[KF_H2PFR] = -3.0233e3 + (3.7797e2 + (-1.82496e1 + (3.9207e-1 - 3.1672e-3 * fl) * fl) * fl) * fl;
        row[KF_H2PES] = -7.342e-3 + (-2.409e-15 + (1.028e-30 + (-4.230e-46 + (1.224e-61 - 1.351e-77 * freq) *
                                                                    freq) * freq) * freq) * freq;
    }
    // CHOP (:16156-16167) and OHOP (:16281-16292): the cross section row at
    // this energy, only where the energy bin lies in the table
    {
        const double evolt = waveno / 8065.479;
        int nb = (int)(evolt * 10.0);
        row[KF_CHN] = (double)nb;
        if (!(nb < 20 || nb >= 105)) {
            const double en = (double)nb * 0.1;
            for (int it = 0; it < 15; ++it) {
                const double a = T.cross_ch[(nb - 1) * 15 + it], b = T.cross_ch[nb * 15 + it];
                row[KF_CH + it] = a + (b - a) * (evolt - en) / 0.1;
            }
        }
        nb = (int)(evolt * 10.0 - 20.0);
        row[KF_OHN] = (double)nb;
        if (!(nb <= 0 || nb >= 130)) {
            const double en = (double)nb * 0.1 + 2.0;
            for (int it = 0; it < 15; ++it) {
                const double a = T.cross_oh[(nb - 1) * 15 + it], b = T.cross_oh[nb * 15 + it];
                row[KF_OH + it] = a + (b - a) * (evolt - en) / 0.1;
            }
        }
    }
    // FELO_SIGMA of every Fe I / Fe II level (:14725)
    for (int ion = 0; ion < 2; ++ion)
        for (int i = 0; i < T.felo_nlev[ion]; ++i) row[KF_FELO + 60 * ion + i] = felo_sigma(T, ion, i, waveno);
    // MBF_HIGH_ION's edge cross sections (:14374-14379)
    for (int e = 0; e < T.hi_nedge; ++e) {
        const double *h = T.hi_edge + 7 * e;
        if (freq < h[0]) continue;
        const double fr = h[0] / freq;
        row[KF_HI + e] = h[1] * (h[2] + fr - h[2] * fr) * sqrt(powi_rt(fr, (int)h[3]));
    }
    // HLINOP (:17420-17439): the lower level and the nearest upper level
    {
        const double NU_LYMAN = FREQ_RYDH;
        const int n = (int)sqrt(NU_LYMAN / freq);
        int ok = !(n == 0 || n > 4);
        if (ok && n == 1 && freq < 2.0e15) ok = 0;
        if (ok && n == 2 && freq < 4.44e14) ok = 0;
        if (ok) {
            const double dn = (double)n;
            row[KF_HLN] = (double)n;
            row[KF_HLMF] = (double)(int)round(sqrt(NU_LYMAN / (NU_LYMAN / (dn * dn) - freq)));
        }
    }
}

// ---- the per-depth factors of a state -----------------------------------------
// One depth j; planes P[k * n + j].  TOPbase levels go through kp_state_tb.
A12_HD inline void kp_state(const KTab &T, const KState &S, const double *qh, int j, double *P)
{
    const int n = S.n;
#define PL(k) P[(size_t)(k) * n + j]
    const double t = S.t[j], tkev = S.tkev[j], hckt = S.hckt[j], xne = S.xne[j], rho = S.rho[j];
    for (int m = 0; m <= 80; ++m) PL(KS_WOCC + m) = occupation_prob(m, xne, qh);
    // HOP's explicit levels 2..15 (:11406-11505)
    for (int m = 0; m < 16; ++m) PL(KS_HOPE + m) = m >= 2 ? exp(-HOP_E[m] * hckt) : 0.0;
    for (int m = 0; m <= 80; ++m) {
        if (m < 7) { PL(KS_HIE + m) = 0.0; continue; }
        const double dm = (double)m;
        const double e_n = ELIM_HI - RYDBERG_H / (dm * dm);
        PL(KS_HIE + m) = exp(-e_n * hckt);
    }
    // HE1OP's cache (:12991-13006)
    const double x3 = xnfp(S, 3, j);
    for (int m = 1; m <= 10; ++m) PL(KS_HE1B + m - 1) = exp(-HE1_CHI[m - 1] / tkev) * HE1_G[m - 1] * x3 / rho;
    for (int m = 4; m <= 27; ++m) {
        const double dm = (double)m;
        PL(KS_HE1BN + m - 4) = exp(-24.587 * (1.0 - 1.0 / (dm * dm)) / tkev) * 4.0 * (dm * dm) * x3 / rho;
    }
    PL(KS_HE1FREET) = xne * xnf(S, 4, j) / rho / sqrt(t);
    const double xr = x3 * (4.0 / 2.0 / 13.595) * tkev / rho;
    PL(KS_HE1BOLTEX) = exp(-23.730 / tkev) * xr;
    PL(KS_HE1EXLIM) = exp(-24.587 / tkev) * xr;
    // HMINOP's cache (:12506-12512)
    const double b1 = bhyd(S, 1, j), x1 = xnfp(S, 1, j);
    PL(KS_HMTHETA) = THETA_COEFF / t;
    PL(KS_HMXHMIN) = exp(HMINUS_EA / tkev) / (T.saha_prefac * t * sqrt(t)) * S.bmin[j] * b1 * x1 * xne;
    // H2MINOP (:12670-12675), H2COLLOP (:16482-16503)
    const double xnh2m_b = x1 * 2.0 * b1;
    PL(KS_XNH2M) = (xnh2m_b * xnh2m_b) * S.eqh2[j];
    const double xnhi = x1 * 2.0 * b1;
    const double xnh2 = t <= 20000.0 ? (xnhi * xnhi) * S.eqh2[j] : 0.0;
    PL(KS_XNH2) = xnh2;
    const double x3f = xnf(S, 3, j);
    const double den[4] = {xnh2 * xnh2, xnh2 * x3f, xnh2 * xnhi, xnhi * x3f};
    for (int ip = 0; ip < 4; ++ip) {
        PL(KS_CIADEN + ip) = den[ip];
        int it = 1;
        double wt = 0.0;
        if (ip < T.cia_npair) {
            const int nt = T.cia_nt[ip];
            const double *tt = T.cia_temp + (size_t)ip * T.cia_ntmax;
            const double tmin = tt[0], tmax = tt[nt - 1];
            const double tc = tmin > (tmax < t ? tmax : t) ? tmin : (tmax < t ? tmax : t);
            for (int i = 1; i <= nt - 1; ++i)
                if (tc >= tt[i - 1]) it = i;
            wt = (tc - tt[it - 1]) / (tt[it] - tt[it - 1]);
        }
        PL(KS_CIAIT + ip) = (double)it;
        PL(KS_CIAWT + ip) = wt;
    }
    // HE2OP's level factors (:13590-13648)
    for (int m = 0; m < 9; ++m) PL(KS_HE2E + m) = exp(-HE2_E[m] * hckt);
    // HLINOP's cache (:17410-17417)
    for (int m = 1; m <= 4; ++m) {
        const double dm = (double)m;
        PL(KS_HLBOLT + m - 1) = exp(-(13.595 - 13.595 / (dm * dm)) / tkev) * 2.0 * (dm * dm) *
                                bhyd(S, m, j) * x1 / rho;
    }
    // STARK's F0 (:17576-17581)
    PL(KS_F0) = 1.25e-9 * pow(xne, 2.0 / 3.0);
    // FELO's Boltzmann factors (:14727-14731)
    for (int ion = 0; ion < 2; ++ion)
        for (int i = 0; i < 60; ++i)
            PL(KS_FELOE + ion * 60 + i) = i < T.felo_nlev[ion] ? exp(-T.felo_e[ion * 60 + i] * hckt) : 0.0;
    // MBF_HIGH_ION's Boltzmann divisors (:14383)
    for (int e = 0; e < 60; ++e) PL(KS_HIEXP + e) = e < T.hi_nedge ? exp(T.hi_edge[7 * e + 5] / tkev) : 1.0;
#undef PL
}

// XNFP(J, species) * TB_BOLTZ(level, J) of MBF_TOPBASE (:14249-14262, the
// first product of its sum at :14317)
A12_HD inline double kp_state_tb(const KTab &T, const KState &S, int level, int j)
{
    const double kt_ry = S.t[j] * 8.617333e-5 / 13.6056923;
    const double boltz = (double)T.tb_g[level] * exp(-T.tb_eexc[level] / kt_ry);
    return xnfp(S, T.tb_idx[T.tb_sp[level]], j) * boltz;
}

// SIGMA_NU of one TOPbase level at one wavelength (:14280-14310); 0 where
// MBF_TOPBASE skips the level (below its species' or its own threshold,
// or before its first tabulated energy).
A12_HD inline double kp_tb_sigma(const KTab &T, int lvl, double wave)
{
    const double freq = CLIGHT_NMS / wave;
    const double eph = freq / 3.289842e15;
    if (eph < T.tb_thr[T.tb_sp[lvl]]) return 0.0;
    if (eph < fabs(T.tb_ebind[lvl])) return 0.0;
    const double log_e = log10(eph);
    const double xi = (log_e - T.tb_logemin[lvl]) / T.tb_logestride[lvl];
    const int k = (int)xi + 1;
    if (k < 1) return 0.0;
    const int np = T.tb_np[lvl];
    const double *sg = T.tb_sigma + T.tb_off[lvl];
    double sig;
    if (k >= np) {
        const double r = T.tb_ephot[T.tb_off[lvl] + np - 1] / eph;
        sig = sg[np - 1] * ((r * r) * r);
    } else {
        sig = sg[k - 1] + (sg[k] - sg[k - 1]) * (xi - (double)(k - 1));
    }
    return sig * 1.0e-18;
}

// MBF_TOPBASE's sum at one cell over the levels in the reference's order,
// sigma(level) = sig(level * nw + w): the first part of ACONT_METAL.
A12_HD inline double kp_cell_tb(const KTab &T, const KState &S, const double *P, const double *sig, int nw,
                                int w, double wave, int j)
{
    const int n = S.n;
    const double freq = CLIGHT_NMS / wave;
    const double stim = 1.0 - exp(-freq * S.hkt[j]);
    const double stim_over_rho = stim / S.rho[j];
    double acm = 0.0;
    for (int l = 0; l < T.tb_nlev_tot; ++l) {
        const double sg = sig[(size_t)l * nw + w];
        if (sg == 0.0) continue;
        acm = acm + P[(size_t)(KS_TB + l) * n + j] * sg * stim_over_rho;
    }
    return acm;
}

// The rest of KAPP at one cell, in three parts that keep fewer values live
// (each a kernel of its own on the device), and KAPP's totals.  o[] holds
// the 17 per-source values in a12_kapp_eval's order.
// Part A: HOP, H2PLOP, HMINOP, H2MINOP, H2COLLOP, HRAYOP.
A12_HD inline void kp_cell_a(const KTab &T, const KState &S, const double *P, const double *row,
                             double wave, int j, double *o)
{
    const int n = S.n;
#define PL(k) P[(size_t)(k) * n + j]
    // RUN_RT_PASS's frequency set-up (:3905-3915)
    const double freq = CLIGHT_NMS / wave;
    const double waveno = 1.0e7 / wave;
    const double freqlg = log(freq);
    const double freq15 = freq / 1.0e15;
    const double ehvkt = exp(-freq * S.hkt[j]);
    const double stim = 1.0 - ehvkt;
    const double bnu = BNU_PREFAC * ((freq15 * freq15) * freq15) * ehvkt / stim;
    const double t = S.t[j], tkev = S.tkev[j], tlog = S.tlog[j], hckt = S.hckt[j], hkt = S.hkt[j];
    const double xne = S.xne[j], rho = S.rho[j];
    const double x1 = xnfp(S, 1, j), x2 = xnfp(S, 2, j), x3 = xnfp(S, 3, j);
    const double b1 = bhyd(S, 1, j);
    const double g1 = coulff(T, tlog, freqlg, 1);

    (void)waveno; (void)freqlg; (void)t; (void)tkev; (void)tlog; (void)hckt; (void)hkt; (void)xne;
    (void)rho; (void)x1; (void)x2; (void)x3; (void)b1; (void)g1; (void)bnu; (void)stim; (void)ehvkt;
    // HOP (:11384-11604)
    double ahyd, shyd = 0.0;
    {
        double H = 0.0, Sx = 0.0;
        for (int m = 15; m >= 1; --m) {
            if (waveno < H_SERIES_LIMITS[m - 1]) break;
            const double X = row[KF_XK + m - 1];
            if (m >= 7) {
                const double A = PL(KS_WOCC + m) * X * HOP_G[m] * PL(KS_HOPE + m) * stim;
                H = H + A;
                Sx = Sx + A * bnu;
            } else if (m >= 2) {
                const double bm = bhyd(S, m, j);
                const double A = X * HOP_G[m] * PL(KS_HOPE + m) * (bm - ehvkt);
                H = H + A;
                Sx = Sx + A * bnu * stim / (bm - ehvkt);
            } else {
                const double A = X * 2.0 * 1.0 * (b1 - ehvkt);
                H = H + A;
                Sx = Sx + A * bnu * stim / (b1 - ehvkt);
            }
        }
        for (int m = 16; m <= 80; ++m) {
            const double dm = (double)m;
            if (waveno < ELIM_HI / (dm * dm)) continue;
            const double w = PL(KS_WOCC + m);
            if (w < 1.0e-6) continue;
            const double A = w * row[KF_XK + m - 1] * 2.0 * (dm * dm) * PL(KS_HIE + m) * stim;
            H = H + A;
            Sx = Sx + A * bnu;
        }
        const double f3i = 1.0 / (freq * freq * freq);
        for (int m = 7; m <= 80; ++m) {
            const double w = PL(KS_WOCC + m);
            if (1.0 - w < 1.0e-6) continue;
            const double dm = (double)m;
            const double A = (1.0 - w) * 2.815e29 * f3i * 2.0 / ((dm * dm) * dm) * PL(KS_HIE + m) * stim;
            H = H + A;
            Sx = Sx + A * bnu;
        }
        H = H * x1 / rho;
        Sx = Sx * x1 / rho;
        const double A = COEFF_FF / sqrt(t) * g1 / freq * xne / freq * x2 / freq * stim / rho;
        H = H + A;
        Sx = Sx + A * bnu;
        ahyd = H;
        if (H > 0.0) shyd = Sx / H;
    }
    // H2PLOP (:12236-12250)
    double ah2p = 0.0;
    if (!(freq > FREQ_RYDH))
        ah2p = exp(-row[KF_H2PES] / tkev + row[KF_H2PFR]) * x1 * 2.0 * b1 * x2 / rho * stim;
    // HMINOP (:12514-12555)
    double ahmin, shmin;
    {
        const double bmin = S.bmin[j];
        const double ffth = linter(HM_THETAFF, row + KF_FFTT, 11, PL(KS_HMTHETA));
        const double hminff = ffth * x1 * 2.0 * b1 * xne / rho;
        const double H = row[KF_HMBF] * 1.0e-18 * (1.0 - ehvkt / bmin) * PL(KS_HMXHMIN) / rho;
        ahmin = H + hminff;
        shmin = (H * bnu * stim / (bmin - ehvkt) + hminff * bnu) / ahmin;
    }
    // H2MINOP (:12689-12701)
    double ah2min;
    {
        int it = 1;
        for (int i = 1; i <= 6; ++i)
            if (t >= H2M_T[i - 1]) it = i;
        double xt = (log(t > 1.0 ? t : 1.0) - T.h2m_logt[it - 1]) / (T.h2m_logt[it] - T.h2m_logt[it - 1]);
        xt = xt < 0.0 ? 0.0 : (xt > 1.0 ? 1.0 : xt);
        const double lkv = row[KF_H2MKT + it - 1] * (1.0 - xt) + row[KF_H2MKT + it] * xt;
        ah2min = exp(lkv) * xne * KBOL * t * PL(KS_XNH2M) / rho;
    }
    // H2COLLOP (:16505-16525)
    double ah2coll = 0.0;
    {
        const double numax = T.cia_nu0 + T.cia_dnu * (double)(T.cia_nnu - 1);
        if (!(waveno > numax) && !(waveno <= T.cia_nu0 && waveno <= 0.0)) {
            double dlow = 0.0, delnu;
            int inu;
            if (waveno <= T.cia_nu0) {
                inu = 1;
                delnu = 0.0;
                dlow = 2.0 * log10(waveno / T.cia_nu0);
            } else {
                inu = (int)((waveno - T.cia_nu0) / T.cia_dnu) + 1;
                inu = inu < T.cia_nnu - 1 ? inu : T.cia_nnu - 1;
                delnu = (waveno - T.cia_nu0 - T.cia_dnu * (double)(inu - 1)) / T.cia_dnu;
            }
            for (int ip = 0; ip < T.cia_npair; ++ip) {
                if (T.cia_use && !T.cia_use[ip]) continue;
                const double den = PL(KS_CIADEN + ip);
                if (den <= 0.0) continue;
                const int it = (int)PL(KS_CIAIT + ip);
                const double wt = PL(KS_CIAWT + ip);
                const double *tab = T.cia_tab + (size_t)ip * T.cia_ntmax * T.cia_nnu;
                const double *ra = tab + (size_t)(it - 1) * T.cia_nnu, *rb = tab + (size_t)it * T.cia_nnu;
                const double ca = ra[inu - 1] * (1.0 - delnu) + ra[inu] * delnu + dlow;
                const double cb = rb[inu - 1] * (1.0 - delnu) + rb[inu] * delnu + dlow;
                const double xlog = ca * (1.0 - wt) + cb * wt;
                ah2coll = ah2coll + pow(10.0, xlog) * den / rho;
            }
        }
    }
    // HRAYOP (:12899)
    const double sigh = row[KF_HRAY] * x1 * 2.0 * b1 / rho;
    o[0] = ahyd; o[1] = shyd; o[2] = ah2p; o[3] = ahmin; o[4] = shmin; o[5] = sigh;
    o[10] = ah2coll; o[11] = ah2min;
#undef PL
}

// Part B: HE1OP, HE2OP, HEMIOP, HERAOP, the rest of the metal continuum
// (acm_tb is MBF_TOPBASE's sum), ELECOP, H2RAOP.
A12_HD inline void kp_cell_b(const KTab &T, const KState &S, const double *P, const double *row,
                             double wave, int j, double acm_tb, double *o)
{
    const int n = S.n;
#define PL(k) P[(size_t)(k) * n + j]
    // RUN_RT_PASS's frequency set-up (:3905-3915)
    const double freq = CLIGHT_NMS / wave;
    const double waveno = 1.0e7 / wave;
    const double freqlg = log(freq);
    const double freq15 = freq / 1.0e15;
    const double ehvkt = exp(-freq * S.hkt[j]);
    const double stim = 1.0 - ehvkt;
    const double bnu = BNU_PREFAC * ((freq15 * freq15) * freq15) * ehvkt / stim;
    const double t = S.t[j], tkev = S.tkev[j], tlog = S.tlog[j], hckt = S.hckt[j], hkt = S.hkt[j];
    const double xne = S.xne[j], rho = S.rho[j];
    const double x1 = xnfp(S, 1, j), x2 = xnfp(S, 2, j), x3 = xnfp(S, 3, j);
    const double b1 = bhyd(S, 1, j);
    const double g1 = coulff(T, tlog, freqlg, 1);

    (void)waveno; (void)freqlg; (void)t; (void)tkev; (void)tlog; (void)hckt; (void)hkt; (void)xne;
    (void)rho; (void)x1; (void)x2; (void)x3; (void)b1; (void)g1; (void)bnu; (void)stim; (void)ehvkt;
    // HE1OP (:13114-13139)
    double ahe1;
    {
        const double freq3 = (freq * freq) * freq;
        const double cfree = COEFF_FF / freq3;
        const double c = 2.815e29 / freq3;
        double ex = PL(KS_HE1BOLTEX);
        if (freq < 2.055e14) ex = PL(KS_HE1EXLIM) / ehvkt;
        double he1 = (ex - PL(KS_HE1EXLIM)) * c;
        const int imin = (int)row[KF_HE1IMIN];
        if (imin > 0)
            for (int m = imin; m <= 10; ++m) he1 = he1 + row[KF_HE1T + m - 1] * PL(KS_HE1B + m - 1);
        if (freq >= 1.25408e16)
            for (int m = 4; m <= 27; ++m) he1 = he1 + row[KF_HE1N + m - 4] * PL(KS_HE1BN + m - 4);
        ahe1 = (he1 + g1 * PL(KS_HE1FREET) * cfree) * stim;
    }
    // HE2OP (:13581-13670)
    double ahe2;
    {
        const double freq3 = 2.815e29 / freq / freq / freq;
        const double xr = xnfp(S, 4, j) / rho;
        const double emax = 438908.85 - waveno > 434519.959 ? 438908.85 - waveno : 434519.959;
        double H = freq3 * 16.0 * 2.0 / 2.0 / (438889.068 * hckt) *
                   (exp(-emax * hckt) - PL(KS_HE2E + 8)) * stim * xr;
        for (int q = 0; q < 9; ++q) {
            if (waveno < HE2_LIM[q]) break;
            double X, A;
            switch (q) {
            case 0: X = freq3 / 59049.0 * 16.0; A = X * 162.0 * PL(KS_HE2E + 0) * stim * xr; break;
            case 1: X = freq3 * 16.0 / 32768.0; A = X * 128.0 * PL(KS_HE2E + 1) * stim * xr; break;
            case 2: X = freq3 * 16.0 / 16807.0; A = X * 98.0 * PL(KS_HE2E + 2) * stim * xr; break;
            case 3:
                X = freq3 * 16.0 / 7776.0 * (1.0986 + (-2.704e13 + 1.229e27 / freq) / freq);
                A = X * 72.0 * PL(KS_HE2E + 3) * stim * xr;
                break;
            case 4:
                X = freq3 * 16.0 / 3125.0 * (1.102 + (-3.909e13 + 2.371e27 / freq) / freq);
                A = X * 50.0 * PL(KS_HE2E + 4) * stim * xr;
                break;
            case 5:
                X = freq3 * 16.0 / 1024.0 * (1.101 + (-5.765e13 + 4.593e27 / freq) / freq);
                A = X * 32.0 * PL(KS_HE2E + 5) * stim * xr;
                break;
            case 6:
                X = freq3 * 16.0 / 243.0 * (1.101 + (-9.863e13 + 1.035e28 / freq) / freq);
                A = X * 18.0 * PL(KS_HE2E + 6) * stim * xr;
                break;
            case 7:
                X = freq3 * 16.0 / 32.0 * (1.105 + (-2.375e14 + 4.077e28 / freq) / freq);
                A = X * 8.0 * PL(KS_HE2E + 7) * stim * xr;
                break;
            default:
                X = freq3 * 16.0 / 1.0 * (0.9916 + (2.719e13 - 2.268e30 / freq) / freq);
                A = X * 2.0 * stim * xr;
                break;
            }
            H = H + A;
        }
        const double g2 = coulff(T, tlog, freqlg, 2);
        const double A = COEFF_FF * 4.0 / sqrt(t) * g2 / freq * xne / freq * xnfp(S, 5, j) / freq * stim / rho;
        ahe2 = H + A;
    }
    // HEMIOP (:13820-13829)
    double ahemin;
    {
        int it = 1;
        for (int i = 1; i <= 11; ++i)
            if (t >= HEM_T[i - 1]) it = i;
        double xt = (log(t > 1.0 ? t : 1.0) - T.hem_logt[it - 1]) / (T.hem_logt[it] - T.hem_logt[it - 1]);
        xt = xt < 0.0 ? 0.0 : (xt > 1.0 ? 1.0 : xt);
        const double lkv = row[KF_HEMKT + it - 1] * (1.0 - xt) + row[KF_HEMKT + it] * xt;
        ahemin = exp(lkv) * xne * KBOL * t * x3 / rho;
    }
    // HERAOP (:13939)
    const double sighe = row[KF_HERA] * x3 / rho;
    // CONT_METAL_OPACITY_TOPBASE (:14160-14194)
    double acm = acm_tb;
    {
        // MBF_HIGH_ION (:14338-14391)
        {
            double xnfs[5] = {0, 0, 0, 0, 0};
            const int base[8] = {22, 29, 37, 56, 79, 106, 137, 352};
            for (int z = 0; z < 5; ++z) {
                double sum = xnf(S, base[0] + z, j);
                for (int b = 1; b < 8; ++b) sum = sum + xnf(S, base[b] + z, j);
                xnfs[z] = sum;
            }
            const double free_ = g1 * 1.0 * xnfs[0] + coulff(T, tlog, freqlg, 2) * 4.0 * xnfs[1] +
                                 coulff(T, tlog, freqlg, 3) * 9.0 * xnfs[2] +
                                 coulff(T, tlog, freqlg, 4) * 16.0 * xnfs[3] +
                                 coulff(T, tlog, freqlg, 5) * 25.0 * xnfs[4];
            double ahot = free_ * COEFF_FF / (freq * freq * freq) * xne / sqrt(t);
            for (int e = 0; e < T.hi_nedge; ++e) {
                const double *h = T.hi_edge + 7 * e;
                if (freq < h[0]) continue;
                const double xx = row[KF_HI + e] * xnfp(S, (int)h[6], j) * h[4];
                if (xx > ahot / 100.0) ahot = ahot + xx / PL(KS_HIEXP + e);
            }
            acm = acm + ahot * stim / rho;
        }
        // FELO_OPACITY, Fe I then Fe II (:14710-14741)
        for (int ion = 0; ion < 2; ++ion) {
            double a = 0.0;
            for (int i = 0; i < T.felo_nlev[ion]; ++i) {
                const double sg = row[KF_FELO + 60 * ion + i];
                if (sg <= 0.0) continue;
                a = a + sg * T.felo_g[ion * 60 + i] * PL(KS_FELOE + 60 * ion + i);
            }
            a = a * stim * xnfp(S, T.felo_idx[ion], j) / rho;
            acm = acm + a;
        }
        // CHOP (:16103-16197) and OHOP (:16228-16323) below the Lyman limit
        if (freq <= FREQ_RYDH) {
            double chop = 0.0, ohop = 0.0;
            if (t < 9000.0) {
                int itp = (int)((t - 1000.0) / 200.0) + 1;
                itp = itp > 1 ? itp : 1;
                itp = itp < 40 ? itp : 40;
                const double tnp = (double)itp * 200.0 + 800.0;
                int itx = (int)((t - 2000.0) / 500.0) + 1;
                itx = itx > 1 ? itx : 1;
                itx = itx < 14 ? itx : 14;
                const double tnx = (double)itx * 500.0 + 1500.0;
                const int nch = (int)row[KF_CHN];
                if (!(nch < 20 || nch >= 105)) {
                    const double part = PART_CH[itp - 1] + (PART_CH[itp] - PART_CH[itp - 1]) * (t - tnp) / 200.0;
                    const double *c = row + KF_CH;
                    chop = exp((c[itx - 1] + (c[itx] - c[itx - 1]) * (t - tnx) / 500.0) * LN10) * part;
                }
                const int noh = (int)row[KF_OHN];
                if (!(noh <= 0 || noh >= 130)) {
                    const double part = PART_OH[itp - 1] + (PART_OH[itp] - PART_OH[itp - 1]) * (t - tnp) / 200.0;
                    const double *c = row + KF_OH;
                    ohop = exp((c[itx - 1] + (c[itx] - c[itx - 1]) * (t - tnx) / 500.0) * LN10) * part;
                }
            }
            acm = acm + (chop * xnfp(S, 846, j) + ohop * xnfp(S, 848, j)) * stim / rho;
        }
    }
    // ELECOP (:17220), H2RAOP (:17336)
    const double sigel = SIGMA_THOMSON * xne / rho;
    const double sigh2 = row[KF_H2RA] * PL(KS_XNH2) / rho;
    o[6] = ahe1; o[7] = ahe2; o[8] = ahemin; o[9] = sighe;
    o[12] = acm; o[13] = sigel; o[14] = sigh2;
#undef PL
}

// Part C: HLINOP.
A12_HD inline void kp_cell_c(const KTab &T, const KState &S, const double *P, const double *row,
                             double wave, int j, double *o)
{
    const int n = S.n;
#define PL(k) P[(size_t)(k) * n + j]
    // RUN_RT_PASS's frequency set-up (:3905-3915)
    const double freq = CLIGHT_NMS / wave;
    const double waveno = 1.0e7 / wave;
    const double freqlg = log(freq);
    const double freq15 = freq / 1.0e15;
    const double ehvkt = exp(-freq * S.hkt[j]);
    const double stim = 1.0 - ehvkt;
    const double bnu = BNU_PREFAC * ((freq15 * freq15) * freq15) * ehvkt / stim;
    const double t = S.t[j], tkev = S.tkev[j], tlog = S.tlog[j], hckt = S.hckt[j], hkt = S.hkt[j];
    const double xne = S.xne[j], rho = S.rho[j];
    const double x1 = xnfp(S, 1, j), x2 = xnfp(S, 2, j), x3 = xnfp(S, 3, j);
    const double b1 = bhyd(S, 1, j);
    const double g1 = 0.0;

    (void)waveno; (void)freqlg; (void)t; (void)tkev; (void)tlog; (void)hckt; (void)hkt; (void)xne;
    (void)rho; (void)x1; (void)x2; (void)x3; (void)b1; (void)g1; (void)bnu; (void)stim; (void)ehvkt;
    // HLINOP (:17419-17500)
    double ahline = 0.0, shline = 0.0;
    {
        const int nl = (int)row[KF_HLN];
        if (nl > 0) {
            const double f0s = PL(KS_F0);
            int m1 = (int)row[KF_HLMF];
            int m2 = m1 + 1;
            m1 = m1 > nl + 1 ? m1 : nl + 1;
            double H = 0.0, Sx = 0.0;
            bool dissolved = false;
            if (m1 > 6) {
                if (m1 <= 80) {
                    if (PL(KS_WOCC + m1) < 1.0e-3) dissolved = true;
                } else {
                    dissolved = true;
                }
                if (!dissolved) {
                    m1 = m1 - 1;
                    m2 = m2 + 3;
                    if (nl >= 4 && m1 <= 8) {
                        const double b3 = bhyd(S, 3, j), b4 = bhyd(S, 4, j);
                        H = stark_mmm(T, 3, 4, freq, t, xne, f0s, hkt) * (1.0 - ehvkt * b4 / b3) * PL(KS_HLBOLT + 2);
                        Sx = H * bnu * stim / (b3 / b4 - ehvkt);
                    }
                }
            }
            if (dissolved) {
                ahline = 0.0;
                shline = bnu;
            } else {
                const double bn = bhyd(S, nl, j);
                for (int m = m1; m <= m2; ++m) {
                    const double bm = m <= 6 ? bhyd(S, m, j) : 1.0;
                    double w = 1.0;
                    if (m >= 2 && m <= 80) w = PL(KS_WOCC + m);
                    if (m > 80) w = 0.0;
                    const double A = w * stark_mmm(T, nl, m, freq, t, xne, f0s, hkt) * (1.0 - ehvkt * bm / bn) *
                                     PL(KS_HLBOLT + nl - 1);
                    H = H + A;
                    Sx = Sx + A * bnu * stim / (bn / bm - ehvkt);
                }
                ahline = H;
                shline = H > 0.0 ? Sx / H : bnu;
            }
        }
    }
    o[15] = ahline; o[16] = shline;
#undef PL
}

// KAPP's totals (:11296-11320) from the 17 per-source values; out[0..22]
// in a12_kapp_eval's order.
A12_HD inline void kp_totals(const double *o, double wave, double hkt, double *out)
{
    const double freq = CLIGHT_NMS / wave;
    const double freq15 = freq / 1.0e15;
    const double ehvkt = exp(-freq * hkt);
    const double stim = 1.0 - ehvkt;
    const double bnu = BNU_PREFAC * ((freq15 * freq15) * freq15) * ehvkt / stim;
    const double ahyd = o[0], shyd = o[1], ah2p = o[2], ahmin = o[3], shmin = o[4], sigh = o[5];
    const double ahe1 = o[6], ahe2 = o[7], ahemin = o[8], sighe = o[9], ah2coll = o[10], ah2min = o[11];
    const double acm = o[12], sigel = o[13], sigh2 = o[14], ahline = o[15], shline = o[16];
    const double A = ah2p + ahe1 + ahe2 + ahemin + ah2min + ah2coll + acm;
    const double acont = A + ahyd + ahmin;
    double scont = bnu;
    if (acont > 0.0) scont = (A * bnu + ahyd * shyd + ahmin * shmin) / acont;
    const double aline = ahline + 0.0 + 0.0;
    double sline = bnu;
    if (aline > 0.0) sline = (ahline * shline + 0.0 * bnu + 0.0 * 0.0) / aline;
    const double sigmac = sigh + sighe + sigel + sigh2 + 0.0;
    for (int k = 0; k < 17; ++k) out[k] = o[k];
    out[17] = acont; out[18] = sigmac; out[19] = scont; out[20] = sline; out[21] = aline; out[22] = 0.0;
}

// All of KAPP at one cell (the host build's path; the device runs the parts
// as separate kernels).
A12_HD inline void kp_cell(const KTab &T, const KState &S, const double *P, const double *row,
                           double wave, int j, double acm_tb, double *out)
{
    double o[17];
    kp_cell_a(T, S, P, row, wave, j, o);
    kp_cell_b(T, S, P, row, wave, j, acm_tb, o);
    kp_cell_c(T, S, P, row, wave, j, o);
    kp_totals(o, wave, S.hkt[j], out);
}

}  // namespace kapp
