/*
 * a12_oracle_atlas.c -- ATLAS12 parts of the CPU oracle (test
 * infrastructure, the checker only): the Kurucz Voigt tables and function,
 * LINOP1's opacity-sampling line opacity, and the frequency loop of
 * RUN_RT_PASS with the TCORR / RADIAP / ROSS mode-1 and mode-2
 * accumulators.  Citations are cconroy20/atlas12 file:line.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "a12_oracle.h"
#include "a12_oracle_atlas.h"

#define KW_MAX 128

static const double C_SQRTPI = 1.7724538509055159;
static const double C_INVSQRTPI = 0.5641895835477563;
static const double C_PI = 3.141592653589793;
static const double C_CLIGHT_NMS = 2.99792458e17;
static const double C_HPLANCK = 6.62607015e-27;
static const double C_CLIGHT = 2.99792458e10;

/* TABVOIGT, src/atlas12_modules.f90:19796-19851: H0 = exp(-v^2),
 * H1 interpolated (MAP1) from the 81-point table of the Harris expansion,
 * H2 = (1 - 2 v^2) exp(-v^2), on n points at vsteps per Doppler width.
 * TABVI/TABH1 are the published table of the method. */
void or_tabvoigt(double vsteps, int n, double *h0, double *h1, double *h2)
{
    static const double TABVI[81] = {
        0.0, 0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9, 1.0, 1.1, 1.2, 1.3, 1.4, 1.5, 1.6,
        1.7, 1.8, 1.9, 2.0, 2.1, 2.2, 2.3, 2.4, 2.5, 2.6, 2.7, 2.8, 2.9, 3.0, 3.1, 3.2, 3.3,
        3.4, 3.5, 3.6, 3.7, 3.8, 3.9, 4.0, 4.2, 4.4, 4.6, 4.8, 5.0, 5.2, 5.4, 5.6, 5.8, 6.0,
        6.2, 6.4, 6.6, 6.8, 7.0, 7.2, 7.4, 7.6, 7.8, 8.0, 8.2, 8.4, 8.6, 8.8, 9.0, 9.2, 9.4,
        9.6, 9.8, 10.0, 10.2, 10.4, 10.6, 10.8, 11.0, 11.2, 11.4, 11.6, 11.8, 12.0};
    static const double TABH1[81] = {
        -1.12838, -1.10596, -1.04048, -0.93703, -0.80346, -0.64945, -0.48552, -0.32192,
        -0.16772, -0.03012, 0.08594, 0.17789, 0.24537, 0.28981, 0.31394, 0.32130, 0.31573,
        0.30094, 0.28027, 0.25648, 0.231726, 0.207528, 0.184882, 0.164341, 0.146128,
        0.130236, 0.116515, 0.104739, 0.094653, 0.086005, 0.078565, 0.072129, 0.066526,
        0.061615, 0.057281, 0.053430, 0.049988, 0.046894, 0.044098, 0.041561, 0.039250,
        0.035195, 0.031762, 0.028824, 0.026288, 0.024081, 0.022146, 0.020441, 0.018929,
        0.017582, 0.016375, 0.015291, 0.014312, 0.013426, 0.012620, 0.0118860, 0.0112145,
        0.0105990, 0.0100332, 0.0095119, 0.0090306, 0.0085852, 0.0081722, 0.0077885,
        0.0074314, 0.0070985, 0.0067875, 0.0064967, 0.0062243, 0.0059688, 0.0057287,
        0.0055030, 0.0052903, 0.0050898, 0.0049006, 0.0047217, 0.0045526, 0.0043924,
        0.0042405, 0.0040964, 0.0039595};
    for (int i = 0; i < n; ++i) h0[i] = (double)i / vsteps;
    (void)or_map1(TABVI, TABH1, 81, h0, h1, n);
    for (int i = 0; i < n; ++i) {
        double vv = ((double)i / vsteps) * ((double)i / vsteps);
        h0[i] = exp(-vv);
        h2[i] = h0[i] - (vv + vv) * h0[i];
    }
}

/* VOIGT(V, A), src/atlas12_modules.f90:21033-21097: Kurucz's three
 * regimes (Harris expansion from the tables for a < 0.2; a quartic
 * correction for a <= 1.4 and a + v <= 3.2; the corrected Lorentz
 * asymptote otherwise).  Tables are 1-based in the reference: index IV. */
double or_voigt_kurucz(double v, double a, const double *h0, const double *h1,
                       const double *h2)
{
    int iv = (int)(v * 200.0 + 1.5);
    const double H0 = iv >= 1 && iv <= 2001 ? h0[iv - 1] : 0.0;
    const double H1 = iv >= 1 && iv <= 2001 ? h1[iv - 1] : 0.0;
    const double H2 = iv >= 1 && iv <= 2001 ? h2[iv - 1] : 0.0;
    if (a < 0.2) {
        if (v <= 10.0) return (H2 * a + H1) * a + H0;
        return 0.5642 * a / (v * v);
    } else if (a <= 1.4 && a + v <= 3.2) {
        double vv = v * v;
        double hh1 = H1 + H0 * 1.12838;
        double hh2 = H2 + hh1 * 1.12838 - H0;
        double hh3 = (1.0 - H2) * 0.37613 - hh1 * 0.66667 * vv + hh2 * 1.12838;
        double hh4 = (3.0 * hh3 - hh1) * 0.37613 + H0 * 0.66667 * vv * vv;
        return ((((hh4 * a + hh3) * a + hh2) * a + hh1) * a + H0) *
               (((-.122727278 * a + .532770573) * a - .96284325) * a + .979895032);
    } else {
        double aa = a * a, vv = v * v;
        double u = (aa + vv) * 1.4142;
        double r = a * 0.79788 / u;
        if (a <= 100.0) {
            double aau = aa / u, vvu = vv / u, uu = u * u;
            r = ((((aau - 10.0 * vvu) * aau * 3.0 + 15.0 * vvu * vvu) + 3.0 * vv - aa) / uu + 1.0) * r;
        }
        return r;
    }
}

/* TABLOG, src/atlas12_modules.f90:2083-2086 */
static double tablog(int16_t i) { return pow(10.0, (double)(i - 16384) * .001); }

/* One wing profile value of LINOP1 (:18263-18298). */
static double linop1_cv(double center, double v, double adamp, const or_atlas_lines_in *in)
{
    if (adamp > 0.20) return center * or_voigt_kurucz(v, adamp, in->h0, in->h1, in->h2);
    if (v > 10.0) return center * C_INVSQRTPI * adamp / (v * v);
    int iv = (int)(v * 200.0 + 1.5);
    return center * ((in->h2[iv - 1] * adamp + in->h1[iv - 1]) * adamp + in->h0[iv - 1]);
}

/* LINOP1 (src/atlas12_modules.f90:18132-18374) over lines [l0, l1),
 * adding into xlines (nrhox, numnu), depth fastest.  The continuum bin and
 * grid index are walked forward from the previous line and restart when
 * IWL decreases (:18186-18227), so each line gets the first bin with
 * IWL < IWAVETAB and the first grid point with WAVESET > WLVAC. */
void or_linop1_range(const or_atlas_lines_in *in, int64_t l0, int64_t l1, double *xlines,
                     int64_t *evals)
{
    const int n = in->nrhox, numnu = in->numnu, nr = in->ldion;
    const double ratiolg = log(1.0 + 1.0 / 2000000.0);
    const double cen_prefac4 = 0.026538 / C_SQRTPI / C_CLIGHT_NMS;
    const double fourpi_c_inv = 1.0 / ((4.0 * C_PI) * C_CLIGHT_NMS);
    const double start = in->waveset[0] - 1.0, stop = in->waveset[numnu - 1] + 1.0;
    double txnxn[KW_MAX];
    int64_t ev = 0;
    for (int j = 0; j < n; ++j)
        txnxn[j] = (in->xnf1[j] + 0.42 * in->xnf3[j] + 0.85 * in->xnf841[j]) *
                   pow(in->t[j] / 10000.0, 0.3);
    int nucont = 0, nu = 0;
    int32_t iwlold = INT32_MIN;
    for (int64_t line = l0; line < l1; ++line) {
        const int32_t *rec = in->linedata + 4 * (size_t)line;
        const int32_t iwl = rec[0];
        int16_t p[6];
        memcpy(p, rec + 1, 12);
        const int16_t ielion = p[0], ielo = p[1], igflog = p[2], igr = p[3], igs = p[4],
                      igw = p[5];
        if (iwl < iwlold) nucont = nu = 0;
        iwlold = iwl;
        /* continuum bin: first k with IWL < IWAVETAB(k) (sentinel at the end) */
        while (nucont < in->nwave1 - 1 && iwl >= in->iwavetab[nucont]) nucont++;
        const int nelion = abs(ielion / 10);
        if (nelion < 1 || nelion > in->mion) continue;
        const double wlvac = exp(iwl * ratiolg);
        if (wlvac < start || wlvac > stop) continue;
        /* first grid point with WAVESET(NU) > WLVAC (NUMNU+1 past the end) */
        while (nu < numnu && wlvac >= in->waveset[nu]) nu++;
        const double cgf = cen_prefac4 * wlvac * tablog(igflog);
        const double elo = tablog(ielo);
        double adamp = 0.0, gammar = 0.0, gammas = 0.0, gammaw = 0.0;
        int ifj[KW_MAX + 2];
        ifj[0] = 0;
        for (int pass = 0; pass < 2; ++pass) {
            for (int K = 8; K <= n; K += 8) {
                int jlo, jhi;
                if (pass == 0) {
                    jlo = jhi = K;
                    ifj[K] = 0;
                } else {
                    if (ifj[K - 8] + ifj[K] == 0) continue;
                    jlo = K - 7;
                    jhi = K - 1;
                }
                for (int J = jlo; J <= jhi; ++J) {
                    const int j = J - 1;
                    const double tc = in->tabcont[(size_t)nucont * n + j];
                    double center = cgf * in->xnfdop[(size_t)(nelion - 1) * nr + j];
                    if (center < tc) continue;
                    center = center * exp(-elo * in->hckt[j]);
                    if (center < tc) continue;
                    if (pass == 0) {
                        ifj[K] = 1;
                        if (adamp == 0.0) {
                            gammar = tablog(igr) * wlvac * fourpi_c_inv;
                            gammas = tablog(igs) * wlvac * fourpi_c_inv;
                            gammaw = tablog(igw) * wlvac * fourpi_c_inv;
                        }
                    }
                    const double dop = in->dopple[(size_t)(nelion - 1) * nr + j];
                    adamp = (gammar + gammas * in->xne[j] + gammaw * txnxn[j]) / dop;
                    const double dopwave = dop * wlvac;
                    /* red wing: IW = NU .. min(NU+100, NUMNU) (1-based) */
                    int hi = nu + 1 + 100 < numnu ? nu + 1 + 100 : numnu;
                    for (int IW = nu + 1; IW <= hi; ++IW) {
                        double cv = linop1_cv(center, (in->waveset[IW - 1] - wlvac) / dopwave,
                                              adamp, in);
                        xlines[(size_t)(IW - 1) * n + j] += cv;
                        ++ev;
                        if (cv < tc) break;
                    }
                    for (int I = 1; I <= 100; ++I) {
                        int IW = nu + 1 - I;
                        if (IW <= 0) break;
                        double cv = linop1_cv(center, (wlvac - in->waveset[IW - 1]) / dopwave,
                                              adamp, in);
                        xlines[(size_t)(IW - 1) * n + j] += cv;
                        ++ev;
                        if (cv < tc) break;
                    }
                }
            }
        }
    }
    *evals = ev;
}

/* LINOP1, src/atlas12_modules.f90:18132-18374.  xlines (nrhox, numnu),
 * depth fastest, FP64.  evals: accumulated contributions. */
void or_linop1(const or_atlas_lines_in *in, double *xlines, int64_t *evals)
{
    int64_t ev = 0;
    memset(xlines, 0, sizeof(double) * (size_t)in->nrhox * in->numnu);
    or_linop1_range(in, 0, in->nlines, xlines, &ev);
    if (evals) *evals = ev;
}

/* The frequency loop of RUN_RT_PASS (src/atlas12_modules.f90:3905-3949)
 * for frequencies given by index list, with the KAPP outputs supplied as
 * tables, and TCORR(2) (:2465-2522), RADIAP(2) (:5084-5096), ROSS(2)
 * (:5180-5190) accumulated in frequency order after their mode-1 zeroing.
 * IFSCAT = 1, IFSURF = 0 (ATLAS12's own mode). */
void or_rt_pass(const or_rt_in *in, or_rt_acc *acc, double *hnu1_out)
{
    const int n = in->nrhox;
    double ehvkt[KW_MAX], stim[KW_MAX], bnu[KW_MAX], aline[KW_MAX];
    double taunu[KW_MAX], snu[KW_MAX], jnu[KW_MAX], hnu[KW_MAX], jmins[KW_MAX];
    double abtot[KW_MAX], alpha[KW_MAX], dabtot[KW_MAX];
    const double euler_m1 = 0.922784335098467;
    const double bnu_prefac = 2.0 * C_HPLANCK / (C_CLIGHT * C_CLIGHT) * 1.0e45;
    memset(acc, 0, sizeof *acc);
    for (int k = 0; k < in->nfreq; ++k) {
        const double wave = in->wave[k];
        const double freq = C_CLIGHT_NMS / wave;
        const double rcowt = in->rcowt[k];
        const double f15 = freq / 1.0e15;
        const size_t o = (size_t)k * n;
        for (int j = 0; j < n; ++j) {
            ehvkt[j] = exp(-freq * in->hkt[j]);
            stim[j] = 1.0 - ehvkt[j];
            bnu[j] = bnu_prefac * (f15 * f15 * f15) * ehvkt[j] / stim[j];
            aline[j] = in->aline[o + j] + in->xlines[o + j] * stim[j];
        }
        or_josh_io io;
        memset(&io, 0, sizeof io);
        io.nrhox = n;
        io.acont = in->acont + o;
        io.aline = aline;
        io.sigmac = in->sigmac + o;
        io.sigmal = in->sigmal + o;
        io.scont = in->scont + o;
        io.sline = in->sline + o;
        io.rhox = in->rhox;
        io.bnu = bnu;
        io.coefj = in->coefj;
        io.coefh = in->coefh;
        io.taunu = taunu;
        io.snu = snu;
        io.jnu = jnu;
        io.hnu = hnu;
        io.jmins = jmins;
        memset(snu, 0, sizeof snu);
        memset(jnu, 0, sizeof jnu);
        memset(hnu, 0, sizeof hnu);
        memset(jmins, 0, sizeof jmins);
        (void)or_josh(&io, 1, 0);
        const double knu1 = io.knu1;
        if (hnu1_out) hnu1_out[k] = hnu[0];
        int neg = 0;
        for (int j = 0; j < n; ++j)
            if (snu[j] < 0.0 || jnu[j] < 0.0 || hnu[j] < 0.0) { neg = 1; break; }
        if (neg)
            for (int j = 0; j < n; ++j) {
                jnu[j] = fmax(jnu[j], 1.0e-99);
                snu[j] = fmax(snu[j], 1.0e-99);
                hnu[j] = fmax(hnu[j], 1.0e-99);
            }
        memcpy(acc->last_hnu, hnu, sizeof(double) * n);
        memcpy(acc->last_taunu, taunu, sizeof(double) * n);
        /* JOSH's ABTOT and ALPHA (:18475-18476) */
        for (int j = 0; j < n; ++j) {
            abtot[j] = io.acont[j] + aline[j] + io.sigmac[j] + io.sigmal[j];
            alpha[j] = (io.sigmac[j] + io.sigmal[j]) / abtot[j];
        }
        /* TCORR(2) */
        or_deriv(in->rhox, abtot, dabtot, n);
        double term2 = 0.0, d = 0.0;
        for (int j = 0; j < n; ++j) {
            acc->rdabh[j] = acc->rdabh[j] + dabtot[j] / abtot[j] * hnu[j] * rcowt;
            acc->rjmins[j] = acc->rjmins[j] + abtot[j] * jmins[j] * rcowt;
            acc->flxrad[j] = acc->flxrad[j] + hnu[j] * rcowt;
        }
        for (int J = 1; J <= n; ++J) {
            const int j = J - 1;
            double term1 = term2;
            if (J != n) d = taunu[j + 1] - taunu[j];
            d = fmax(1.0e-10, d);
            if (d <= 0.01) {
                term2 = (euler_m1 - log(d)) * d / 4.0 + d * d / 12.0 - d * d * d / 96.0 +
                        d * d * d * d / 720.0;
            } else {
                double ex = 0.0;
                if (d < 10.0) ex = or_expi(3, d);
                if (in->teff <= 4250.0 && d > 0.005 && d < 0.02) ex = 0.0;
                term2 = 0.5 * (d + ex - 0.5) / d;
            }
            double diagj = term1 + term2;
            double dbdt = bnu[j] * freq * in->hkt[j] / in->t[j] / fmax(stim[j], 1.0e-30);
            const double rdiagjnu = abtot[j] * (diagj - 1.0) / (1.0 - alpha[j] * diagj) *
                                    (1.0 - alpha[j]) * dbdt * rcowt;
            acc->rdiagj[j] = acc->rdiagj[j] + rdiagjnu;
            acc->last_rdiagjnu[j] = rdiagjnu;
        }
        /* RADIAP(2) */
        for (int j = 0; j < n; ++j) {
            acc->raden[j] = acc->raden[j] + jnu[j] * rcowt;
            acc->h_total[j] = acc->h_total[j] + hnu[j] * rcowt;
            acc->accrad[j] = acc->accrad[j] + abtot[j] * hnu[j] * rcowt;
        }
        acc->pradk0 = acc->pradk0 + knu1 * rcowt;
        /* ROSS(2) */
        for (int j = 0; j < n; ++j) {
            double dbdt = bnu[j] * freq * in->hkt[j] / in->t[j] / stim[j];
            acc->abross[j] = acc->abross[j] + dbdt / abtot[j] * rcowt;
        }
    }
}


/* ------------------------------------------------------------------ */
/* SELECTLINES (src/atlas12_modules.f90:19171-19781), paths 1 and 5.   */
/* ------------------------------------------------------------------ */

/* XNFDOPMAX(NELION, NU) = max_J XNFDOP(J, NELION) / TABCONT(J, NU), from
 * 0 (:19247-19255); stored [nel][nu]. */
static double *select_xnfdopmax(const or_select_in *in)
{
    const int nw1 = in->nwave + 1;
    double *m = (double *)calloc((size_t)in->mion * nw1, sizeof(double));
    for (int nu = 0; nu < nw1; ++nu)
        for (int nel = 0; nel < in->mion; ++nel) {
            double v = 0.0;
            for (int j = 0; j < in->nrhox; ++j)
                v = fmax(v, in->xnfdop[(size_t)nel * in->ldion + j] / in->tabcont[(size_t)nu * in->nrhox + j]);
            m[(size_t)nel * nw1 + nu] = v;
        }
    return m;
}

static int32_t pack_pair(int16_t lo, int16_t hi)
{
    return (int32_t)((uint32_t)(uint16_t)lo | ((uint32_t)(uint16_t)hi << 16));
}

/* Float to default INTEGER as x86-64's truncating conversion gives it:
 * the "integer indefinite" INT32_MIN for NaN and out-of-range values
 * (the reference's INT(LOG10(0.0) ...) for an H2O line with E_lo = 0). */
static int32_t f2i_trunc(double x)
{
    if (!(x > -2147483649.0 && x < 2147483648.0)) return INT32_MIN;
    return (int32_t)x;
}

static int16_t abs16(int16_t v)   /* ABS of an INTEGER(2), wrapping */
{
    return (int16_t)(v < 0 ? -(int32_t)v : v);
}

void or_sort_indirect_by_int(int32_t *keys, int32_t *order, int64_t n)
{
    /* 1-based heap over keys[0..n-1] (:1985-2039) */
    if (n <= 1) return;
#define K(i) keys[(i) - 1]
#define O(i) order[(i) - 1]
#define SWAP(a, b) do { int32_t t_ = K(a); K(a) = K(b); K(b) = t_; \
                        t_ = O(a); O(a) = O(b); O(b) = t_; } while (0)
    for (int64_t i = n / 2; i >= 1; --i) {
        int64_t parent = i;
        for (;;) {
            int64_t child = 2 * parent;
            if (child > n) break;
            if (child + 1 <= n && K(child + 1) > K(child)) child = child + 1;
            if (K(parent) >= K(child)) break;
            SWAP(parent, child);
            parent = child;
        }
    }
    for (int64_t i = n; i >= 2; --i) {
        SWAP(1, i);
        int64_t parent = 1;
        for (;;) {
            int64_t child = 2 * parent;
            if (child > i - 1) break;
            if (child + 1 <= i - 1 && K(child + 1) > K(child)) child = child + 1;
            if (K(parent) >= K(child)) break;
            SWAP(parent, child);
            parent = child;
        }
    }
#undef SWAP
#undef O
#undef K
}

/* SELECTLINES' MOLCODES and ISOX tables (:19181-19200) */
static const int SEL_MOLCODES[46] = {
    8410, 8411, 8460, 8461, 8470, 8471, 8480, 8481, 8482, 8510, 8511, 8512, 8530, 8531, 8532,
    8580, 8581, 8582, 8583, 8584, 8620, 8621, 8622, 8623, 8640, 8641, 8642, 8643, 8680, 8681,
    8682, 8690, 8691, 8692, 8693, 8700, 8701, 8702, 8703, 8704, 8705, 8890, 8891, 8892, 8896,
    8960};
static const int SEL_ISOX[46] = {
    0, -4469, -5, -1955, -2, -2444, -1, -3398, -2690, -105, -996, -947, -35, -1331, -1516,
    -13, -2189, -2870, -1681, -4398, -1362, -77, -1022, -1626, -1237, -38, -1658, -2553, -10,
    -1960, -3910, -7, -1957, -2449, -4399, -6, -1956, -3403, -5353, -2695, -4645, -36, -1332,
    -1517, -2725, -2};

typedef struct {
    const or_select_in *in;
    const double *xm;
    int nw1;
    double hcktn;
    int nu;         /* the walk's NU, 0-based */
    double freq;    /* FREQ of the walk's last step */
    int32_t *out;
    int64_t nk;
} sel_walk;

static void sel_step(sel_walk *w, int32_t iwl)
{
    while (iwl >= w->in->iwavetab[w->nu]) {
        w->freq = C_CLIGHT_NMS / w->in->wavetab[w->nu];
        ++w->nu;
    }
}

static double sel_x(const sel_walk *w, int nelion)
{
    return w->xm[(size_t)(nelion - 1) * w->nw1 + w->nu];
}

static void sel_put(sel_walk *w, int32_t iwl, int16_t ielion, int16_t ielo, int16_t igflog,
                    int16_t igr, int16_t igs, int16_t igw)
{
    int32_t *o = w->out + 4 * w->nk++;
    o[0] = iwl;
    o[1] = pack_pair(ielion, ielo);
    o[2] = pack_pair(igflog, igr);
    o[3] = pack_pair(igs, igw);
}

/* The 16-byte paths: (1) predicted (:19262-19313), (2) observed low lines
 * (:19316-19350; the walk goes on from path 1), (3) hilines (:19355-19395),
 * (7) H3+ (:19723-19767).  Returns -1 for a NELION outside 1..mion. */
static int sel_rec16(sel_walk *w, const int32_t *rec, int64_t n, int path)
{
    const double cen_prefac = 0.026538 / C_SQRTPI;   /* :19179 */
    for (int64_t l = 0; l < n; ++l) {
        const int32_t *r = rec + 4 * l;
        const int32_t iwl = r[0];
        const int16_t ielion = (int16_t)(r[1] & 0xffff), ielo = (int16_t)((uint32_t)r[1] >> 16);
        int16_t igflog = (int16_t)(r[2] & 0xffff);
        const int16_t igr = (int16_t)((uint32_t)r[2] >> 16);
        sel_step(w, iwl);
        int nelion = abs(ielion / 10);
        if (path == 7) {
            const int kgflog = igflog;
            igflog = (int16_t)(kgflog - 1272 > 1 ? kgflog - 1272 : 1);
            nelion = 895;
        }
        if (nelion < 1 || nelion > w->in->mion) {
            if (path == 1) continue;
            return -1;
        }
        const double x = sel_x(w, nelion);
        if (path <= 2 ? x <= 1.0e-37 : x == 0.0) continue;
        const double cen = cen_prefac * tablog(igflog) * x / w->freq;
        if (cen < 1.0) continue;
        const double tablog8 = tablog(ielo);
        if (cen * exp(-tablog8 * w->hcktn) < 1.0) continue;
        if (path == 7)
            sel_put(w, iwl, ielion, ielo, igflog, igr, 1, 9384);
        else
            memcpy(w->out + 4 * w->nk++, r, 16);
    }
    return 0;
}

int64_t or_atlas_select_all(const or_select_in *in, const or_select_lists *L, int32_t *out)
{
    static const int TIO_ISOCORR[5] = {-1101, -1138, -131, -1259, -1272};   /* :19204 */
    static const int H2O_ISOCORR[4] = {-1, -3398, -2690, -5000};            /* :19206 */
    const double cen_prefac = 0.026538 / C_SQRTPI;                        /* :19179 */
    const double ratiolg = log(1.0 + 1.0 / 2000000.0);
    double *xm = select_xnfdopmax(in);
    sel_walk w = {in, xm, in->nwave + 1, in->hckt[in->nrhox - 1], 0, 0.0, out, 0};
    int64_t rc = 0;

    if (sel_rec16(&w, L->pred, L->npred, 1) < 0) { rc = -1; goto done; }
    if (sel_rec16(&w, L->lowobs, L->nlowobs, 2) < 0) { rc = -1; goto done; }
    w.nu = 0;
    if (sel_rec16(&w, L->hilines, L->nhilines, 3) < 0) { rc = -1; goto done; }

    /* (4) ASCII diatomics (:19417-19536): keys NINT(LOG(wl)/RATIOLG),
     * heapsorted, then walked in key order */
    if (L->ndia > 0) {
        const int64_t nd = L->ndia;
        int32_t *keys = (int32_t *)malloc(sizeof(int32_t) * nd);
        int32_t *order = (int32_t *)malloc(sizeof(int32_t) * nd);
        for (int64_t i = 0; i < nd; ++i) {
            keys[i] = (int32_t)lround(log(L->dia[i].wlvac_nm) / ratiolg);
            order[i] = (int32_t)(i + 1);
        }
        or_sort_indirect_by_int(keys, order, nd);
        w.nu = 0;
        int molcodeold = 0, imol = -1;
        for (int64_t ii = 0; ii < nd; ++ii) {
            const or_diatomic_rec *d = L->dia + (order[ii] - 1);
            const int32_t iwl = keys[ii];
            sel_step(&w, iwl);
            const int molcode = d->molcode;
            if (molcode != molcodeold) {
                molcodeold = molcode;
                imol = -1;
                for (int k = 0; k < 46; ++k)
                    if (molcode == SEL_MOLCODES[k]) { imol = k; break; }
                if (imol < 0) { rc = -2; break; }
            }
            if (imol < 0) { rc = -2; break; }
            const double gflog = d->gflog_dex - SEL_ISOX[imol] * 1.0e-3;
            const int16_t ielion = (int16_t)molcode;
            const int16_t ielo = (int16_t)(int32_t)lround(log10(fmax(d->elo_cm, 1.0e-10)) * 1000.0 + 16384.0);
            int16_t igflog = (int16_t)(int32_t)lround(gflog * 1000.0 + 16384.0);
            const int16_t igr = (int16_t)(int32_t)lround(d->gammar_log * 1000.0 + 16384.0);
            const int16_t igw = (int16_t)(int32_t)lround(d->gammaw_log * 1000.0 + 16384.0);
            const int kgflog = igflog;
            igflog = (int16_t)(kgflog + SEL_ISOX[imol] > 1 ? kgflog + SEL_ISOX[imol] : 1);
            const int nelion = abs(ielion / 10);
            const double x = sel_x(&w, nelion);
            if (x == 0.0) continue;
            const double cen = cen_prefac * tablog(igflog) * x / w.freq;
            if (cen < 1.0) continue;
            const double tablog8 = tablog(ielo);
            if (cen * exp(-tablog8 * w.hcktn) < 1.0) continue;
            sel_put(&w, iwl, ielion, ielo, igflog, igr, 1, igw);
        }
        free(keys);
        free(order);
        if (rc < 0) goto done;
    }

    /* (5) TiO MOL2 records, repacked (:19557-19647); (6) H2O (:19650-19720);
     * neither above 8000 K */
    if (in->teff <= 8000.0) {
        w.nu = 0;
        for (int64_t l = 0; l < L->nmol2; ++l) {
            const or_mol2_rec *m = L->mol2 + l;
            if (m->code != 822) continue;
            const int iso = m->iso - 45;
            if (iso < 1 || iso > 5) continue;
            const double wlv = (double)m->wl;
            if (wlv <= 0.0) continue;
            const int32_t iwl = (int32_t)(log(wlv) / ratiolg + 0.5);
            sel_step(&w, iwl);
            const int kgflog = (int)lround((double)m->loggf * 1000.0) + 16384;
            const int16_t igflog = (int16_t)(kgflog + TIO_ISOCORR[iso - 1] > 1 ? kgflog + TIO_ISOCORR[iso - 1] : 1);
            if (895 > in->mion) { rc = -1; goto done; }
            const double x = sel_x(&w, 895);
            if (x == 0.0) continue;
            const double cen = cen_prefac * tablog(igflog) * x / w.freq;
            if (cen < 1.0) continue;
            const double elotio = fmin(fabs((double)m->e1), fabs((double)m->e2));
            if (cen * exp(-elotio * w.hcktn) < 1.0) continue;
            const int16_t ielo = (int16_t)(int)(log10(fmax(elotio, 1.0e-10)) * 1000.0 + 16384.5);
            const double gr = m->gr100 == 0 ? 2.223e13 / (wlv * wlv)
                                            : pow(10.0, (double)m->gr100 * 0.01);
            const int16_t igr = (int16_t)(int)(log10(gr) * 1000.0 + 16384.5);
            sel_put(&w, iwl, (int16_t)(8949 + iso), ielo, igflog, igr, 1, 9384);
        }

        w.nu = 0;
        int16_t igr_walk = 0;
        for (int64_t l = 0; l < L->nh2o; ++l) {
            const or_h2o_rec *h = L->h2o + l;
            const int32_t iwl = h->iwl;
            while (iwl >= in->iwavetab[w.nu]) {
                w.freq = C_CLIGHT_NMS / in->wavetab[w.nu];
                /* GAMMAR = 2.474D-22 * FREQ**2 * 0.001 (a REAL(4) literal) */
                const double gammar = 2.474e-22 * (w.freq * w.freq) * (double)0.001f;
                igr_walk = (int16_t)f2i_trunc(log10(gammar) * 1000.0 + 16384.5);
                ++w.nu;
            }
            const int iso = h->ielo > 0 ? (h->igflog > 0 ? 1 : 2) : (h->igflog > 0 ? 3 : 4);
            const double elo = (double)abs16(h->ielo);
            const int kgflog = abs16(h->igflog);
            const int16_t igflog = (int16_t)(kgflog + H2O_ISOCORR[iso - 1] > 1 ? kgflog + H2O_ISOCORR[iso - 1] : 1);
            if (940 > in->mion) { rc = -1; goto done; }
            const double x = sel_x(&w, 940);
            if (x == 0.0) continue;
            const double cen = cen_prefac * tablog(igflog) * x / w.freq;
            if (cen < 1.0) continue;
            if (cen * exp(-elo * w.hcktn) < 1.0) continue;
            const int16_t ielo = (int16_t)f2i_trunc(log10(elo) * 1000.0 + 16384.5);
            sel_put(&w, iwl, (int16_t)(-(9399 + iso)), ielo, igflog, igr_walk, 1, 9384);
        }
    }

    w.nu = 0;
    if (sel_rec16(&w, L->h3p, L->nh3p, 7) < 0) { rc = -1; goto done; }
    rc = w.nk;
done:
    free(xm);
    return rc;
}

int64_t or_atlas_select(const or_select_in *in, const int32_t *pred, int64_t npred,
                        const or_mol2_rec *mol2, int64_t nmol2, int32_t *out)
{
    or_select_lists L;
    memset(&L, 0, sizeof L);
    L.pred = pred;
    L.npred = npred;
    L.mol2 = mol2;
    L.nmol2 = nmol2;
    return or_atlas_select_all(in, &L, out);
}


/* ------------------------------------------------------------------ */
/* SCALE_MODEL (src/atlas12_modules.f90:6967-7060)                     */
/* ------------------------------------------------------------------ */
/* REAL(8)**4 with the constant exponent written in the source: the
 * compiler expands it into products taken left to right, ((x*x)*x)*x
 * (checked against amdflang -O3 on 20,000 arguments; a run-time exponent
 * goes by repeated squaring instead, powi_f64). */
static double powi4(double x)
{
    return ((x * x) * x) * x;
}

int or_scale_model(int n, int kw, double tau1lg, double steplg, const double *sc,
                   double teff_new, double logg_new, const double *const in[7],
                   double *const out[10], double *so)
{
    const double sigma_sb = 5.670374419e-5, fourpi = 4.0 * C_PI;
    const double teff = sc[0], grav = sc[1], pradk0 = sc[2], pcon = sc[3], pturb0 = sc[4];
    const double gnew = pow(10.0, logg_new);
    double tauros[KW_MAX];
    /* step 1: TAUSTD(I) = 10**(TAU1LG + (I-1)*STEPLG) (:6979-6981) */
    for (int i = 0; i < kw; ++i) out[0][i] = pow(10.0, tau1lg + (double)i * steplg);
    /* step 2: TAUROS = INTEG(RHOX, ABROSS) from ABROSS(1)*RHOX(1) (:6984) */
    or_integ(in[0], in[4], tauros, n, in[4][0] * in[0][0]);
    /* step 3: every field onto TAUSTD with MAP1 (:6987-6998) */
    for (int f = 0; f < 7; ++f) (void)or_map1(tauros, in[f], n, out[0], out[1 + f], kw);
    for (int i = 0; i < kw; ++i) out[8][i] = out[6][i] + pradk0;  /* PRADK = PRAD + PRADK0 */
    so[0] = teff;
    so[1] = grav;
    so[4] = pradk0;
    /* step 4 (:7017-7033): sentinel, or no meaningful change */
    if (teff_new == 0.0) return 0;
    if (fabs(teff_new - teff) < 1.0 && fabs(gnew - grav) / grav < 1.0e-3) return 0;
    const double r4 = powi4(teff_new / teff);
    for (int i = 0; i < kw; ++i) {
        out[2][i] = out[2][i] * teff_new / teff;
        out[8][i] = out[8][i] * r4;
        out[6][i] = out[6][i] * r4;
    }
    const double pradk0n = pradk0 * r4;
    const double pzero = pcon + pradk0n + pturb0;
    for (int i = 0; i < kw; ++i) {
        double ptotal = out[3][i] + out[6][i] + 0.0;  /* P + PRAD + PTURB (PTURB = 0) */
        out[1][i] = ptotal / gnew;
        out[9][i] = ptotal + pzero;
    }
    so[0] = teff_new;
    so[1] = gnew;
    so[2] = log10(gnew);
    so[3] = sigma_sb / fourpi * powi4(teff_new);
    so[4] = pradk0n;
    so[5] = pzero;
    return 1;
}


/* ------------------------------------------------------------------ */
/* TCORR mode 3, sections A-F (src/atlas12_modules.f90:2528-2752)      */
/* ------------------------------------------------------------------ */
void or_tcorr3(const or_tcorr_in *in, double *rdabh, double *oldt1, double *out)
{
    const int n = in->n;
    const double fourpi = 4.0 * C_PI, sigma_sb = 5.670374419e-5, flux = in->flux;
    const double abfl = 1.0e-30, rdfl = 1.0e-30, delfl = 1.0e-10, hoff = 1.0e-3;
    double dtdrhx[KW_MAX], ddlt[KW_MAX], dabros[KW_MAX], cnvflx[KW_MAX], ddel[KW_MAX];
    double codrhx[KW_MAX], g[KW_MAX], gflux[KW_MAX], dtau[KW_MAX];
    double *t1 = out, *dtflux = out + n, *dtlamb = out + 2 * n, *dtsurf = out + 3 * n;
    double *flxerr = out + 4 * n, *flxdrv = out + 5 * n, *hratio = out + 6 * n;
    or_deriv(in->rhox, in->t, dtdrhx, n);        /* :2531-2533 */
    or_deriv(in->rhox, in->dltdlp, ddlt, n);
    or_deriv(in->rhox, in->abross, dabros, n);
    for (int j = 0; j < n; ++j)                  /* (A) :2539-2542 */
        cnvflx[j] = (in->ifconv == 1 && j + 1 >= 3) ? in->flxcnv[j] : 0.0;
    for (int j = 0; j < n; ++j) {                /* (B) :2551-2593 */
        ddel[j] = 1.0;
        const double abs_ = fmax(in->abross[j], abfl);
        rdabh[j] = rdabh[j] - in->flxrad[j] * dabros[j] / abs_;
        double del = 1.0, d = 0.0;
        if (cnvflx[j] > 0.0 && in->flxcnv0[j] > 0.0) {
            del = in->dltdlp[j] - in->grdadb[j];
            del = fmax(del, delfl);
            const double vco = 0.5 * in->mixlth *
                               sqrt(fmax(-0.5 * in->ptotal[j] / in->rho[j] * in->dlrdlt[j], 0.0));
            const double fluxco = 0.5 * in->rho[j] * in->heatcp[j] * in->t[j] * in->mixlth / fourpi;
            if (in->mixlth > 0.0 && vco > 0.0) {
                const double t4 = ((in->t[j] * in->t[j]) * in->t[j]) * in->t[j];
                d = 8.0 * sigma_sb * t4 / (abs_ * in->hscale[j] * in->rho[j]) / (fluxco * fourpi) / vco;
            }
            const double taub = abs_ * in->rho[j] * in->mixlth * in->hscale[j];
            d = d * (taub * taub) / (2.0 + taub * taub);
            d = d * d / 2.0;
            ddel[j] = (1.0 + d / (d + del)) / del;
        }
        double cnvfl = 0.0;
        if (fmax(in->flxrad[j], 1.0e-30) > 0.0)
            if (cnvflx[j] / fmax(in->flxrad[j], 1.0e-30) > 1.0e-3 &&
                in->flxcnv0[j] / fmax(in->flxrad[j], 1.0e-30) > 1.0e-3)
                cnvfl = cnvflx[j];
        codrhx[j] = (rdabh[j] + cnvfl * (dtdrhx[j] / in->t[j] * (1.0 - 9.0 * d / (d + del)) +
                                         1.5 * ddlt[j] / del * (1.0 + d / (d + del)))) /
                    (in->flxrad[j] + cnvflx[j] * 1.5 * in->dltdlp[j] / del * (1.0 + d / (d + del)));
    }
    codrhx[0] = 0.0;
    if (n > 1) codrhx[1] = 0.0;
    or_integ(in->rhox, codrhx, g, n, 0.0);       /* (C) :2613-2630 */
    for (int j = 0; j < n; ++j) {
        g[j] = exp(g[j]);
        gflux[j] = g[j] * (in->flxrad[j] + cnvflx[j] - flux) /
                   (in->flxrad[j] + cnvflx[j] * 1.5 * in->dltdlp[j] * ddel[j]);
    }
    or_integ(in->tauros, gflux, dtau, n, 0.0);
    for (int j = 0; j < n; ++j) {
        dtau[j] = dtau[j] / g[j];
        dtau[j] = fmax(-in->tauros[j] / 3.0, fmin(in->tauros[j] / 3.0, dtau[j]));
        dtflux[j] = -dtau[j] * dtdrhx[j] / fmax(in->abross[j], abfl);
    }
    const double maxdt = 0.04 * in->teff;
    for (int j = 0; j < n; ++j)                  /* (D) :2644-2685 */
        flxerr[j] = (in->flxrad[j] + cnvflx[j] - flux) / flux * 100.0;
    or_deriv(in->tauros, flxerr, flxdrv, n);
    for (int j = 0; j < n; ++j) hratio[j] = cnvflx[j] / (cnvflx[j] + fmax(in->flxrad[j], 1.0e-30));
    for (int j = 0; j < n; ++j) {
        if (hratio[j] < hoff)
            flxdrv[j] = in->rjmins[j] / fmax(in->abross[j], abfl) / flux * 100.0;
        dtlamb[j] = -flxdrv[j] * flux / 100.0 / fmax(fabs(in->rdiagj[j]), rdfl) *
                    copysign(1.0, in->rdiagj[j]) * fmax(in->abross[j], abfl);
        if (hratio[j] >= hoff) {
            dtlamb[j] = 0.0;
            for (int k = 1; k <= 5; ++k)
                if (j - k >= 0) dtlamb[j - k] = dtlamb[j - k] / 2.0;
        }
        dtlamb[j] = fmax(-maxdt, fmin(maxdt, dtlamb[j]));
    }
    /* (E) :2693-2712 */
    double dtsur = (flux - in->flxrad[0]) / flux * 0.25 * in->t[0];
    dtsur = fmax(-maxdt, fmin(maxdt, dtsur));
    double dum[KW_MAX], tinteg[KW_MAX], x1 = 0.1, x2 = 2.0, tone, ttwo;
    for (int j = 0; j < n; ++j) dum[j] = dtflux[j] + dtlamb[j];
    or_integ(in->tauros, dum, tinteg, n, 0.0);
    (void)or_map1(in->tauros, tinteg, n, &x1, &tone, 1);
    (void)or_map1(in->tauros, tinteg, n, &x2, &ttwo, 1);
    double tav = (ttwo - tone) / 2.0;
    if (dtsur * tav <= 0.0) tav = 0.0;
    if (fabs(tav) > fabs(dtsur)) tav = dtsur;
    dtsur = dtsur - tav;
    /* (F) :2717-2752 */
    for (int j = 0; j < n; ++j) {
        dtsurf[j] = dtsur;
        t1[j] = dtflux[j] + dtlamb[j] + dtsur;
    }
    for (int j = 0; j < n; ++j) {
        const int J = j + 1;
        if (in->ifconv == 1 && hratio[j] > 0.0) {
        } else if (in->ifconv == 1 && J >= n / 3) {
        } else if (in->iter == 1) {
        } else if (oldt1[j] * t1[j] > 0.0 && fabs(oldt1[j]) > fabs(t1[j])) {
            t1[j] = t1[j] * 1.25;
        } else if (oldt1[j] * t1[j] < 0.0) {
            t1[j] = t1[j] * 0.5;
        }
        oldt1[j] = t1[j];
    }
}

/* ---- NMOLEC gas-phase Newton, src/atlas12_modules.f90:9078-9217 ------- */
#define MOL_MAXEQ 64
int or_molec_newton(const or_moltab *tb, double xntot, const double *xab, const double *equilj,
                    double *xn, int *status)
{
    const int ne = tb->nequa, ne1 = ne + 1;
    double deq[MOL_MAXEQ * MOL_MAXEQ], eq[MOL_MAXEQ], eqold[MOL_MAXEQ];
#define DEQ(i, j) deq[((j) - 1) * ne + ((i) - 1)]
#define XN(k) xn[(k) - 1]
#define EQ(k) eq[(k) - 1]
    for (int k = 0; k < ne; ++k) eqold[k] = 0.0;
    const int charge = tb->idequa[ne - 1] >= 100;
    int it = 0;
    *status = 0;
    for (;;) {
        ++it;
        for (int k = 0; k < ne * ne; ++k) deq[k] = 0.0;
        EQ(1) = -xntot;                        /* total particles :9084 */
        for (int k = 2; k <= ne; ++k) {
            EQ(1) = EQ(1) + XN(k);
            DEQ(1, k) = 1.0;
            EQ(k) = XN(k) - xab[k - 1] * XN(1);  /* element k :9092 */
            DEQ(k, k) = 1.0;
            DEQ(k, 1) = -xab[k - 1];
        }
        if (charge) {                          /* charge :9099-9102 */
            EQ(ne) = -XN(ne);
            DEQ(ne, ne) = -1.0;
        }
        for (int jm = 0; jm < tb->nummol; ++jm) {  /* molecules :9105-9162 */
            const int l1 = tb->locj[jm], l2 = tb->locj[jm + 1] - 1;
            if (l2 - l1 + 1 == 1) continue;
            double term = equilj[jm];
            for (int lk = l1; lk <= l2; ++lk) {
                const int k = tb->kcomps[lk - 1];
                if (k == ne1) term = term / XN(ne);
                else term = term * XN(k);
            }
            EQ(1) = EQ(1) + term;
            for (int lk = l1; lk <= l2; ++lk) {
                int k = tb->kcomps[lk - 1];
                double d;
                if (k >= ne1) {
                    k = ne;
                    d = -term / XN(k);
                } else {
                    d = term / XN(k);
                }
                EQ(k) = EQ(k) + term;
                DEQ(1, k) = DEQ(1, k) + d;
                for (int lm = l1; lm <= l2; ++lm) {
                    int m = tb->kcomps[lm - 1];
                    if (m == ne1) m = ne;
                    DEQ(m, k) = DEQ(m, k) + d;
                }
            }
            /* negative ions: the last component is the electron equation
             * (IDEQUA beyond NEQUA reads as 0) :9145-9160 */
            const int kl = tb->kcomps[l2 - 1];
            if (kl <= ne && tb->idequa[kl - 1] == 100) {
                for (int lk = l1; lk <= l2; ++lk) {
                    const int k = tb->kcomps[lk - 1];
                    const double d = term / XN(k);
                    if (k == ne) EQ(k) = EQ(k) - term - term;
                    for (int lm = l1; lm <= l2; ++lm) {
                        const int m = tb->kcomps[lm - 1];
                        if (m == ne) DEQ(m, k) = DEQ(m, k) - d - d;
                    }
                }
            }
        }
        or_solvit(deq, ne, eq);                /* :9165 */
        int iferr = 0;                         /* damped update :9168-9189 */
        double scale = 100.0;
        for (int k = 1; k <= ne; ++k) {
            const double ratio = fabs(EQ(k) / XN(k));
            if (ratio > 1.0e-4) iferr = 1;
            if (eqold[k - 1] * EQ(k) < 0.0) EQ(k) = EQ(k) * 0.69;
            const double xneq = XN(k) - EQ(k);
            const double xn100 = XN(k) / 100.0;
            if (fabs(xneq) >= xn100) {
                XN(k) = fabs(xneq);
            } else {
                XN(k) = XN(k) / scale;
                if (eqold[k - 1] * EQ(k) < 0.0) scale = sqrt(scale);
            }
            eqold[k - 1] = EQ(k);
        }
        if (XN(1) != XN(1) || EQ(1) != EQ(1)) { /* :9193-9202 */
            *status = 1;
            break;
        }
        if (iferr == 0) break;
        if (it >= 5000) {
            *status = 2;
            break;
        }
    }
#undef DEQ
#undef XN
#undef EQ
    return it;
}
