// a12_kapp_host.h -- host side of KAPP: checks the caller's a12_kapp_tables,
// forms the first-call logarithms of HMINOP/H2MINOP/HEMIOP and the
// physical constants that need libm, and packs everything into three flat
// arrays (double, int, float) whose offsets a KTab resolves on either side.
// Included by a12_kapp.hip.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "a12gpu.h"

namespace kapp {

struct KPack {
    std::vector<double> d;
    std::vector<int> i;
    std::vector<float> f;
    // offsets into d / i / f of every KTab array
    size_t gff, cia_temp, cia_tab, cross_ch, cross_oh, hi_edge, felo_nuth, felo_g, felo_e, felo_x,
        felo_logsig, felo_seaton, tb_thr, tb_ebind, tb_eexc, tb_logemin, tb_logestride, tb_ephot,
        tb_sigma, st_dens[4], st_temp[4], st_ldalpha[4], st_kalpha[4], st_prof[4], hm_wfflog, hm_fflog,
        h2m_logw, h2m_logt, h2m_logk, hem_logw, hem_logt, hem_logk;
    size_t cia_nt, cia_use, felo_ifancy, felo_npts, tb_idx, tb_nlev, tb_g, tb_np, tb_off, tb_sp, st_maxdens[4];
    size_t k_freqn, k_xn, k_xl, k_ekarzas;
    KTab scal;   // the scalar members (counts, constants); pointers unset
};

inline size_t put(std::vector<double> &v, const double *p, size_t n)
{
    const size_t o = v.size();
    v.insert(v.end(), p, p + n);
    return o;
}
inline size_t put(std::vector<int> &v, const int32_t *p, size_t n)
{
    const size_t o = v.size();
    v.insert(v.end(), p, p + n);
    return o;
}
inline size_t put(std::vector<float> &v, const float *p, size_t n)
{
    const size_t o = v.size();
    v.insert(v.end(), p, p + n);
    return o;
}

// The constants mod_constants forms with libm (SAHA_PREFAC's **1.5) and
// COULFF's two log10 parameters.
inline double saha_prefac() { return 2.0 * std::pow(2.0 * PI * EMASS * KBOL / (HPLANCK * HPLANCK), 1.5); }

// Packs the caller's tables; returns an empty string or what is wrong.
inline std::string kapp_pack(const a12_kapp_tables *t, KPack &P)
{
    if (!t) return "null tables";
    if (!t->k_freqn || !t->k_xn || !t->k_xl || !t->k_ekarzas) return "Karzas tables missing";
    if (!t->gff_logtab) return "free-free Gaunt table missing";
    if (t->cia_npair < 1 || t->cia_npair > 4 || t->cia_ntmax < 2 || t->cia_nnu < 2 || !t->cia_nt ||
        !t->cia_temp || !t->cia_tab || !(t->cia_dnu > 0.0))
        return "CIA tables missing or inconsistent";
    for (int p = 0; p < t->cia_npair; ++p)
        if (t->cia_nt[p] < 2 || t->cia_nt[p] > t->cia_ntmax) return "CIA temperature count out of range";
    if (!t->cross_ch || !t->cross_oh) return "CH/OH cross sections missing";
    if (t->hi_nedge < 0 || t->hi_nedge > 60 || (t->hi_nedge > 0 && !t->hi_edge))
        return "hotop edges missing or more than 60";
    for (int e = 0; e < t->hi_nedge; ++e) {
        const int id = (int)t->hi_edge[7 * e + 6];
        if (id < 1 || id > 1006) return "hotop species index out of range";
    }
    for (int k = 0; k < 2; ++k) {
        if (t->felo_nlev[k] < 0 || t->felo_nlev[k] > 60 || t->felo_xnfp_idx[k] < 1 || t->felo_xnfp_idx[k] > 1006)
            return "Fe I/II level table out of range";
    }
    if (!t->felo_nuth || !t->felo_g || !t->felo_e || !t->felo_ifancy || !t->felo_npts || !t->felo_x ||
        !t->felo_logsig || !t->felo_seaton)
        return "Fe I/II tables missing";
    for (int k = 0; k < 120; ++k)
        if (t->felo_npts[k] < 0 || t->felo_npts[k] > 200) return "Fe I/II fit length out of range";
    if (t->tb_nspecies < 0 || t->tb_nspecies > 30 || t->tb_nlevels < 0 || t->tb_npoints < 0)
        return "TOPbase counts out of range";
    if (t->tb_nspecies > 0 && (!t->tb_xnfp_idx || !t->tb_nlev || !t->tb_lowest_thr || !t->tb_g || !t->tb_np ||
                               !t->tb_ebind || !t->tb_eexc || !t->tb_logemin || !t->tb_logestride ||
                               !t->tb_ephot || !t->tb_sigma))
        return "TOPbase tables missing";
    long nl = 0, npt = 0;
    for (int s = 0; s < t->tb_nspecies; ++s) {
        if (t->tb_xnfp_idx[s] < 1 || t->tb_xnfp_idx[s] > 1006 || t->tb_nlev[s] < 0) return "TOPbase species out of range";
        nl += t->tb_nlev[s];
    }
    if (nl != t->tb_nlevels) return "TOPbase level count differs from the species' sum";
    for (int l = 0; l < t->tb_nlevels; ++l) {
        if (t->tb_np[l] < 1) return "TOPbase level without points";
        npt += t->tb_np[l];
    }
    if (npt != t->tb_npoints) return "TOPbase point count differs from the levels' sum";
    for (int s = 0; s < 4; ++s) {
        const a12_kapp_stark &st = t->stark[s];
        if (!st.loaded) continue;
        const int nt = st.n_upper_max - st.n_upper_min + 1;
        if (nt < 1 || st.n_dens < 2 || st.n_dens > 20 || !st.dens || !st.temp || !st.log_dalpha ||
            !st.max_dens_idx || !st.k_alpha || !st.profiles)
            return "Stark table inconsistent";
        for (int q = 0; q < nt; ++q)
            if (st.max_dens_idx[q] < 1 || st.max_dens_idx[q] > st.n_dens) return "Stark density index out of range";
    }

    P = KPack();
    KTab &S = P.scal;
    std::memset(&S, 0, sizeof(S));
    P.k_freqn = put(P.f, t->k_freqn, 15 * 29);
    P.k_xn = put(P.f, t->k_xn, 15 * 29);
    P.k_xl = put(P.f, t->k_xl, 6 * 6 * 29);
    P.k_ekarzas = put(P.f, t->k_ekarzas, 29);
    P.gff = put(P.d, t->gff_logtab, 146 * 81);
    S.cia_npair = t->cia_npair;
    S.cia_ntmax = t->cia_ntmax;
    S.cia_nnu = t->cia_nnu;
    S.cia_nu0 = t->cia_nu0;
    S.cia_dnu = t->cia_dnu;
    P.cia_nt = put(P.i, t->cia_nt, t->cia_npair);
    std::vector<int32_t> use(t->cia_npair, 1);
    if (t->cia_use)
        for (int p = 0; p < t->cia_npair; ++p) use[p] = t->cia_use[p] != 0;
    P.cia_use = put(P.i, use.data(), use.size());
    P.cia_temp = put(P.d, t->cia_temp, (size_t)t->cia_npair * t->cia_ntmax);
    P.cia_tab = put(P.d, t->cia_tab, (size_t)t->cia_npair * t->cia_ntmax * t->cia_nnu);
    P.cross_ch = put(P.d, t->cross_ch, 105 * 15);
    P.cross_oh = put(P.d, t->cross_oh, 130 * 15);
    S.hi_nedge = t->hi_nedge;
    const double zero7[7] = {0, 0, 0, 0, 0, 0, 1};
    P.hi_edge = t->hi_nedge > 0 ? put(P.d, t->hi_edge, 7 * (size_t)t->hi_nedge) : put(P.d, zero7, 7);
    for (int k = 0; k < 2; ++k) {
        S.felo_nlev[k] = t->felo_nlev[k];
        S.felo_idx[k] = t->felo_xnfp_idx[k];
    }
    P.felo_nuth = put(P.d, t->felo_nuth, 120);
    P.felo_g = put(P.d, t->felo_g, 120);
    P.felo_e = put(P.d, t->felo_e, 120);
    P.felo_x = put(P.d, t->felo_x, 120 * 200);
    P.felo_logsig = put(P.d, t->felo_logsig, 120 * 200);
    P.felo_seaton = put(P.d, t->felo_seaton, 120 * 4);
    P.felo_ifancy = put(P.i, t->felo_ifancy, 120);
    P.felo_npts = put(P.i, t->felo_npts, 120);
    const int32_t izero = 0;
    const double dzero = 0.0;
    const bool tb = t->tb_nspecies > 0 && t->tb_nlevels > 0;
    S.tb_nsp = tb ? t->tb_nspecies : 0;
    S.tb_nlev_tot = tb ? t->tb_nlevels : 0;
    P.tb_idx = tb ? put(P.i, t->tb_xnfp_idx, t->tb_nspecies) : put(P.i, &izero, 1);
    P.tb_nlev = tb ? put(P.i, t->tb_nlev, t->tb_nspecies) : put(P.i, &izero, 1);
    P.tb_thr = tb ? put(P.d, t->tb_lowest_thr, t->tb_nspecies) : put(P.d, &dzero, 1);
    P.tb_g = tb ? put(P.i, t->tb_g, t->tb_nlevels) : put(P.i, &izero, 1);
    P.tb_np = tb ? put(P.i, t->tb_np, t->tb_nlevels) : put(P.i, &izero, 1);
    std::vector<int32_t> off(tb ? t->tb_nlevels : 1, 0);
    for (int l = 1; l < t->tb_nlevels; ++l) off[l] = off[l - 1] + t->tb_np[l - 1];
    P.tb_off = put(P.i, off.data(), off.size());
    std::vector<int32_t> sp(tb ? t->tb_nlevels : 1, 0);
    for (int q = 0, l = 0; tb && q < t->tb_nspecies; ++q)
        for (int m = 0; m < t->tb_nlev[q]; ++m) sp[l++] = q;
    P.tb_sp = put(P.i, sp.data(), sp.size());
    P.tb_ebind = tb ? put(P.d, t->tb_ebind, t->tb_nlevels) : put(P.d, &dzero, 1);
    P.tb_eexc = tb ? put(P.d, t->tb_eexc, t->tb_nlevels) : put(P.d, &dzero, 1);
    P.tb_logemin = tb ? put(P.d, t->tb_logemin, t->tb_nlevels) : put(P.d, &dzero, 1);
    P.tb_logestride = tb ? put(P.d, t->tb_logestride, t->tb_nlevels) : put(P.d, &dzero, 1);
    P.tb_ephot = tb ? put(P.d, t->tb_ephot, t->tb_npoints) : put(P.d, &dzero, 1);
    P.tb_sigma = tb ? put(P.d, t->tb_sigma, t->tb_npoints) : put(P.d, &dzero, 1);
    for (int s = 0; s < 4; ++s) {
        const a12_kapp_stark &st = t->stark[s];
        S.st_loaded[s] = st.loaded ? 1 : 0;
        S.st_numin[s] = st.n_upper_min;
        S.st_numax[s] = st.n_upper_max;
        S.st_ndens[s] = st.n_dens;
        const int nt = st.loaded ? st.n_upper_max - st.n_upper_min + 1 : 0;
        P.st_dens[s] = st.loaded ? put(P.d, st.dens, 20) : put(P.d, &dzero, 1);
        P.st_temp[s] = st.loaded ? put(P.d, st.temp, 10) : put(P.d, &dzero, 1);
        P.st_ldalpha[s] = st.loaded ? put(P.d, st.log_dalpha, 60) : put(P.d, &dzero, 1);
        P.st_kalpha[s] = st.loaded ? put(P.d, st.k_alpha, nt) : put(P.d, &dzero, 1);
        P.st_prof[s] = st.loaded ? put(P.d, st.profiles, (size_t)nt * st.n_dens * 600) : put(P.d, &dzero, 1);
        P.st_maxdens[s] = st.loaded ? put(P.i, st.max_dens_idx, nt) : put(P.i, &izero, 1);
    }
    // HMINOP's first call (:12497-12504)
    double v[22 * 17];
    for (int iw = 0; iw < 22; ++iw) v[iw] = std::log(91.134 / HM_WAVEK[iw]);
    P.hm_wfflog = put(P.d, v, 22);
    for (int iw = 0; iw < 22; ++iw)
        for (int it = 0; it < 11; ++it) v[iw * 11 + it] = std::log(HM_FF[iw * 11 + it] * 1.0e-26);
    P.hm_fflog = put(P.d, v, 22 * 11);
    // H2MINOP's (:12656-12667) and HEMIOP's (:13788-13799)
    for (int i = 0; i < 10; ++i) v[i] = std::log(H2M_W[i]);
    P.h2m_logw = put(P.d, v, 10);
    for (int i = 0; i < 7; ++i) v[i] = std::log(H2M_T[i]);
    P.h2m_logt = put(P.d, v, 7);
    for (int i = 0; i < 70; ++i) v[i] = std::log(H2M_K[i]);
    P.h2m_logk = put(P.d, v, 70);
    for (int i = 0; i < 17; ++i) v[i] = std::log(HEM_W[i]);
    P.hem_logw = put(P.d, v, 17);
    for (int i = 0; i < 12; ++i) v[i] = std::log(HEM_T[i]);
    P.hem_logt = put(P.d, v, 12);
    for (int i = 0; i < 204; ++i) v[i] = std::log(HEM_K[i]);
    P.hem_logk = put(P.d, v, 204);
    S.saha_prefac = saha_prefac();
    S.lg_ryh_over_k = std::log10(FREQ_RYDH * HOVERK);
    S.lg_k_over_h = -std::log10(HOVERK);
    return "";
}

// A KTab over packed arrays at d / i / f (host or device copies of P).
inline KTab kapp_tab(const KPack &P, const double *d, const int *i, const float *f)
{
    KTab T = P.scal;
    T.k_freqn = f + P.k_freqn;
    T.k_xn = f + P.k_xn;
    T.k_xl = f + P.k_xl;
    T.k_ekarzas = f + P.k_ekarzas;
    T.gff = d + P.gff;
    T.cia_nt = i + P.cia_nt;
    T.cia_use = i + P.cia_use;
    T.cia_temp = d + P.cia_temp;
    T.cia_tab = d + P.cia_tab;
    T.cross_ch = d + P.cross_ch;
    T.cross_oh = d + P.cross_oh;
    T.hi_edge = d + P.hi_edge;
    T.felo_nuth = d + P.felo_nuth;
    T.felo_g = d + P.felo_g;
    T.felo_e = d + P.felo_e;
    T.felo_x = d + P.felo_x;
    T.felo_logsig = d + P.felo_logsig;
    T.felo_seaton = d + P.felo_seaton;
    T.felo_ifancy = i + P.felo_ifancy;
    T.felo_npts = i + P.felo_npts;
    T.tb_idx = i + P.tb_idx;
    T.tb_nlev = i + P.tb_nlev;
    T.tb_g = i + P.tb_g;
    T.tb_np = i + P.tb_np;
    T.tb_off = i + P.tb_off;
    T.tb_sp = i + P.tb_sp;
    T.tb_thr = d + P.tb_thr;
    T.tb_ebind = d + P.tb_ebind;
    T.tb_eexc = d + P.tb_eexc;
    T.tb_logemin = d + P.tb_logemin;
    T.tb_logestride = d + P.tb_logestride;
    T.tb_ephot = d + P.tb_ephot;
    T.tb_sigma = d + P.tb_sigma;
    for (int s = 0; s < 4; ++s) {
        T.st_dens[s] = d + P.st_dens[s];
        T.st_temp[s] = d + P.st_temp[s];
        T.st_ldalpha[s] = d + P.st_ldalpha[s];
        T.st_kalpha[s] = d + P.st_kalpha[s];
        T.st_prof[s] = d + P.st_prof[s];
        T.st_maxdens[s] = i + P.st_maxdens[s];
    }
    T.hm_wfflog = d + P.hm_wfflog;
    T.hm_fflog = d + P.hm_fflog;
    T.h2m_logw = d + P.h2m_logw;
    T.h2m_logt = d + P.h2m_logt;
    T.h2m_logk = d + P.h2m_logk;
    T.hem_logw = d + P.hem_logw;
    T.hem_logt = d + P.hem_logt;
    T.hem_logk = d + P.hem_logk;
    return T;
}

// Checks a state; returns an empty string or what is wrong.
inline std::string kapp_check_state(const a12_kapp_state *s)
{
    if (!s) return "null state";
    if (s->nrhox < 1 || s->nrhox > 128 || s->ld < s->nrhox || s->mion < 848)
        return "nrhox, ld or mion out of range (1 <= nrhox <= ld, nrhox <= 128, mion >= 848)";
    if (!s->t || !s->tkev || !s->tlog || !s->hkt || !s->hckt || !s->xne || !s->rho || !s->bmin ||
        !s->bhyd || !s->equilh2 || !s->xnf || !s->xnfp)
        return "null state array";
    return "";
}

// The state's arrays packed depth-fastest with leading dimension nrhox:
// t tkev tlog hkt hckt xne rho bmin eqh2 (9 x n), bhyd (6 x n), xnf, xnfp
// (mion x n each).
inline void kapp_pack_state(const a12_kapp_state *s, std::vector<double> &out)
{
    const int n = s->nrhox, ld = s->ld, mion = s->mion;
    out.assign((size_t)(9 + 6 + 2 * mion) * n, 0.0);
    const double *v[9] = {s->t, s->tkev, s->tlog, s->hkt, s->hckt, s->xne, s->rho, s->bmin, s->equilh2};
    for (int k = 0; k < 9; ++k) std::memcpy(&out[(size_t)k * n], v[k], sizeof(double) * n);
    for (int m = 0; m < 6; ++m) std::memcpy(&out[(size_t)(9 + m) * n], s->bhyd + (size_t)m * ld, sizeof(double) * n);
    for (int m = 0; m < mion; ++m) {
        std::memcpy(&out[(size_t)(15 + m) * n], s->xnf + (size_t)m * ld, sizeof(double) * n);
        std::memcpy(&out[(size_t)(15 + mion + m) * n], s->xnfp + (size_t)m * ld, sizeof(double) * n);
    }
}

inline KState kapp_state_view(const double *p, int n, int mion)
{
    KState S;
    S.n = n;
    S.mion = mion;
    S.t = p;
    S.tkev = p + n;
    S.tlog = p + 2 * n;
    S.hkt = p + 3 * n;
    S.hckt = p + 4 * n;
    S.xne = p + 5 * n;
    S.rho = p + 6 * n;
    S.bmin = p + 7 * n;
    S.eqh2 = p + 8 * n;
    S.bhyd = p + 9 * n;
    S.xnf = p + (size_t)15 * n;
    S.xnfp = p + (size_t)(15 + mion) * n;
    return S;
}

inline int kapp_nplanes(const KTab &T) { return KS_TB + T.tb_nlev_tot; }

// The largest XNF/XNFP species index the sources read with these tables.
inline int kapp_max_species(const KPack &P)
{
    int m = 848;
    for (int s = 0; s < P.scal.tb_nsp; ++s) m = std::max(m, P.i[P.tb_idx + s]);
    for (int e = 0; e < P.scal.hi_nedge; ++e) m = std::max(m, (int)P.d[P.hi_edge + 7 * e + 6]);
    for (int k = 0; k < 2; ++k) m = std::max(m, P.scal.felo_idx[k]);
    return m;
}

}  // namespace kapp
