/*
 * a12_oracle_tcorr.c -- CPU restatement of the end of TCORR(3): the
 * ROSSTAB Rosseland-opacity table, TTAUP's hydrostatic march, the
 * CZ_CONSTRUCT efficient-convection corrector and sections (G2)-(K) of
 * TCORR(3) (RHOX correction, apply, monotonicity, floor, thermodynamic
 * vectors, remap onto TAUSTD).
 *
 * TEST INFRASTRUCTURE ONLY (see a12_oracle.h): the checker for the HIP
 * library's device TCORR(3).  Every function restates the cited reference
 * lines statement by statement, with the reference's evaluation order
 * (amdflang -O3 -ffp-contract=off evaluates X**4 as ((X*X)*X)*X, X**2 as
 * X*X and X**1.0D0 as X; the tests pin these against the reference's
 * own outputs).
 */
#include <float.h>
#include <math.h>
#include <string.h>

#include "a12_oracle.h"
#include "a12_oracle_atlas.h"

#define KW_MAX 128

/* src/atlas12_modules.f90:315-344 */
#define C_PI 3.141592653589793
#define C_FOURPI (4.0 * C_PI)
#define C_KBOL 1.380649e-16
#define C_HPLANCK 6.62607015e-27
#define C_CLIGHT 2.99792458e10
#define C_EV_ERG 1.602176634e-12
#define C_KBOL_EV (C_KBOL / C_EV_ERG)
#define C_SIGMA_SB 5.670374419e-5

/* ---- ROSSTAB, src/atlas12_modules.f90:5937-6268 ----------------------- */

/* MODE 1, STORE (:6003-6042): append the n layers' (T, P, ABROSS); the
 * first store fixes the normalisation from the surface and bottom layers. */
int or_rosstab_store(or_rosstab *tb, int n, const double *t, const double *p,
                     const double *abross)
{
    if (tb->nross == 0) {
        tb->zerot = log10(t[0]);
        tb->zerop = log10(p[0]);
        tb->slopet = log10(t[n - 1]) - tb->zerot;
        tb->slopep = log10(p[n - 1]) - tb->zerop;
    }
    int added = 0;
    for (int j = 0; j < n; ++j) {
        tb->nross = tb->nross + 1;
        if (tb->nross > tb->maxtab) {  /* :6014-6018: overflow, nothing written */
            tb->nross = tb->maxtab;
            break;
        }
        const int k = tb->nross - 1;
        tb->tabt[k] = (log10(t[j]) - tb->zerot) / tb->slopet;
        tb->tabp[k] = (log10(p[j]) - tb->zerop) / tb->slopep;
        tb->ross[k] = log10(abross[j]);
        ++added;
    }
    return added;
}

/* MODE 2, LOOKUP (:6047-6268).  irosstab 1: nearest stored point in each
 * of the four (T, P) quadrants and a bilinear blend (inverse-distance
 * weights when a quadrant is empty); 2: the 12 nearest points with Shepard
 * weights 1/(d^2 + 1e-4); 3: a moving-least-squares plane through them
 * with Wendland weights.  Returns 10**R. */
double or_rosstab_lookup(const or_rosstab *tb, int irosstab, double temp, double pressure)
{
    const double templog = (log10(temp) - tb->zerot) / tb->slopet;
    const double presslog = (log10(pressure) - tb->zerop) / tb->slopep;
    double r;
    if (irosstab == 1) {
        double rmin_pp = 1.0e30, rmin_pm = 1.0e30, rmin_mp = 1.0e30, rmin_mm = 1.0e30;
        int idx_pp = 0, idx_pm = 0, idx_mp = 0, idx_mm = 0;
        for (int i = 1; i <= tb->nross; ++i) {
            const double dt = tb->tabt[i - 1] - templog;
            const double dp = tb->tabp[i - 1] - presslog;
            const double radius2 = dt * dt + dp * dp;
            if (dt >= 0.0) {
                if (dp >= 0.0) {
                    if (radius2 < rmin_pp) { rmin_pp = radius2; idx_pp = i; }
                } else {
                    if (radius2 < rmin_pm) { rmin_pm = radius2; idx_pm = i; }
                }
            } else {
                if (dp >= 0.0) {
                    if (radius2 < rmin_mp) { rmin_mp = radius2; idx_mp = i; }
                } else {
                    if (radius2 < rmin_mm) { rmin_mm = radius2; idx_mm = i; }
                }
            }
        }
        if (idx_pp != 0 && idx_pm != 0 && idx_mp != 0 && idx_mm != 0) {
            const double tpp = tb->tabt[idx_pp - 1], ppp = tb->tabp[idx_pp - 1], rpp = tb->ross[idx_pp - 1];
            const double tpm = tb->tabt[idx_pm - 1], ppm = tb->tabp[idx_pm - 1], rpm = tb->ross[idx_pm - 1];
            const double tmp = tb->tabt[idx_mp - 1], pmp = tb->tabp[idx_mp - 1], rmp = tb->ross[idx_mp - 1];
            const double tmm = tb->tabt[idx_mm - 1], pmm = tb->tabp[idx_mm - 1], rmm = tb->ross[idx_mm - 1];
            const double r_hi = ((templog - tmp) * rpp + (tpp - templog) * rmp) / (tpp - tmp);
            const double p_hi = ((templog - tmp) * ppp + (tpp - templog) * pmp) / (tpp - tmp);
            const double r_lo = ((templog - tmm) * rpm + (tpm - templog) * rmm) / (tpm - tmm);
            const double p_lo = ((templog - tmm) * ppm + (tpm - templog) * pmm) / (tpm - tmm);
            r = ((presslog - p_lo) * r_hi + (p_hi - presslog) * r_lo) / (p_hi - p_lo);
        } else {
            const double d_pp = sqrt(rmin_pp) + 1.0e-5, d_pm = sqrt(rmin_pm) + 1.0e-5;
            const double d_mp = sqrt(rmin_mp) + 1.0e-5, d_mm = sqrt(rmin_mm) + 1.0e-5;
            const double rwt = 1.0 / d_pp + 1.0 / d_pm + 1.0 / d_mp + 1.0 / d_mm;
            /* ROSS(max(1, IDX)): an empty quadrant reads entry 1 */
            r = (tb->ross[(idx_pp > 1 ? idx_pp : 1) - 1] / d_pp +
                 tb->ross[(idx_pm > 1 ? idx_pm : 1) - 1] / d_pm +
                 tb->ross[(idx_mp > 1 ? idx_mp : 1) - 1] / d_mp +
                 tb->ross[(idx_mm > 1 ? idx_mm : 1) - 1] / d_mm) / rwt;
        }
    } else {
        enum { KFIT = 12 };
        double dist2[KFIT];
        int kidx[KFIT];
        int kfound = 0, kmax = 1;
        double dmax2 = 0.0;
        for (int i = 1; i <= tb->nross; ++i) {
            const double dt = tb->tabt[i - 1] - templog;
            const double dp = tb->tabp[i - 1] - presslog;
            const double radius2 = dt * dt + dp * dp;
            if (kfound < KFIT) {
                kfound = kfound + 1;
                dist2[kfound - 1] = radius2;
                kidx[kfound - 1] = i;
                if (radius2 > dmax2) {
                    dmax2 = radius2;
                    kmax = kfound;
                }
            } else if (radius2 < dmax2) {
                dist2[kmax - 1] = radius2;
                kidx[kmax - 1] = i;
                dmax2 = dist2[0];
                kmax = 1;
                for (int k = 2; k <= KFIT; ++k)
                    if (dist2[k - 1] > dmax2) {
                        dmax2 = dist2[k - 1];
                        kmax = k;
                    }
            }
        }
        if (irosstab == 2) {
            double rs = 0.0, rwt = 0.0;
            for (int k = 0; k < kfound; ++k) {
                const double w = 1.0 / (dist2[k] + 1.0e-4);  /* (d^2 + eps)**1.0D0 */
                rs = rs + w * tb->ross[kidx[k] - 1];
                rwt = rwt + w;
            }
            r = rs / rwt;
        } else {
            const double d2sup = dmax2 * 1.1025;
            double a11 = 0.0, a12 = 0.0, a13 = 0.0, a22 = 0.0, a23 = 0.0, a33 = 0.0;
            double b1 = 0.0, b2 = 0.0, b3 = 0.0;
            for (int k = 0; k < kfound; ++k) {
                const double dnrm = sqrt(dist2[k] / d2sup);
                const double om = 1.0 - dnrm;
                const double w = ((om * om) * om) * om * (1.0 + 4.0 * dnrm);
                const double dt = tb->tabt[kidx[k] - 1] - templog;
                const double dp = tb->tabp[kidx[k] - 1] - presslog;
                const double rr = tb->ross[kidx[k] - 1];
                a11 = a11 + w;
                a12 = a12 + w * dt;
                a13 = a13 + w * dp;
                a22 = a22 + w * dt * dt;
                a23 = a23 + w * dt * dp;
                a33 = a33 + w * dp * dp;
                b1 = b1 + w * rr;
                b2 = b2 + w * rr * dt;
                b3 = b3 + w * rr * dp;
            }
            const double ridge = 1.0e-6 * (a22 + a33) + 1.0e-30;
            a22 = a22 + ridge;
            a33 = a33 + ridge;
            const double det = a11 * (a22 * a33 - a23 * a23) - a12 * (a12 * a33 - a23 * a13) +
                               a13 * (a12 * a23 - a22 * a13);
            const double det1 = b1 * (a22 * a33 - a23 * a23) - a12 * (b2 * a33 - a23 * b3) +
                                a13 * (b2 * a23 - a22 * b3);
            const double deta = fmax(fabs(a11 * a22 * a33), 1.0e-300);
            if (fabs(det) > 1.0e-10 * deta && a11 > 0.0) r = det1 / det;
            else r = b1 / fmax(a11, 1.0e-300);
        }
    }
    return pow(10.0, r);
}

/* ---- TTAUP, src/atlas12_modules.f90:6303-6523 ------------------------- */
/* Hydrostatic equilibrium on a log-uniform tau grid: predictor (surface
 * estimate, linear, 4th-order Adams-Bashforth), corrector (surface,
 * trapezoidal, 4th-order Adams-Moulton) with stall-driven relaxation, the
 * best-iterate fallback and the surface bootstrap.  Returns the number of
 * depths whose corrector did not converge. */
int or_ttaup(const or_rosstab *tb, int irosstab, const double *t_in, const double *tau,
             double *abstd, double *ptotal, double *p_out, const double *prad_in,
             const double *pturb_in, double grav, int numtau)
{
    const double relax_ladder[4] = {0.5, 0.3, 0.1, 0.05};
    const double dlgtau = log(tau[1] / tau[0]);
    double plog1 = 0.0, plog2 = 0.0, plog3 = 0.0, plog4 = 0.0;
    double dplog1 = 0.0, dplog2 = 0.0, dplog3 = 0.0;
    double plog = 0.0, pnew = 0.0, dplog = 0.0;
    int nfail = 0;
    abstd[0] = 0.1;
    if (prad_in[0] > 0.0) abstd[0] = fmin(0.1, grav * tau[0] / prad_in[0] / 2.0);
    double abstd1_init = abstd[0];
    for (int J = 1; J <= numtau; ++J) {
        const int j = J - 1;
        if (J == 1) plog = log(grav / abstd[0] * tau[0]);
        else if (J <= 4) plog = plog1 + dplog1;
        else plog = (3.0 * plog4 + 8.0 * dplog1 - 4.0 * dplog2 + 8.0 * dplog3) / 3.0;
        double error = 1.0, prev_error = 1.0, best_error = DBL_MAX, best_plog = plog;
        int n = 0, converged = 0, n_stall = 0, relax_level = 1;
        double relax;
        for (;;) {
            ptotal[j] = exp(plog);
            p_out[j] = ptotal[j] + (prad_in[0] - prad_in[j]) + (pturb_in[0] - pturb_in[j]);
            if (p_out[j] <= 0.0) {
                p_out[j] = ptotal[j] * 1.0e-4;
                if (p_out[j] <= 0.0) p_out[j] = 1.0e-10;
            }
            abstd[j] = or_rosstab_lookup(tb, irosstab, t_in[j], p_out[j]);
            dplog = grav / abstd[j] * tau[j] / ptotal[j] * dlgtau;
            dplog = fmax(0.0, fmin(10.0, dplog));
            n = n + 1;
            if (n == 1) continue;
            if (J == 1) pnew = log(grav / abstd[0] * tau[0]);
            else if (J <= 4) pnew = (plog + 2.0 * plog1 + dplog + dplog1) / 3.0;
            else
                pnew = (126.0 * plog1 - 14.0 * plog3 + 9.0 * plog4 + 42.0 * dplog +
                        108.0 * dplog1 - 54.0 * dplog2 + 24.0 * dplog3) / 121.0;
            prev_error = error;
            error = fabs(pnew - plog);
            if (error < best_error) {
                best_error = error;
                best_plog = pnew;
            }
            if (n > 3 && error > 0.9 * prev_error) {
                n_stall = n_stall + 1;
                if (n_stall >= 4 && relax_level < 4) {
                    relax_level = relax_level + 1;
                    n_stall = 0;
                }
            } else {
                n_stall = n_stall - 1 > 0 ? n_stall - 1 : 0;
            }
            relax = relax_ladder[relax_level - 1];
            plog = (1.0 - relax) * plog + relax * pnew;
            if (error <= 1.0e-4) {
                converged = 1;
                break;
            }
            if (n >= 50) break;
        }
        if (!converged) {
            ++nfail;
            plog = best_plog;
            ptotal[j] = exp(plog);
            p_out[j] = ptotal[j] + (prad_in[0] - prad_in[j]) + (pturb_in[0] - pturb_in[j]);
            if (p_out[j] <= 0.0) p_out[j] = fmax(ptotal[j] * 1.0e-4, 1.0e-10);
            abstd[j] = or_rosstab_lookup(tb, irosstab, t_in[j], p_out[j]);
            dplog = grav / abstd[j] * tau[j] / ptotal[j] * dlgtau;
            dplog = fmax(0.0, fmin(10.0, dplog));
        }
        if (J == 1 && (abstd[0] / abstd1_init > 2.0 || abstd1_init / abstd[0] > 2.0)) {
            abstd1_init = abstd[0];
            plog = log(grav / abstd[0] * tau[0]);
            ptotal[0] = exp(plog);
            p_out[0] = ptotal[0];
            if (p_out[0] > 0.0) {
                abstd[0] = or_rosstab_lookup(tb, irosstab, t_in[0], p_out[0]);
                plog = log(grav / abstd[0] * tau[0]);
                ptotal[0] = exp(plog);
                p_out[0] = ptotal[0];
            }
            dplog = grav / abstd[0] * tau[0] / ptotal[0] * dlgtau;
            dplog = fmax(0.0, fmin(10.0, dplog));
        }
        plog4 = plog3;
        plog3 = plog2;
        plog2 = plog1;
        plog1 = plog;
        dplog3 = dplog2;
        dplog2 = dplog1;
        dplog1 = dplog;
    }
    return nfail;
}

/* ---- CZ_CONSTRUCT, src/atlas12_modules.f90:3012-3287 ------------------ */

/* CZC_DELREQ_LAYER (:3298-3358): the superadiabatic excess that carries
 * F_conv = FLUX - min(FLXRAD, 0.999 FLUX) at layer j (0-based). */
static double czc_delreq_layer(const or_czc_in *in, int j, double delin_j)
{
    const double f_req = in->flux - fmin(in->flxrad[j], 0.999 * in->flux);
    const double fluxco = 0.5 * in->rho[j] * in->heatcp[j] * in->t[j] * in->mixlth / C_FOURPI;
    const double vco =
        0.5 * in->mixlth * sqrt(fmax(-0.5 * in->ptotal[j] / in->rho[j] * in->dlrdlt[j], 0.0));
    if (fluxco <= 0.0 || vco <= 0.0) return delin_j;
    double delta = pow(f_req / (fluxco * vco), 2.0 / 3.0);
    delta = fmin(delta, 0.15 / in->mixlth);
    const double deltat_b = in->t[j] * in->mixlth * delta;
    const double tj = in->t[j], pj = in->p[j];
    const double rosst0 = or_rosstab_lookup(in->tab, in->irosstab, tj, pj);
    const double dplus = or_rosstab_lookup(in->tab, in->irosstab, tj + deltat_b, pj) / rosst0;
    const double dminus = or_rosstab_lookup(in->tab, in->irosstab, tj - deltat_b, pj) / rosst0;
    double abconv;
    if (in->use4pt) {
        const double dplus_i =
            or_rosstab_lookup(in->tab, in->irosstab, tj + deltat_b / 3.0, pj) / rosst0;
        const double dminus_i =
            or_rosstab_lookup(in->tab, in->irosstab, tj - deltat_b / 3.0, pj) / rosst0;
        abconv = 4.0 / (1.0 / dplus + 1.0 / dplus_i + 1.0 / dminus + 1.0 / dminus_i) *
                 in->abross[j];
    } else {
        abconv = 2.0 / (1.0 / dplus + 1.0 / dminus) * in->abross[j];
    }
    const double t4 = ((tj * tj) * tj) * tj;
    double dresp = 8.0 * C_SIGMA_SB * t4 / (abconv * in->hscale[j] * in->rho[j]) /
                   (fluxco * C_FOURPI) / vco;
    const double taub = abconv * in->rho[j] * in->mixlth * in->hscale[j];
    dresp = dresp * (taub * taub) / (2.0 + taub * taub);
    dresp = dresp * dresp / 2.0;
    return delta + sqrt(2.0 * dresp * delta);
}

/* CZ_CONSTRUCT(T1) with its SAVE state in *st; returns 1 when it changed
 * T1 (engaged), 0 otherwise.  CZC_* parameters: :1505-1590. */
int or_cz_construct(const or_czc_in *in, or_czc_state *st, double *t1)
{
    const double hr_min = 0.60, hr_w_lo = 0.50, hr_w_hi = 0.85, del_max = 0.05;
    const double err_on = 5.0, err_off = 2.0, dt_release = 1.5, reeng_fac = 2.0;
    const double dt_max = 25.0;
    const int gap_max = 2, j_min = 4, nrefine = 2;
    const int n = in->n;
    double hrat[KW_MAX], delin[KW_MAX], delreq[KW_MAX], tkur[KW_MAX], tw[KW_MAX];
    double wblnd[KW_MAX], gin[KW_MAX], dtk[KW_MAX], gadj[KW_MAX], dtw[KW_MAX], errg[KW_MAX];
    int goodj[KW_MAX];
#define F1(a, J) (a)[(J) - 1]
    if (in->iter <= 1) {
        st->prev_engaged = 0;
        st->jc_sticky = 0;
        st->nbot_bad = 0;
        st->last_dtmax = 1.0e30;
        st->rel_err = 0.0;
        st->run_engaged = 0;
        st->run_jc = 0;
    }
    if (!in->use_constructor) return 0;
    if (in->ifconv != 1 || in->mixlth <= 0.0) return 0;
    if (n < j_min + 5) return 0;
    for (int J = 1; J <= n; ++J) {
        F1(hrat, J) = F1(in->flxcnv, J) / (F1(in->flxcnv, J) + fmax(F1(in->flxrad, J), 1.0e-30));
        F1(delin, J) = F1(in->dltdlp, J) - F1(in->grdadb, J);
        F1(goodj, J) = (F1(hrat, J) > hr_min) && (F1(delin, J) < del_max);
    }
    if (!F1(goodj, n)) {
        if (st->prev_engaged && st->nbot_bad < 3) {
            st->nbot_bad = st->nbot_bad + 1;
            F1(goodj, n) = 1;
        } else {
            st->prev_engaged = 0;
            st->jc_sticky = 0;
            st->nbot_bad = 0;
            return 0;
        }
    } else {
        st->nbot_bad = 0;
    }
    int jc = n, ngap = 0;
    for (int J = n - 1; J >= j_min; --J) {
        if (F1(goodj, J)) {
            jc = J;
            ngap = 0;
        } else {
            ngap = ngap + 1;
            if (ngap > gap_max) break;
        }
    }
    const int jc_found = jc;
    if (st->jc_sticky > 0) {
        const int s = st->jc_sticky > j_min ? st->jc_sticky : j_min;
        jc = jc < s ? jc : s;
    }
    for (int J = jc; J <= jc_found - 1; ++J)
        if (F1(delin, J) > del_max) jc = J + 1;
    if (n - jc < 4) {
        st->prev_engaged = 0;
        st->jc_sticky = 0;
        return 0;
    }
    double maxerr = 0.0;
    for (int J = jc; J <= n; ++J) maxerr = fmax(maxerr, fabs(F1(in->flxerr, J)));
    int engage;
    if (st->prev_engaged) {
        engage = maxerr > err_off;
        if (st->last_dtmax < dt_release) {
            engage = 0;
            st->rel_err = maxerr;
        }
    } else {
        engage = maxerr > fmax(err_on, reeng_fac * st->rel_err);
    }
    st->prev_engaged = engage;
    if (!engage) {
        st->jc_sticky = 0;
        st->last_dtmax = 1.0e30;
        return 0;
    }
    st->jc_sticky = jc;
    st->rel_err = 0.0;
    st->run_engaged = 1;
    st->run_jc = jc;
    for (int J = jc; J <= n; ++J) F1(delreq, J) = czc_delreq_layer(in, J - 1, F1(delin, J));
    for (int J = 1; J <= n; ++J) {
        F1(tkur, J) = F1(in->t, J) + F1(t1, J);
        F1(tw, J) = F1(tkur, J);
        F1(gadj, J) = 0.0;
    }
    for (int J = (jc - 1 > 1 ? jc - 1 : 1); J <= n; ++J) {
        double wint = (F1(hrat, J) - hr_w_lo) / (hr_w_hi - hr_w_lo);
        wint = fmax(0.0, fmin(1.0, wint));
        F1(wblnd, J) = wint * wint * (3.0 - 2.0 * wint);
    }
    or_deriv(in->rhox, tkur, dtk, n);
    for (int J = jc; J <= n - 1; ++J)
        F1(gin, J) = F1(in->ptotal, J) / F1(tkur, J) / in->grav * F1(dtk, J);
    F1(gin, n) = log(F1(tkur, n) / F1(tkur, n - 1)) / log(F1(in->ptotal, n) / F1(in->ptotal, n - 1));
    for (int K = 0; K <= nrefine; ++K) {
        for (int J = jc; J <= n; ++J) {
            const double lnpr = log(F1(in->ptotal, J) / F1(in->ptotal, J - 1));
            const double slopin = log(F1(tkur, J) / F1(tkur, J - 1)) / lnpr;
            double gtarg;
            if (J == jc)
                gtarg = 0.5 * (F1(in->grdadb, J - 1) + F1(in->grdadb, J)) + F1(delreq, J);
            else
                gtarg = 0.5 * (F1(in->grdadb, J - 1) + F1(in->grdadb, J)) +
                        0.5 * (F1(delreq, J - 1) + F1(delreq, J));
            const double wint = 0.5 * (F1(wblnd, J - 1) + F1(wblnd, J));
            const double gbari = (gtarg - slopin) * wint + F1(gadj, J);
            F1(tw, J) = F1(tw, J - 1) * (F1(tkur, J) / F1(tkur, J - 1)) * exp(gbari * lnpr);
            F1(tw, J) = fmax(F1(tw, J), F1(tw, J - 1));
        }
        if (K == nrefine) break;
        or_deriv(in->rhox, tw, dtw, n);
        for (int J = jc; J <= n - 1; ++J)
            F1(errg, J) = F1(wblnd, J) * (F1(in->grdadb, J) + F1(delreq, J)) +
                          (1.0 - F1(wblnd, J)) * F1(gin, J) -
                          F1(in->ptotal, J) / F1(tw, J) / in->grav * F1(dtw, J);
        F1(errg, n) = F1(wblnd, n) * (F1(in->grdadb, n) + F1(delreq, n)) +
                      (1.0 - F1(wblnd, n)) * F1(gin, n) -
                      log(F1(tw, n) / F1(tw, n - 1)) / log(F1(in->ptotal, n) / F1(in->ptotal, n - 1));
        F1(gadj, jc) = F1(gadj, jc) + 0.70 * F1(errg, jc);
        for (int J = jc + 1; J <= n; ++J)
            F1(gadj, J) = F1(gadj, J) + 0.35 * (F1(errg, J - 1) + F1(errg, J));
    }
    double dtmax_blk = 0.0;
    for (int J = jc; J <= n; ++J) dtmax_blk = fmax(dtmax_blk, fabs(F1(tw, J) - F1(tkur, J)));
    double scl = 1.0;
    if (dtmax_blk > dt_max) scl = dt_max / dtmax_blk;
    for (int J = jc; J <= n; ++J) F1(t1, J) = F1(t1, J) + (F1(tw, J) - F1(tkur, J)) * scl;
    st->last_dtmax = dtmax_blk * scl;
#undef F1
    return 1;
}

/* ---- TCORR(3) sections (G2)-(K), src/atlas12_modules.f90:2816-2967 ----- */
/* (G2) CZ_CONSTRUCT; (H) the RHOX correction from two TTAUP marches on
 * TAUSTD (T and T + T1); (I) T += T1, the optional 3-point smoothing,
 * monotonicity, the temperature floor; (J) TK, HKT, HCKT, TKEV, TLOG; (K)
 * RHOX += DRHOX and the MAP1 remap of the structure and the hydrogen
 * departure coefficients onto TAUSTD.  The Newton branch (G) is off in the
 * reference (USE_DTGRAD = .FALSE., :1457) and CZC_POLISH re-runs the EOS:
 * both stay with the driver.  Returns the number of TTAUP depths that did
 * not converge. */
int or_tcorr3_finish(const or_tcorr_fin_in *in, or_tcorr_fin_io *io, or_czc_state *cz,
                     double *t1, double *drhox_out)
{
    const int n = in->n;
    double taustd[KW_MAX], tplus[KW_MAX], tnew1[KW_MAX], prdnew[KW_MAX], tnew2[KW_MAX];
    double ab1[KW_MAX], ptot1[KW_MAX], p1[KW_MAX], ab2[KW_MAX], ptot2[KW_MAX], p2[KW_MAX];
    double ppp[KW_MAX], rrr[KW_MAX], drhox[KW_MAX], remap[10][KW_MAX], tsmooth[KW_MAX];
    int nfail = 0;
    /* (G2), on the module arrays as TCORR leaves them here */
    or_czc_in ci;
    ci.n = n;
    ci.ifconv = in->ifconv;
    ci.iter = in->iter;
    ci.use4pt = in->use4pt;
    ci.use_constructor = in->use_constructor;
    ci.irosstab = in->irosstab;
    ci.flux = in->flux;
    ci.mixlth = in->mixlth;
    ci.grav = in->grav;
    ci.flxcnv = in->flxcnv;
    ci.flxrad = in->flxrad;
    ci.dltdlp = in->dltdlp;
    ci.grdadb = in->grdadb;
    ci.flxerr = in->flxerr;
    ci.t = io->t;
    ci.rhox = io->rhox;
    ci.ptotal = in->ptotal;
    ci.rho = in->rho;
    ci.heatcp = in->heatcp;
    ci.dlrdlt = in->dlrdlt;
    ci.p = io->p;
    ci.vturb = io->vturb;
    ci.abross = io->abross;
    ci.hscale = in->hscale;
    ci.tab = in->tab;
    (void)or_cz_construct(&ci, cz, t1);
    /* (H) */
    for (int j = 0; j < n; ++j) {
        tplus[j] = io->t[j] + t1[j];
        taustd[j] = pow(10.0, in->tau1lg + (double)j * in->steplg);
    }
    (void)or_map1(io->tauros, io->t, n, taustd, tnew1, n);
    (void)or_map1(io->tauros, io->prad, n, taustd, prdnew, n);
    nfail += or_ttaup(in->tab, in->irosstab, tnew1, taustd, ab1, ptot1, p1, prdnew, io->pturb,
                      in->grav, n);
    (void)or_map1(io->tauros, tplus, n, taustd, tnew2, n);
    nfail += or_ttaup(in->tab, in->irosstab, tnew2, taustd, ab2, ptot2, p2, prdnew, io->pturb,
                      in->grav, n);
    for (int j = 0; j < n; ++j) ppp[j] = (ptot2[j] - ptot1[j]) / ptot1[j];
    (void)or_map1(taustd, ppp, n, io->tauros, rrr, n);
    for (int j = 0; j < n; ++j) drhox[j] = rrr[j] * io->rhox[j];
    if (drhox_out) memcpy(drhox_out, drhox, sizeof(double) * (size_t)n);
    /* (I) */
    for (int j = 0; j < n; ++j) io->t[j] = io->t[j] + t1[j];
    if (in->j1smooth > 0) {
        for (int J = in->j1smooth; J <= in->j2smooth; ++J)
            tsmooth[J - 1] = in->wtjm1 * io->t[J - 2] + in->wtj * io->t[J - 1] +
                             in->wtjp1 * io->t[J];
        for (int J = in->j1smooth; J <= in->j2smooth; ++J) io->t[J - 1] = tsmooth[J - 1];
    }
    for (int I = 2; I <= n; ++I) {
        const int J = n + 1 - I;
        io->t[J - 1] = fmin(io->t[J - 1], io->t[J] - 1.0);
    }
    for (int j = 0; j < n; ++j) io->t[j] = fmax(io->t[j], in->tfloor);
    /* (J) -- recomputed below after the remap; the values between are not
     * visible outside TCORR */
    /* (K) */
    for (int j = 0; j < n; ++j) io->rhox[j] = io->rhox[j] + drhox[j];
    const double *src[10] = {io->rhox, io->t, io->p, io->xne, io->abross, io->prad, io->vturb,
                             io->bmin, io->pturb, io->accrad};
    for (int k = 0; k < 10; ++k) (void)or_map1(io->tauros, src[k], n, taustd, remap[k], n);
    for (int j = 0; j < n; ++j)
        if (io->tauros[0] > taustd[j] && remap[4][j] < 0.0) remap[4][j] = io->abross[0];
    for (int j = 0; j < n; ++j) {
        io->rhox[j] = remap[0][j];
        io->t[j] = remap[1][j];
        io->tk[j] = C_KBOL * io->t[j];
        io->hkt[j] = C_HPLANCK / io->tk[j];
        io->hckt[j] = io->hkt[j] * C_CLIGHT;
        io->tkev[j] = C_KBOL_EV * io->t[j];
        io->tlog[j] = log(io->t[j]);
        io->p[j] = remap[2][j];
        io->xne[j] = remap[3][j];
        io->abross[j] = remap[4][j];
        io->prad[j] = remap[5][j];
        io->pradk[j] = io->prad[j] + in->pradk0;
        io->vturb[j] = remap[6][j];
        io->bmin[j] = remap[7][j];
        io->pturb[j] = remap[8][j];
    }
    for (int i = 0; i < 6; ++i) {
        (void)or_map1(io->tauros, io->bhyd + (size_t)i * n, n, taustd, remap[0], n);
        memcpy(io->bhyd + (size_t)i * n, remap[0], sizeof(double) * (size_t)n);
    }
    memcpy(io->tauros, taustd, sizeof(double) * (size_t)n);
    return nfail;
}

/* ---- CONVEC, src/atlas12_modules.f90:10323-10953 ---------------------- */
/* The mixing-length convection of one iteration given the four EOS
 * evaluations CONVEC makes (EDENS and RHO at T x 1.001, T x 0.999,
 * P x 1.001, P x 0.999; the EOS itself stays with the driver): the
 * finite-difference thermodynamic derivatives, the per-depth MLT solve
 * with its bubble-opacity iteration (Picard, Steffensen every third step,
 * cycle-mean exit), the radiative-gap fill, the bottom clamp, the 1-2-1
 * smoothing, overshooting and the surface zeroing.  T, TK, HKT, HCKT,
 * TKEV, TLOG and P come back as the reference leaves them after its
 * perturb-and-restore sequence (a few ulps from the input). */
void or_convec(const or_convec_in *in, or_convec_io *io, int mlt_only)
{
    const int n = in->n;
    const double fourpi = C_FOURPI, sigma_sb = C_SIGMA_SB;
    double dtdrhx[KW_MAX], dilut[KW_MAX], e1[KW_MAX], e2[KW_MAX], e3[KW_MAX], e4[KW_MAX];
    double cnvint[KW_MAX], delhgt[KW_MAX];
    or_deriv(in->rhox, io->t, dtdrhx, n);                      /* :10452 */
    for (int j = 0; j < n; ++j) dilut[j] = 1.0 - exp(-in->tauros[j]);
    const double c4p = ((1.001 * 1.001) * 1.001) * 1.001 - 1.0;
    const double c4m = ((0.999 * 0.999) * 0.999) * 0.999 - 1.0;
    for (int j = 0; j < n; ++j) {                              /* :10468-10535 */
        double tl = io->tlog[j] + 0.0009995003, t = io->t[j] * 1.001, tk = io->tk[j] * 1.001;
        double hkt = io->hkt[j] / 1.001, hckt = io->hckt[j] / 1.001, tkev = io->tkev[j] * 1.001;
        e1[j] = in->edens[0][j] + 3.0 * in->pradk[j] / in->rho4[0][j] * (1.0 + dilut[j] * c4p);
        tl = tl - 0.0009995003 - 0.0010005003;
        t = t / 1.001 * 0.999;
        tk = tk / 1.001 * 0.999;
        hkt = hkt * 1.001 / 0.999;
        hckt = hckt * 1.001 / 0.999;
        tkev = tkev / 1.001 * 0.999;
        e2[j] = in->edens[1][j] + 3.0 * in->pradk[j] / in->rho4[1][j] * (1.0 + dilut[j] * c4m);
        tl = tl + 0.0010005003;
        t = t / 0.999;
        tk = tk / 0.999;
        hkt = hkt * 0.999;
        hckt = hckt * 0.999;
        tkev = tkev / 0.999;
        double p = io->p[j] * 1.001;
        e3[j] = in->edens[2][j] + 3.0 * in->pradk[j] / in->rho4[2][j];
        p = p / 1.001 * 0.999;
        e4[j] = in->edens[3][j] + 3.0 * in->pradk[j] / in->rho4[3][j];
        p = p / 0.999;
        io->t[j] = t;
        io->tk[j] = tk;
        io->hkt[j] = hkt;
        io->hckt[j] = hckt;
        io->tkev[j] = tkev;
        io->tlog[j] = tl;
        io->p[j] = p;
    }
    const double *t = io->t, *p = io->p, *rho = in->rho, *ptotal = in->ptotal;
    for (int J = 1; J <= n; ++J) {                             /* :10540-10798 */
        const int j = J - 1;
        const double dedt = (e1[j] - e2[j]) / t[j] * 500.0;
        const double drdt = (in->rho4[0][j] - in->rho4[1][j]) / t[j] * 500.0;
        const double dedpg = (e3[j] - e4[j]) / p[j] * 500.0;
        const double drdpg = (in->rho4[2][j] - in->rho4[3][j]) / p[j] * 500.0;
        const double dpdpg = 1.0;
        const double dpdt = 4.0 * in->pradk[j] / t[j] * dilut[j];
        io->dltdlp[j] = ptotal[j] / t[j] / in->grav * dtdrhx[j];
        if (J == n && J >= 2)
            if (t[j] > t[j - 1] && ptotal[j] > ptotal[j - 1])
                io->dltdlp[j] = log(t[j] / t[j - 1]) / log(ptotal[j] / ptotal[j - 1]);
        const double heatcv = dedt - dedpg * drdt / drdpg;
        io->heatcp[j] = dedt - dedpg * dpdt / dpdpg -
                        ptotal[j] / (rho[j] * rho[j]) * (drdt - drdpg * dpdt / dpdpg);
        if (heatcv <= 0.0) io->velsnd[j] = 0.0;
        else io->velsnd[j] = sqrt(fmax(io->heatcp[j] / heatcv * dpdpg / drdpg, 0.0));
        io->dlrdlt[j] = t[j] / rho[j] * (drdt - drdpg * dpdt / dpdpg);
        io->grdadb[j] = -ptotal[j] / rho[j] / t[j] * io->dlrdlt[j] / io->heatcp[j];
        io->hscale[j] = ptotal[j] / rho[j] / in->grav;
        io->vconv[j] = 0.0;
        io->flxcnv[j] = 0.0;
        io->abconv[j] = in->abross[j];
        io->deltat[j] = 0.0;
        if (in->mixlth == 0.0) continue;
        if (J < 4) continue;
        const double del = io->dltdlp[j] - io->grdadb[j];
        if (del < 0.0) continue;
        const double vco = 0.5 * in->mixlth * sqrt(fmax(-0.5 * ptotal[j] / rho[j] * io->dlrdlt[j], 0.0));
        if (vco == 0.0) continue;
        const double fluxco = 0.5 * rho[j] * io->heatcp[j] * t[j] * in->mixlth / fourpi;
        const double rosst = or_rosstab_lookup(in->tab, in->irosstab, t[j], p[j]);
        double olddelt = 0.0;
        const int its30 = in->ifconv == 0 ? 1 : 30;
        const double omega = 0.7;
        double sa = 0.0, sb = 0.0, sc = 0.0, hist[12];
        int shave = 0, hist_n = 0;
        for (int id = 1; id <= its30; ++id) {
            double dt_j = io->deltat[j];
            const double dplus = or_rosstab_lookup(in->tab, in->irosstab, t[j] + dt_j, p[j]) / rosst;
            const double dminus = or_rosstab_lookup(in->tab, in->irosstab, t[j] - dt_j, p[j]) / rosst;
            if (in->use4pt) {
                const double dpi = or_rosstab_lookup(in->tab, in->irosstab, t[j] + dt_j / 3.0, p[j]) / rosst;
                const double dmi = or_rosstab_lookup(in->tab, in->irosstab, t[j] - dt_j / 3.0, p[j]) / rosst;
                io->abconv[j] = 4.0 / (1.0 / dplus + 1.0 / dpi + 1.0 / dminus + 1.0 / dmi) * in->abross[j];
            } else {
                io->abconv[j] = 2.0 / (1.0 / dplus + 1.0 / dminus) * in->abross[j];
            }
            const double t4 = ((t[j] * t[j]) * t[j]) * t[j];
            double d = 8.0 * sigma_sb * t4 / (io->abconv[j] * io->hscale[j] * rho[j]) /
                       (fluxco * fourpi) / vco;
            const double taub = io->abconv[j] * rho[j] * in->mixlth * io->hscale[j];
            d = d * (taub * taub) / (2.0 + taub * taub);
            d = d * d / 2.0;
            const double q = del / (d + del);
            const double ddd = q * q;
            double delta;
            if (ddd >= 0.5) {
                delta = (1.0 - sqrt(1.0 - ddd)) / ddd;
            } else {
                delta = 0.5;
                double term = 0.5, up = -1.0, down = 2.0;
                for (;;) {
                    up = up + 2.0;
                    down = down + 2.0;
                    term = up / down * ddd * term;
                    delta = delta + term;
                    if (term <= 1.0e-6) break;
                }
            }
            delta = delta * (del * del) / (d + del);
            io->vconv[j] = vco * sqrt(delta);
            io->flxcnv[j] = fluxco * io->vconv[j] * delta;
            io->flxcnv[j] = fmax(io->flxcnv[j], 0.0);
            dt_j = t[j] * in->mixlth * delta;
            dt_j = fmin(dt_j, t[j] * 0.15);
            dt_j = dt_j * omega + olddelt * (1.0 - omega);
            io->deltat[j] = dt_j;
            const double dstep = dt_j - olddelt;
            if (fabs(dstep) < 0.5 || fabs(dstep) < 1.0e-3 * fmax(fabs(dt_j), 1.0)) break;
            sa = sb;
            sb = sc;
            sc = dt_j;
            shave = shave + 1 < 3 ? shave + 1 : 3;
            if (shave == 3 && id % 3 == 0) {
                const double den = sc - 2.0 * sb + sa;
                if (fabs(den) > 1.0e-6 * fmax(fabs(sc), 1.0)) {
                    const double num = sc - sb;
                    const double x = sc - num * num / den;
                    if (x > 0.0 && fabs(x - sc) < 0.5 * fmax(fabs(sc), 1.0)) {
                        io->deltat[j] = x;
                        olddelt = x;
                        shave = 0;
                        continue;
                    }
                }
            }
            for (int ii = (hist_n < 11 ? hist_n : 11); ii >= 1; --ii) hist[ii] = hist[ii - 1];
            hist[0] = dt_j;
            hist_n = hist_n + 1 < 12 ? hist_n + 1 : 12;
            int cycle_exit = 0;
            double cmean = 0.0;
            if (hist_n >= 8) {
                for (int ii = 2; ii <= 6; ++ii) {
                    if (hist_n < 2 * ii) continue;
                    int all = 1;
                    const double tol = 5.0e-4 * fmax(fabs(hist[0]), 1.0);
                    for (int k = 0; k < ii; ++k)
                        if (!(fabs(hist[k] - hist[ii + k]) <= tol)) all = 0;
                    if (all) {
                        double cmin = hist[0], cmax = hist[0], s = 0.0;
                        for (int k = 0; k < ii; ++k) {
                            cmin = fmin(cmin, hist[k]);
                            cmax = fmax(cmax, hist[k]);
                            s = s + hist[k];
                        }
                        cmean = s / (double)ii;
                        if (cmax - cmin > 2.0e-3 * fmax(fabs(cmean), 1.0)) {
                            cycle_exit = 1;
                            break;
                        }
                    }
                }
            }
            if (cycle_exit) {
                io->deltat[j] = cmean;
                const double dp2 = or_rosstab_lookup(in->tab, in->irosstab, t[j] + cmean, p[j]) / rosst;
                const double dm2 = or_rosstab_lookup(in->tab, in->irosstab, t[j] - cmean, p[j]) / rosst;
                if (in->use4pt) {
                    const double dpi = or_rosstab_lookup(in->tab, in->irosstab, t[j] + cmean / 3.0, p[j]) / rosst;
                    const double dmi = or_rosstab_lookup(in->tab, in->irosstab, t[j] - cmean / 3.0, p[j]) / rosst;
                    io->abconv[j] = 4.0 / (1.0 / dp2 + 1.0 / dpi + 1.0 / dm2 + 1.0 / dmi) * in->abross[j];
                } else {
                    io->abconv[j] = 2.0 / (1.0 / dp2 + 1.0 / dm2) * in->abross[j];
                }
                break;
            }
            olddelt = dt_j;
        }
    }
    if (mlt_only) return;
    double *f = io->flxcnv;
    int jtop = 0, jbot = 0;                                    /* gap fill :10813-10841 */
    for (int J = 1; J <= n; ++J)
        if (f[J - 1] > 0.0) {
            if (jtop == 0) jtop = J;
            jbot = J;
        }
    if (jtop > 0 && jbot > jtop + 1)
        for (int J = jtop + 1; J <= jbot - 1; ++J)
            if (f[J - 1] == 0.0) {
                int ja = J - 1, jb = J + 1;
                while (ja > jtop && f[ja - 1] == 0.0) ja = ja - 1;
                while (jb < jbot && f[jb - 1] == 0.0) jb = jb + 1;
                if (f[ja - 1] > 0.0 && f[jb - 1] > 0.0) {
                    const double w = (double)(J - ja) / (double)(jb - ja);
                    f[J - 1] = f[ja - 1] * (1.0 - w) + f[jb - 1] * w;
                }
            }
    if (in->ifcorr == 1) {                                     /* bottom clamp :10867-10878 */
        jbot = 0;
        for (int J = 1; J <= n; ++J)
            if (f[J - 1] > 0.0) jbot = J;
        if (jbot >= n - 3 && jbot < n)
            for (int J = jbot + 1; J <= n; ++J) {
                if (io->dltdlp[J - 1] - io->grdadb[J - 1] > -0.01)
                    f[J - 1] = fmax(in->flux - in->flxrad[J - 1], 0.0);
                else
                    break;
            }
    }
    for (int j = 0; j < n; ++j) io->flxcnv0[j] = f[j];        /* smoothing :10885-10893 */
    if (in->smooth)
        for (int J = 2; J <= n - 1; ++J)
            f[J - 1] = 0.25 * io->flxcnv0[J - 2] + 0.50 * io->flxcnv0[J - 1] + 0.25 * io->flxcnv0[J];
    for (int j = 0; j < n; ++j) io->flxcnv0[j] = f[j];
    if (in->overwt > 0.0) {                                    /* overshooting :10901-10927 */
        double wtcnv = 0.0;
        for (int j = 0; j < n; ++j) wtcnv = fmax(wtcnv, f[j] / in->flux);
        wtcnv = fmin(wtcnv, 1.0) * in->overwt;
        for (int j = 0; j < n; ++j)
            delhgt[j] = fmin(fmin(io->hscale[j] * 0.5e-5 * wtcnv, in->height[n - 1] - in->height[j]),
                             in->height[j] - in->height[0]);
        for (int j = 0; j < n; ++j) {
            io->flxcnv0[j] = f[j];
            io->flxcnv1[j] = 0.0;
        }
        or_integ(in->height, f, cnvint, n, 0.0);
        for (int J = n / 2; J <= n - 1; ++J) {
            if (delhgt[J - 1] == 0.0) continue;
            double x1 = in->height[J - 1] - delhgt[J - 1], x2 = in->height[J - 1] + delhgt[J - 1];
            double c1, c2;
            (void)or_map1(in->height, cnvint, n, &x1, &c1, 1);
            (void)or_map1(in->height, cnvint, n, &x2, &c2, 1);
            io->flxcnv1[J - 1] = io->flxcnv1[J - 1] + (c2 - c1) / delhgt[J - 1] / 2.0;
        }
        for (int j = 0; j < n; ++j) f[j] = fmax(io->flxcnv0[j], io->flxcnv1[j]);
    }
    for (int J = 1; J <= in->nconv && J <= n; ++J) f[J - 1] = 0.0;   /* :10930-10932 */
}
