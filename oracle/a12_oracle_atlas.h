/*
 * a12_oracle_atlas.h -- ATLAS12 part of the CPU oracle (test
 * infrastructure only; see a12_oracle.h).
 */
#ifndef A12_ORACLE_ATLAS_H
#define A12_ORACLE_ATLAS_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Inputs of LINOP1 (src/atlas12_modules.f90:18132-18374).  Depth arrays
 * have nrhox entries; (ldion, mion) arrays hold XNFDOP(J, NELION) and
 * DOPPLE(J, NELION) with J fastest (ldion >= nrhox). */
typedef struct {
    int nrhox, numnu, nlines, mion, ldion, nwave1;
    const double *waveset;     /* (numnu) [nm] */
    const int32_t *linedata;   /* (4, nlines) LINEDATA records */
    const double *tabcont;     /* (nrhox, nwave1) TABCONT, J fastest */
    const int32_t *iwavetab;   /* (nwave1) with its 2**30 sentinel */
    const double *xnfdop, *dopple; /* (ldion, mion) */
    const double *t, *xne, *hckt, *xnf1, *xnf3, *xnf841; /* (nrhox) */
    const double *h0, *h1, *h2; /* TABVOIGT tables (2001) */
} or_atlas_lines_in;

void or_tabvoigt(double vsteps, int n, double *h0, double *h1, double *h2);
double or_voigt_kurucz(double v, double a, const double *h0, const double *h1,
                       const double *h2);
void or_linop1(const or_atlas_lines_in *in, double *xlines, int64_t *evals);
/* LINOP1 over lines [l0, l1), adding into xlines (not cleared). */
void or_linop1_range(const or_atlas_lines_in *in, int64_t l0, int64_t l1, double *xlines,
                     int64_t *evals);
/* LINOP1 with the lines split into contiguous chunks over OpenMP threads,
 * each into its own XLINES, summed at the end (a timing port: the sum
 * order differs from the reference's). */
void or_linop1_omp(const or_atlas_lines_in *in, double *xlines, int64_t *evals);

/* Frequency loop of RUN_RT_PASS with tabulated KAPP outputs. */
typedef struct {
    int nrhox, nfreq;
    double teff;
    const double *wave, *rcowt;              /* (nfreq) */
    const double *t, *hkt, *rhox;            /* (nrhox) */
    const double *acont, *sigmac, *scont, *sline, *aline, *sigmal; /* (nrhox, nfreq) */
    const double *xlines;                    /* (nrhox, nfreq) */
    const double *coefj, *coefh;
} or_rt_in;

typedef struct {
    double rjmins[128], rdabh[128], rdiagj[128], flxrad[128];
    double raden[128], h_total[128], accrad[128], abross[128];
    double pradk0;
    /* the last frequency's RDIAGJNU (TCORR(2)'s per-frequency term, :2507-2510),
     * HNU and TAUNU, as the reference leaves them in its module arrays */
    double last_rdiagjnu[128], last_hnu[128], last_taunu[128];
} or_rt_acc;

void or_rt_pass(const or_rt_in *in, or_rt_acc *acc, double *hnu1_out);

/* SELECTLINES' filter (src/atlas12_modules.f90:19171-19781) on the
 * predicted 16-byte records (path 1) followed by the TiO MOL2 records
 * (path 5).  tabcont: (nrhox, nwave + 1), J fastest; xnfdop: (ldion, mion). */
typedef struct {
    int nrhox, mion, ldion, nwave;
    double teff;
    const double *tabcont;
    const int32_t *iwavetab;   /* (nwave + 1), sentinel last */
    const double *wavetab;     /* (nwave) [nm] */
    const double *xnfdop;
    const double *hckt;        /* (nrhox) */
} or_select_in;

/* MOL2 record (32 bytes). */
typedef struct {
    float wl, loggf, e1, e2;
    int32_t code, iso, gr100, flag;
} or_mol2_rec;

/* Writes the kept records (4 x int32 each) to out in store order; returns
 * their number, or -1 for a NELION outside 1..mion in the TiO path. */
int64_t or_atlas_select(const or_select_in *in, const int32_t *pred, int64_t npred,
                        const or_mol2_rec *mol2, int64_t nmol2, int32_t *out);

/* diatomic_record_t (src/mod_mklinelist.f90:71-78), 48 bytes. */
typedef struct {
    double wlvac_nm, gflog_dex, elo_cm, gammar_log, gammaw_log;
    int32_t molcode, pad;
} or_diatomic_rec;

/* h2opokazatel.bin record (src/atlas12_modules.f90:19659-19660), 8 bytes. */
typedef struct {
    int32_t iwl;
    int16_t ielo, igflog;
} or_h2o_rec;

/* The seven lists SELECTLINES reads, in store order; 16-byte records are
 * int32 (n, 4); dia in read_diatomics_for_atlas's order (unsorted). */
typedef struct {
    const int32_t *pred;            int64_t npred;
    const int32_t *lowobs;          int64_t nlowobs;
    const int32_t *hilines;         int64_t nhilines;
    const or_diatomic_rec *dia;     int64_t ndia;
    const or_mol2_rec *mol2;        int64_t nmol2;
    const or_h2o_rec *h2o;          int64_t nh2o;
    const int32_t *h3p;             int64_t nh3p;
} or_select_lists;

/* SELECTLINES (src/atlas12_modules.f90:19171-19781) over all seven lists.
 * Returns the number of kept records written to out, -1 for a NELION
 * outside 1..mion (the observed-low-line and hilines paths index without a
 * bounds check), -2 for a diatomic MOLCODE outside its table. */
int64_t or_atlas_select_all(const or_select_in *in, const or_select_lists *L, int32_t *out);

/* sort_indirect_by_int (src/atlas12_modules.f90:1985-2039): the heapsort
 * whose order among equal keys the diatomic store inherits.  keys and
 * order (n each) are permuted together. */
void or_sort_indirect_by_int(int32_t *keys, int32_t *order, int64_t n);

/* SCALE_MODEL(TEFF_NEW, LOGG_NEW), src/atlas12_modules.f90:6967-7060.
 * in: RHOX, T, P, XNE, ABROSS, PRAD, VTURB, each (n); sc: TEFF, GRAV,
 * PRADK0, PCON, PTURB0.  out (kw each): TAUSTD, RHOX, T, P, XNE, ABROSS,
 * PRAD, VTURB, PRADK, PTOTAL (PTOTAL only when rescaled); so: TEFF, GRAV,
 * GLOG, FLUX, PRADK0, PZERO (FLUX, GLOG, PZERO as given when not rescaled).
 * Returns 1 when rescaled, 0 when only regridded. */
int or_scale_model(int n, int kw, double tau1lg, double steplg, const double *sc,
                   double teff_new, double logg_new, const double *const in[7],
                   double *const out[10], double *so);

/* TCORR(3) sections A-F, src/atlas12_modules.f90:2528-2752: the
 * flux-constancy (DTFLUX), Lambda (DTLAMB) and surface (DTSURF)
 * corrections and the iteration damping.  Arrays (n).  rdabh and oldt1 are
 * TCORR's SAVE state, updated in place.  out (7 x n): T1, DTFLUX, DTLAMB,
 * DTSURF, FLXERR, FLXDRV, HRATIO. */
typedef struct {
    int n, ifconv, iter;
    double flux, teff, mixlth;
    const double *tauros, *rhox, *t, *dltdlp, *abross, *flxcnv, *flxcnv0, *grdadb, *ptotal,
        *rho, *dlrdlt, *heatcp, *hscale, *flxrad, *rjmins, *rdiagj;
} or_tcorr_in;
void or_tcorr3(const or_tcorr_in *in, double *rdabh, double *oldt1, double *out);

/* ROSSTAB's table (module state ROSSTAB_*, src/atlas12_modules.f90:
 * 1779-1785; maxtab = kw * 60).  Arrays of maxtab entries. */
typedef struct {
    int nross, maxtab;
    double zerot, slopet, zerop, slopep;
    double *tabt, *tabp, *ross;
} or_rosstab;
/* ROSSTAB(0, 0, 0), the store (:6003-6042); returns the entries added. */
int or_rosstab_store(or_rosstab *tb, int n, const double *t, const double *p,
                     const double *abross);
/* ROSSTAB(T, P, V), the lookup (:6047-6268) in mode irosstab (1, 2, 3). */
double or_rosstab_lookup(const or_rosstab *tb, int irosstab, double temp, double pressure);
/* TTAUP (:6303-6523); returns the number of depths whose corrector hit
 * its cap (the best iterate is used there). */
int or_ttaup(const or_rosstab *tb, int irosstab, const double *t_in, const double *tau,
             double *abstd, double *ptotal, double *p_out, const double *prad_in,
             const double *pturb_in, double grav, int numtau);

/* CZ_CONSTRUCT's SAVE state (:3054-3069) and the module flags it sets. */
typedef struct {
    int prev_engaged, jc_sticky, nbot_bad, run_engaged, run_jc;
    double last_dtmax, rel_err;
} or_czc_state;
/* The module arrays CZ_CONSTRUCT and CZC_DELREQ_LAYER read (n each). */
typedef struct {
    int n, ifconv, iter, use4pt, use_constructor, irosstab;
    double flux, mixlth, grav;
    const double *flxcnv, *flxrad, *dltdlp, *grdadb, *flxerr, *t, *rhox, *ptotal, *rho,
        *heatcp, *dlrdlt, *p, *vturb, *abross, *hscale;
    const or_rosstab *tab;
} or_czc_in;
/* CZ_CONSTRUCT(T1), :3012-3287; returns 1 when engaged (T1 changed). */
int or_cz_construct(const or_czc_in *in, or_czc_state *st, double *t1);

/* TCORR(3) sections (G2)-(K), :2816-2967.  Read-only inputs (n each) and
 * the module arrays the sections update in place. */
typedef struct {
    int n, ifconv, iter, use4pt, use_constructor, irosstab, j1smooth, j2smooth;
    double flux, mixlth, grav, tau1lg, steplg, pradk0, tfloor, wtjm1, wtj, wtjp1;
    const double *flxcnv, *flxrad, *dltdlp, *grdadb, *flxerr, *ptotal, *rho, *heatcp,
        *dlrdlt, *hscale;
    const or_rosstab *tab;
} or_tcorr_fin_in;
typedef struct {
    double *tauros, *rhox, *t, *p, *xne, *abross, *prad, *vturb, *bmin, *pturb;
    const double *accrad;                 /* remapped by (K) but not kept  */
    double *bhyd;                         /* (n, 6) BHYD, J fastest        */
    double *tk, *hkt, *hckt, *tkev, *tlog, *pradk;  /* out                 */
} or_tcorr_fin_io;
/* t1: TCORR's T1 after section (F) in, after (G2) out; drhox_out (n) or
 * NULL.  Returns the TTAUP depths that did not converge. */
int or_tcorr3_finish(const or_tcorr_fin_in *in, or_tcorr_fin_io *io, or_czc_state *cz,
                     double *t1, double *drhox_out);

/* CONVEC (src/atlas12_modules.f90:10323-10953) given the four EOS
 * evaluations it makes: edens[k], rho4[k] (n) = the module EDENS and RHO
 * after COMPUTE_ONE_POP at T x 1.001, T x 0.999 (P), then P x 1.001,
 * P x 0.999 (T), in that order (:10468-10528).  height: COMPUTE_HEIGHT's
 * HEIGHT; flxrad: the RT pass's FLXRAD. */
typedef struct {
    int n, ifconv, ifcorr, nconv, use4pt, smooth, irosstab;
    double grav, mixlth, flux, overwt;
    const double *ptotal, *rho, *pradk, *tauros, *rhox, *abross, *vturb, *height, *flxrad;
    const double *edens[4], *rho4[4];
    const or_rosstab *tab;
} or_convec_in;
/* In/out: T, TK, HKT, HCKT, TKEV, TLOG, P (the drift of CONVEC's perturb-
 * and-restore), FLXCNV1 (kept unless overshooting); out: the rest. */
typedef struct {
    double *t, *tk, *hkt, *hckt, *tkev, *tlog, *p;
    double *dltdlp, *heatcp, *velsnd, *dlrdlt, *grdadb, *hscale, *vconv, *flxcnv, *flxcnv0,
        *flxcnv1, *abconv, *deltat;
} or_convec_io;
void or_convec(const or_convec_in *in, or_convec_io *io, int mlt_only);

/* NMOLEC's molecule table (READMOL): 1-based as in the reference.  locj
 * (nummol + 1) start of each entry's components in kcomps; kcomps the
 * equation index of each component (nequa + 1 = electron, divides);
 * idequa (nequa) element of each equation (100 = electrons). */
typedef struct {
    int nequa, nummol;
    const int *locj, *kcomps, *idequa;
} or_moltab;
/* The gas-phase Newton solve of one depth, src/atlas12_modules.f90:
 * 9078-9217.  xab (nequa) element abundances (entry 1 unused), equilj
 * (nummol), xn (nequa) start in, solution out.  Returns the iteration
 * count; *status 0 converged, 1 NaN, 2 iteration cap (5000). */
int or_molec_newton(const or_moltab *tb, double xntot, const double *xab, const double *equilj,
                    double *xn, int *status);

#ifdef __cplusplus
}
#endif
#endif
