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
} or_rt_acc;

void or_rt_pass(const or_rt_in *in, or_rt_acc *acc, double *hnu1_out);

#ifdef __cplusplus
}
#endif
#endif
