/*
 * a12_oracle.h -- CPU restatement of the ATLAS12/SYNTHE hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * library in atlas12_amd/: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product path never calls
 * it and never falls back to it.
 *
 * Every routine restates one reference routine (cited file:line, relative
 * to the cconroy20/atlas12 tree).  Arrays follow the reference's Fortran
 * layout (column-major, 0-based here).  Two precision modes exist for the
 * line-opacity build:
 *   OR_PREC_REF : the reference's own mixed FP32/FP64 expression types,
 *                 naive FP32 accumulation in line order (src/synthe.f90).
 *   OR_PREC_F64 : the same algorithm with every quantity in FP64; this is
 *                 the "FP64 oracle" errors are stated against.
 *
 * Pinning status of each routine is recorded beside it (see tests/).
 */
#ifndef A12_ORACLE_H
#define A12_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* lte_line_t, src/mod_mklinelist.f90:30-41 (32 bytes). */
typedef struct {
    int32_t nbuff;
    float   cgf;
    int32_t nelion;
    float   elo;
    float   gamrf, gamsf, gamwf;
    float   code;
} or_line_t;

enum { OR_PREC_REF = 0, OR_PREC_F64 = 1 };
enum { OR_NXTAU = 51, OR_MAXPROF = 1000000 };

/* ---- numerics (src/atlas12_modules.f90) ------------------------------ */
void   or_deriv(const double *x, const double *f, double *dfdx, int n);
void   or_parcoe(const double *f, const double *x, double *a, double *b,
                 double *c, int n);
void   or_integ(const double *x, const double *f, double *fint, int n,
                double start);
int    or_map1(const double *xold, const double *fold, int nold,
               const double *xnew, double *fnew, int nnew);
/* returns the number of pivots that hit the 1e-30 floor */
int    or_solvit(double *a, int n, double *b);
double or_expi(int n, double x);

/* ---- Voigt -------------------------------------------------------------- */
float  or_voigt_ref(float v, float a);          /* FP64 inside, FP32 out */
double or_voigt_f64(double v, double a);         /* FP64 in and out     */

/* ---- JOSH ------------------------------------------------------------- */
typedef struct {
    int nrhox;
    /* inputs (nrhox each) */
    const double *acont, *aline, *sigmac, *sigmal, *scont, *sline, *rhox, *bnu;
    const double *coefj, *coefh;       /* 51x51 column-major */
    /* outputs (nrhox each); any may be NULL when ifsurf == 1 */
    double *taunu, *snu, *jnu, *hnu, *jmins;
    double knu1;                       /* K(1) (ifsurf 0) */
    int    maxj;                       /* split index used */
    int    vef_sweeps;                 /* deep-branch sweeps taken (0 if none) */
    int    vef_unconverged;            /* 1 if the deep branch hit its cap */
    int    pivot_floor_hits;
} or_josh_io;
/* src/atlas12_modules.f90:18398-18776.  Returns HNU(1). */
double or_josh(or_josh_io *io, int ifscat, int ifsurf);

/* ---- SYNTHE line opacity (src/synthe.f90:476-689) --------------------- */
typedef struct {
    int     nrhox;         /* layers */
    int     length;        /* wavelength points */
    double  wbegin;        /* first grid wavelength [nm] */
    double  ratio;         /* 1 + 1/resolu */
    double  resolu;
    int     nedge;         /* continuum edges */
    const double *wledge;  /* [nedge] |wledge_m| */
    int     nf;            /* leading dimension of continall (>= 3*(nedge-1)) */
    const double *continall; /* (nf, nrhox) log10 kappa_tot at edge freqs */
    int     mw6;           /* 6*mw */
    const double *xnfpel;  /* (mw6, nrhox) xnfpel_m flattened */
    const double *dopple;  /* (mw6, nrhox) dopple_m flattened */
    const double *rho, *t, *xne, *xnfh, *xnfhe1, *xnfh2, *hckt; /* xf_* [nrhox] */
} or_synthe_atm;

/* Continuum at every grid point of one depth, FP32 as the reference
 * (src/synthe.f90:481-516).  out_f64 (optional) gets the unrounded FP64
 * value 10**x for the FP64 oracle. */
void or_synthe_continuum(const or_synthe_atm *atm, int j, float *out,
                         double *out_f64);

/* Opacity matrix (length, nrhox) lambda-fastest.  prec = OR_PREC_REF writes
 * out32; OR_PREC_F64 writes out64.  evals (optional) receives the number of
 * accumulated profile contributions; nstep_out (optional, nlines*nrhox,
 * line-fastest) receives the profile extent per (line, depth): -1 when the
 * line is cut, else nstep (0 means centre only). */
void or_synthe_line_opacity(const or_synthe_atm *atm, const or_line_t *lines,
                            int nlines, int prec, float *out32, double *out64,
                            int64_t *evals, int32_t *nstep_out);

/* Depths [j0, j1) only (outputs indexed by absolute depth). */
void or_synthe_line_opacity_range(const or_synthe_atm *atm, const or_line_t *lines,
                                  int nlines, int prec, int j0, int j1, float *out32,
                                  double *out64, int64_t *evals, int32_t *nstep_out);
/* OpenMP over depths (one whole depth per thread): identical results. */
void or_synthe_line_opacity_omp(const or_synthe_atm *atm, const or_line_t *lines,
                                int nlines, int prec, int j0, int j1, float *out32,
                                double *out64, int64_t *evals);
int or_omp_threads(int n);
struct or_synthe_rt_s;
void or_synthe_transfer_omp(const struct or_synthe_rt_s *rt, const float *opac,
                            const double *opac64, double *h_line, double *h_cont,
                            int i0, int i1);

/* ---- SYNTHE transfer (src/synthe.f90:698-718, :797-902) --------------- */
typedef struct or_synthe_rt_s {
    int     nrhox, length;
    double  wbegin, ratio;
    int     nedge;
    const double *wledge;        /* [nedge] */
    int     nf;
    const double *contabs;       /* (nf, nrhox) log10 absorption */
    const double *contscat;      /* (nf, nrhox) log10 scattering */
    const double *hkt_syn;       /* synthe_module HKT = xf_hkt [nrhox] */
    const double *hkt_atl;       /* mod_atlas_data HKT [nrhox] */
    const double *rhox;          /* mod_atlas_data RHOX [nrhox] */
    const double *coefj, *coefh; /* 51x51 column-major */
} or_synthe_rt;

/* One wavelength point; opac_col holds opacity(iwave, 1:nrhox) in FP32 (or
 * FP64 when opac_col64 != NULL).  Writes H_line and H_cont as printed in
 * the .spec file. */
void or_synthe_transfer_point(const or_synthe_rt *rt, int iwave,
                              const float *opac_col, const double *opac_col64,
                              double *h_line, double *h_cont);
/* TAUNU (i1 - i0, nrhox) of the line-included JOSH call at wavelength points
 * [i0, i1) (src/synthe.f90:886-890, the .linform table). */
void or_synthe_taunu(const or_synthe_rt *rt, const float *opac, const double *opac64,
                     double *taunu, int i0, int i1);
void or_synthe_transfer(const or_synthe_rt *rt, const float *opac,
                        const double *opac64, double *h_line, double *h_cont,
                        int i0, int i1);

#ifdef __cplusplus
}
#endif
#endif
