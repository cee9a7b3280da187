/*
 * a12gpu.h -- C-ABI of the MI355X (gfx950) implementation of the
 * ATLAS12/SYNTHE data-parallel hot path.
 *
 * The reference's hot routines are Fortran subroutines that communicate
 * through module arrays (no argument lists).  Each entry point below takes
 * exactly those arrays, in their Fortran (column-major) layout, so a
 * Fortran driver reaches it through iso_c_binding without changing its
 * data (see atlas12_amd/fortran/a12gpu_iface.f90).  REAL(4) <-> float,
 * REAL(8) <-> double, default INTEGER <-> int32_t.
 *
 * All entry points return A12_OK (0) or a negative a12_status; the message
 * is available from a12_last_error().  A context is not thread-safe: use
 * one per host thread.  Results are deterministic (no atomics; fixed
 * reduction order).  The library never frees caller memory.
 */
#ifndef A12GPU_H
#define A12GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    A12_OK = 0,
    A12_ERR_ARG = -1,        /* bad dimension, null pointer, bad enum  */
    A12_ERR_HIP = -2,        /* HIP runtime error                       */
    A12_ERR_OOM = -3,        /* device allocation failed                */
    A12_ERR_STATE = -4,      /* call out of order (e.g. run before load)*/
    A12_ERR_NONFINITE = -5   /* non-finite input detected               */
} a12_status;

/* Arithmetic of the line-opacity build. */
typedef enum {
    A12_PREC_MIXED = 0,  /* FP32 Voigt and Neumaier-compensated FP32 sums;
                            FP64 Boltzmann factor and far-wing anchor   */
    A12_PREC_FP64 = 1    /* every quantity in FP64 (the FP64 GPU path) */
} a12_prec;

typedef struct a12_ctx a12_ctx;

/* lte_line_t, src/mod_mklinelist.f90:30-41 (32 bytes, BIND(C) mirror). */
typedef struct {
    int32_t nbuff;   /* grid index of the line centre (1-based)       */
    float   cgf;     /* line-strength factor                          */
    int32_t nelion;  /* species slot (nelem-1)*6 + ion stage          */
    float   elo;     /* lower-level energy [cm^-1]                    */
    float   gamrf, gamsf, gamwf; /* damping / (4 pi nu)               */
    float   code;    /* Kurucz species code                           */
} a12_lte_line;

/* State read by SYNTHE's two hot loops (src/synthe.f90:476-718), as the
 * driver holds it after run_xnfpelsyn (src/synthe_module.f90:250-382) and
 * READIN (src/atlas12_modules.f90:171, :1719).  Arrays are Fortran layout. */
typedef struct {
    int32_t nrhox;          /* layers (<= 128)                                */
    int32_t length;         /* wavelength points                              */
    double  wbegin;         /* first grid wavelength [nm]                     */
    double  ratio;          /* 1 + 1/resolu                                   */
    double  resolu;         /* resolving power                                */
    int32_t nedge;          /* continuum edges                                */
    const double *wledge;   /* (nedge) |wledge_m|                              */
    int32_t nf;             /* leading dimension of the continuum tables       */
    const double *continall;/* (nf, nrhox) log10(kappa_abs + kappa_scat)       */
    const double *contabs;  /* (nf, nrhox) log10 kappa_abs                     */
    const double *contscat; /* (nf, nrhox) log10 kappa_scat                    */
    int32_t mw6;            /* 6*mw                                           */
    const double *xnfpel;   /* (mw6, nrhox) xnfpel_m                           */
    const double *dopple;   /* (mw6, nrhox) dopple_m                           */
    const double *t, *xne, *rho, *xnfh, *xnfhe1, *xnfh2, *hckt, *hkt;
                            /* xf_t ... xf_hckt, xf_hkt (nrhox)                */
    const double *hkt_atlas;/* mod_atlas_data HKT (nrhox)                       */
    const double *rhox;     /* mod_atlas_data RHOX (nrhox)                      */
} a12_synthe_atm;

/* ---- context ------------------------------------------------------------ */
int a12_ctx_create(int32_t device, a12_ctx **out);
int a12_ctx_destroy(a12_ctx *ctx);
const char *a12_last_error(const a12_ctx *ctx);
/* hipStream_t the context launches on (as void*). */
void *a12_stream(a12_ctx *ctx);

/* JOSH operator matrices COEFJ, COEFH (51,51), as BLOCKJ/BLOCKH load them
 * (src/atlas12_modules.f90:18792-18860).  Required before any transfer. */
int a12_set_operators(a12_ctx *ctx, const double *coefj, const double *coefh);

/* ---- SYNTHE line opacity ------------------------------------------------ */
/* Replaces the LTE part of SYNTHE's depth loop, src/synthe.f90:476-689
 * (continuum :481-516, per-depth prep :518-551, line loop :557-684).
 * opacity_out: (length, nrhox), float for A12_PREC_MIXED, double for
 * A12_PREC_FP64.  evals_out (optional): accumulated profile contributions.
 * Host pointers in, host pointer out (synchronous). */
int a12_synthe_line_opacity(a12_ctx *ctx, const a12_synthe_atm *atm,
                            const a12_lte_line *lines, int32_t nlines,
                            int32_t prec, void *opacity_out, int64_t *evals_out);

/* Staged form of the same operation, for callers that keep inputs
 * resident: load copies atmosphere and lines to the device and sorts the
 * lines by centre; run executes the device pipeline into the context's
 * device opacity buffer (no host traffic); fetch copies it back. */
int a12_synthe_load(a12_ctx *ctx, const a12_synthe_atm *atm,
                    const a12_lte_line *lines, int32_t nlines);
int a12_synthe_opacity_run(a12_ctx *ctx, int32_t prec);
int a12_synthe_opacity_fetch(a12_ctx *ctx, void *opacity_out);
/* Tile shapes of the line-opacity gather (no counterpart in the reference;
 * results do not depend on the shape): shape idx -> {points per thread,
 * lines per LDS stage, 0 for the barrier-free per-wave form}; returns the
 * number of shapes, -1 past the end.
 * set_shape picks the shape of later runs (single and batched). */
int a12_synthe_shape_info(int32_t idx, int32_t pts_chunk[2]);
int a12_synthe_set_shape(a12_ctx *ctx, int32_t shape);

/* Batched SYNTHE: nmodel models, each with its own atmosphere (its own
 * nrhox, continuum tables and window: atms[k].length, wbegin; one
 * resolving power), over one shared line list sorted by nbuff.  Model k
 * takes lines line_lo[k] .. line_hi[k]-1 of the list, with nbuff
 * relative to its window = list nbuff - nb_shift[k] (windows on one
 * log-lambda grid); each model's result is what a12_synthe_line_opacity
 * (and a12_synthe_transfer_run) gives for that model and those lines,
 * replacing src/synthe.f90:476-718 per model.  The batch runs in groups
 * whose work arrays fit a budget (default half the free memory, at most
 * 64 GiB): per group the set-up kernels of each model and one persistent
 * gather launch over all (model, tile, depth) items.  transfer != 0 adds
 * the FP64 transfer of every model.  use_graph != 0 captures the whole
 * sequence into a hipGraph on the first call and replays it on later calls
 * with the same precision, shape and transfer flag. */
int a12_synthe_batch_load(a12_ctx *ctx, int32_t nmodel, const a12_synthe_atm *atms,
                          const int32_t *nb_shift, const int64_t *line_lo,
                          const int64_t *line_hi, const a12_lte_line *lines, int64_t nlines);
int a12_synthe_batch_set_budget(a12_ctx *ctx, int64_t bytes);
int a12_synthe_batch_run(a12_ctx *ctx, int32_t prec, int32_t transfer, int32_t use_graph);
/* Model k's opacity (length, nrhox) as a12_synthe_opacity_fetch, its
 * transfer output and evals [all, Voigt core]; any pointer may be NULL. */
int a12_synthe_batch_fetch(a12_ctx *ctx, int32_t model, void *opacity_out, double *hnu_line,
                           double *hnu_cont, int64_t evals[2]);
/* info: nmodel, groups of the last run, work bytes, budget, work bytes of
 * the largest model (the smallest budget that runs). */
int a12_synthe_batch_info(a12_ctx *ctx, int64_t info[5]);

/* Device pointer of the opacity buffer of the last run and its precision. */
const void *a12_synthe_opacity_device(a12_ctx *ctx, int32_t *prec_out);
/* Profile contributions accumulated by the last run (exact count). */
int64_t a12_synthe_evals(a12_ctx *ctx);

/* ---- SYNTHE transfer ---------------------------------------------------- */
/* Replaces the wavelength loop, src/synthe.f90:698-718 with
 * setup_opacity_sv :797-833 and process_wavelength_point :848-902 (two
 * JOSH(1,1) calls per point).  Uses the opacity of the last
 * a12_synthe_opacity_run on this context (already on the device).
 * Writes H_line and H_cont per point as the .spec file records them. */
int a12_synthe_transfer_run(a12_ctx *ctx);
int a12_synthe_transfer_fetch(a12_ctx *ctx, double *hnu_line, double *hnu_cont);
/* Host-array form: opacity (length, nrhox) float or double per prec. */
int a12_synthe_transfer(a12_ctx *ctx, const a12_synthe_atm *atm,
                        const void *opacity, int32_t prec,
                        double *hnu_line, double *hnu_cont);

/* ---- JOSH, batched ------------------------------------------------------ */
/* JOSH(IFSCAT, IFSURF) for nbatch independent frequencies,
 * src/atlas12_modules.f90:18398-18776.  Inputs are (nrhox, nbatch) arrays
 * (depth fastest): acont, aline, sigmac, sigmal, scont, sline, bnu; rhox is
 * shared (nrhox).  Outputs (nrhox, nbatch): taunu, snu, jnu, hnu, jmins
 * (any may be NULL for ifsurf = 1), knu1 (nbatch), hnu1 (nbatch).
 * Host pointers (synchronous). */
int a12_josh_batch(a12_ctx *ctx, int32_t ifscat, int32_t ifsurf,
                   int32_t nrhox, int32_t nbatch, const double *rhox,
                   const double *acont, const double *aline,
                   const double *sigmac, const double *sigmal,
                   const double *scont, const double *sline,
                   const double *bnu, double *hnu1, double *knu1,
                   double *taunu, double *snu, double *jnu, double *hnu,
                   double *jmins);

/* ---- SOLVIT, batched ---------------------------------------------------- */
/* SOLVIT(A, N, B, IPIVOT), src/atlas12_modules.f90:5739-5847, on nsys
 * independent systems of order n (1 <= n <= 64): a is (n, n, nsys), each
 * system column-major as the reference holds A; b is (n, nsys); x (n, nsys)
 * receives the solutions (b is not modified).  It runs the device forms
 * that JOSH (n = 51, the system in registers) and NMOLEC (the system in LDS)
 * use; both equal the serial routine bit for bit.  Host pointers
 * (synchronous). */
int a12_solvit_batched(a12_ctx *ctx, int32_t n, int32_t nsys, const double *a,
                       const double *b, double *x);

/* ---- ATLAS12 opacity sampling ----------------------------------------- */
/* State LINOP1 and the RUN_RT_PASS frequency loop read (module arrays of
 * mod_atlas_data, src/atlas12_modules.f90:1858-1919, and the driver's grid,
 * src/atlas12c.f90:451-491).  (ld, *) arrays keep the Fortran leading
 * dimension (kw). */
typedef struct {
    int32_t nrhox;           /* layers (<= 128)                                */
    int32_t numnu;           /* frequencies NUMNU                              */
    int32_t mion;            /* species slots (1006)                           */
    int32_t ldion;           /* leading dimension of xnfdop/dopple (kw)        */
    int32_t ldtab;           /* leading dimension of tabcont (kw)              */
    int32_t nwave1;          /* NWAVE + 1 continuum bins (362)                 */
    const double *waveset;   /* (numnu) WAVESET [nm]                           */
    const double *rcoset;    /* (numnu) RCOSET quadrature weights              */
    const double *tabcont;   /* (ldtab, nwave1) TABCONT                        */
    const int32_t *iwavetab; /* (nwave1) IWAVETAB with its sentinel            */
    const double *xnfdop;    /* (ldion, mion) XNFDOP                           */
    const double *dopple;    /* (ldion, mion) DOPPLE                           */
    const double *t, *xne, *hckt, *hkt, *rhox;    /* (nrhox)                    */
    const double *xnf1, *xnf3, *xnf841;           /* XNF(:,1), (:,3), (:,841)   */
    const double *h0tab, *h1tab, *h2tab;          /* (2001) TABVOIGT tables     */
    double teff;             /* TEFF (TCORR's cool-star rule)                  */
} a12_atlas_atm;

/* The line store LINEDATA(4, nlines), as SELECTLINES leaves it, on the
 * grid of atm (only numnu, waveset, nwave1, iwavetab and mion are read):
 * the per-record decoding of LINOP1 (src/atlas12_modules.f90:18176-18236:
 * TABLOG fields, continuum bin, vacuum wavelength, grid index NU, CGF, the
 * damping constants), then a stable sort by NU.  Once per store; nothing in
 * it depends on the atmosphere.  Invalidates any state set before. */
int a12_atlas_load_lines(a12_ctx *ctx, const a12_atlas_atm *atm, const int32_t *linedata,
                         int32_t nlines);
/* The per-iteration state LINOP1 and the RUN_RT_PASS loop read (module
 * arrays T, XNE, HCKT, HKT, RHOX, XNF(:,1/3/841) -> TXNXN as LINOP1 forms it
 * at :18163-18164, XNFDOP, DOPPLE, TABCONT, RCOSET, TABVOIGT, TEFF) on the
 * grid of the loaded store (A12_ERR_STATE if numnu, waveset, nwave1,
 * iwavetab or mion differ).  Once per iteration, before linop1/set_kapp;
 * it invalidates XLINES, the KAPP tables and the RT-pass sums of the
 * previous state. */
int a12_atlas_set_state(a12_ctx *ctx, const a12_atlas_atm *atm);
/* Both of the above in one call. */
int a12_atlas_load(a12_ctx *ctx, const a12_atlas_atm *atm, const int32_t *linedata,
                   int32_t nlines);
/* LINOP1, src/atlas12_modules.f90:18132-18374, into the context's device
 * XLINES (nrhox, numnu). */
int a12_atlas_linop1_run(a12_ctx *ctx, int32_t prec);
/* Tile shapes of the LINOP1 gather (no counterpart in the reference: a
 * launch-shape choice).  a12_atlas_shape_info returns the number of shapes
 * and, for 0 <= idx < that number, writes {points per tile, lanes per
 * block, lines per LDS batch, form} (form 0: staged; 1: a line's items
 * dealt out over the wave instead of walked per lane; 2: per wavefront
 * without barriers; 3: point form, one lane per (depth, points per tile / 8
 * points) with register sums, the third value then the XCD run length of
 * its workgroup numbering).  a12_atlas_set_shape picks one (default 0);
 * a12_atlas_autotune runs LINOP1 once, times the gather in every shape on
 * the loaded store (ms[shape], -1 where not applicable) and keeps the
 * fastest.  XLINES is bitwise the same in every shape. */
int a12_atlas_shape_info(int32_t idx, int32_t *pts_block_batch);
int a12_atlas_set_shape(a12_ctx *ctx, int32_t shape);
int a12_atlas_autotune(a12_ctx *ctx, int32_t prec, float *ms, int32_t *best);
/* Copies XLINES (nrhox, numnu), depth fastest, and the contribution count. */
int a12_atlas_linop1_fetch(a12_ctx *ctx, double *xlines, int64_t *evals);
/* KAPP's outputs per frequency, each (nrhox, numnu) depth fastest:
 * ACONT, SIGMAC, SCONT, SLINE, ALINE, SIGMAL (src/atlas12_modules.f90:3918). */
int a12_atlas_set_kapp(a12_ctx *ctx, const double *acont, const double *sigmac,
                       const double *scont, const double *sline, const double *aline,
                       const double *sigmal);
/* The frequency loop of RUN_RT_PASS (src/atlas12_modules.f90:3905-3949):
 * JOSH(1, 0) per frequency on KAPP + XLINES*STIM, the negative floors, and
 * the TCORR/RADIAP/ROSS mode-1 and mode-2 accumulators.  Uses the device
 * XLINES of the last linop1 run, or xlines (nrhox, numnu) when given. */
int a12_atlas_rt_run(a12_ctx *ctx, const double *xlines);
/* acc (9, nrhox): RJMINS, RDABH, RDIAGJ, FLXRAD, RADEN, H_TOTAL, ACCRAD,
 * ABROSS sums, then PRADK0 in acc[8*nrhox] (the rest of that row is 0);
 * hnu1 (numnu): HNU(1) per frequency after the floors. */
int a12_atlas_rt_fetch(a12_ctx *ctx, double *acc, double *hnu1);

/* No counterpart in the reference: the RT pass of a12_atlas_rt_run for
 * callers that cannot allocate device memory (code under HIP graph
 * capture, tight iteration loops, harnesses that forbid hipMalloc).  Every
 * pointer is device memory the caller owns:
 *   d_waveset, d_rcoset (numnu); d_depth3 (nrhox, 3): T, HKT, RHOX;
 *   d_kapp (nrhox, numnu, 6): ACONT, SIGMAC, SCONT, SLINE, ALINE, SIGMAL;
 *   d_xlines (nrhox, numnu); out d_hnu1 (numnu), d_acc (9, nrhox) as
 *   a12_atlas_rt_fetch lays it out.
 * Launches on the context's stream and returns without synchronising.
 * The per-chunk partial sums go to one fixed device array of the library
 * (9 * A12_RT_NOALLOC_MAX_CELLS doubles, 84.9 MB per GPU, held until the
 * process ends): nrhox * ceil(numnu / 2) must not exceed
 * A12_RT_NOALLOC_MAX_CELLS (A12_ERR_ARG otherwise).  All calls on one GPU
 * share that array, so calls that overlap in time on one GPU (two streams
 * or two host threads) race; calls in one stream do not.  Results are the
 * bits of a12_atlas_rt_run on the same inputs.  Requires
 * a12_set_operators. */
#define A12_RT_NOALLOC_MAX_CELLS 1179648
int a12_atlas_rt_pass_noalloc(a12_ctx *ctx, int32_t nrhox, int32_t numnu, double teff,
                              const double *d_waveset, const double *d_rcoset,
                              const double *d_depth3, const double *d_kapp,
                              const double *d_xlines, double *d_hnu1, double *d_acc);

/* TCORR(3) sections A-F, src/atlas12_modules.f90:2528-2752: the
 * flux-constancy (DTFLUX), Lambda-iteration (DTLAMB) and surface (DTSURF)
 * temperature corrections and the iteration damping, after ROSS(3),
 * RADIAP(3), COMPUTE_HEIGHT and CONVEC (src/atlas12c.f90:582-587).  The
 * sections that follow (CZ_CONSTRUCT, the RHOX correction, the remap onto
 * TAUSTD, :2816-2967) are a12_atlas_tcorr3_finish, on out's T1 and FLXERR.
 * Arrays (nrhox) are the module arrays of those names.  flxrad, rjmins,
 * rdiagj NULL: the sums of the last a12_atlas_rt_run, read on the device
 * (rdabh then optional, receiving the updated sum).  rdabh (TCORR's RDABH)
 * and oldt1 (OLDT1) are TCORR's SAVE state, updated in place.
 * out (7, nrhox): T1 (after damping), DTFLUX, DTLAMB, DTSURF, FLXERR (%),
 * FLXDRV, HRATIO.  Host pointers, synchronous. */
typedef struct {
    int32_t nrhox, ifconv, iter;
    double flux, teff, mixlth;
    const double *tauros, *rhox, *t, *dltdlp, *abross, *flxcnv, *flxcnv0, *grdadb, *ptotal,
        *rho, *dlrdlt, *heatcp, *hscale;
    const double *flxrad, *rjmins, *rdiagj;
} a12_tcorr_in;
int a12_atlas_tcorr3(a12_ctx *ctx, const a12_tcorr_in *in, double *rdabh, double *oldt1,
                     double *out);

/* ROSSTAB's Rosseland-opacity table (module state ROSSTAB_*,
 * src/atlas12_modules.f90:1779-1785) kept on the device by the context.
 * a12_rosstab_store is ROSSTAB(0, 0, 0) (:6003-6042; the driver's call at
 * src/atlas12c.f90:583): it appends the nrhox layers (T, P, ABROSS),
 * the first store fixing the normalisation, and stops at
 * A12_ROSSTAB_MAXTAB entries as the reference does.  a12_rosstab_reset
 * empties it (a new model); a12_rosstab_load replaces it by a given table
 * (zero: ZEROT, SLOPET, ZEROP, SLOPEP; tab (3, nross): TABT, TABP, ROSS,
 * entry fastest); a12_rosstab_fetch copies it out (tab may be NULL).
 * a12_rosstab_lookup is ROSSTAB(T, P, V) (:6047-6268) for nq queries in
 * mode 1 (the reference's IROSSTAB default: nearest point per (T, P)
 * quadrant and a bilinear blend), 2 (Shepard) or 3 (moving least squares);
 * one wavefront per query.  Host pointers, synchronous. */
#define A12_ROSSTAB_MAXTAB 4800
int a12_rosstab_reset(a12_ctx *ctx);
int a12_rosstab_store(a12_ctx *ctx, int32_t nrhox, const double *t, const double *p,
                      const double *abross);
int a12_rosstab_load(a12_ctx *ctx, int32_t nross, const double *zero, const double *tab);
int a12_rosstab_fetch(a12_ctx *ctx, int32_t *nross, double *zero, double *tab);
int a12_rosstab_lookup(a12_ctx *ctx, int32_t mode, int32_t nq, const double *temp,
                       const double *pressure, double *kappa);

/* TTAUP(T, TAU, ABSTD, PTOTAL, P, PRAD, PTURB, VTURB, GRAV, NUMTAU),
 * src/atlas12_modules.f90:6303-6523: the hydrostatic march on a log-uniform
 * tau grid with the context's ROSSTAB table in mode irosstab (predictor,
 * corrector with stall-driven relaxation, best-iterate fallback, surface
 * bootstrap).  One wavefront; *nfail (optional) = depths whose corrector
 * hit its cap of 50.  Host pointers, synchronous. */
int a12_ttaup(a12_ctx *ctx, int32_t irosstab, int32_t n, const double *t, const double *tau,
              const double *prad, const double *pturb, double grav, double *abstd,
              double *ptotal, double *p, int32_t *nfail);

/* CZ_CONSTRUCT's SAVE state (src/atlas12_modules.f90:3054-3069) and the
 * module flags CZC_RUN_ENGAGED / CZC_RUN_JC it sets; the caller keeps it
 * between iterations (start: all 0, last_dtmax 1e30; ITER <= 1 resets it
 * as the reference does). */
typedef struct {
    int32_t prev_engaged, jc_sticky, nbot_bad, run_engaged, run_jc;
    double last_dtmax, rel_err;
} a12_czc_state;

/* The module arrays CZ_CONSTRUCT and CZC_DELREQ_LAYER read (nrhox each)
 * and the switches USE_CZ_CONSTRUCTOR, USE_4POINT_MLT, IROSSTAB. */
typedef struct {
    int32_t nrhox, ifconv, iter, use_constructor, use4pt, irosstab;
    double flux, mixlth, grav;
    const double *flxcnv, *flxrad, *dltdlp, *grdadb, *flxerr, *t, *rhox, *ptotal, *rho,
        *heatcp, *dlrdlt, *p, *vturb, *abross, *hscale;
} a12_czc_in;

/* CZ_CONSTRUCT(T1), src/atlas12_modules.f90:3012-3287 with
 * CZC_DELREQ_LAYER :3298-3358: the efficient-convection temperature
 * corrector folded into t1 (nrhox, in place); *engaged (optional) 1 when
 * it changed t1.  Uses the context's ROSSTAB table.  One wavefront. */
int a12_cz_construct(a12_ctx *ctx, const a12_czc_in *in, a12_czc_state *st, double *t1,
                     int32_t *engaged);

/* TCORR(3) after its table, src/atlas12_modules.f90:2816-2967:
 *   (G2) CZ_CONSTRUCT on t1;
 *   (H)  TAUSTD = 10**(TAU1LG + (J-1) STEPLG), TTAUP on T and on T + T1
 *        mapped onto it, DRHOX from the pressure change mapped back;
 *   (I)  T += T1, the optional 3-point smoothing (J1SMOOTH..J2SMOOTH),
 *        monotonicity (T(J) <= T(J+1) - 1), the floor TFLOOR_ATM;
 *   (J)  TK, HKT, HCKT, TKEV, TLOG;
 *   (K)  RHOX += DRHOX, MAP1 of RHOX, T, P, XNE, ABROSS, PRAD, VTURB, BMIN,
 *        PTURB and BHYD(:, 1:6) onto TAUSTD, the ABROSS guard, PRADK =
 *        PRAD + PRADK0, TAUROS = TAUSTD.
 * The Newton step (G) is off in the reference (USE_DTGRAD = .FALSE.,
 * :1457), and CZC_POLISH (:2974, final iterations, only after the
 * constructor engaged: st->run_engaged) re-runs the EOS and CONVEC: both
 * stay with the driver.  t1 (nrhox): T1 after section (F) in, after (G2)
 * out.  io arrays are the module arrays, updated in place; drhox (nrhox,
 * optional) out; *nfail (optional) TTAUP depths not converged.  Uses the
 * context's ROSSTAB table.  One wavefront, host pointers, synchronous. */
typedef struct {
    int32_t nrhox, ifconv, iter, use_constructor, use4pt, irosstab, j1smooth, j2smooth;
    double flux, mixlth, grav, tau1lg, steplg, pradk0, tfloor, wtjm1, wtj, wtjp1;
    const double *flxcnv, *flxrad, *dltdlp, *grdadb, *flxerr, *ptotal, *rho, *heatcp,
        *dlrdlt, *hscale;
} a12_tcorr_fin_in;
typedef struct {
    double *tauros, *rhox, *t, *p, *xne, *abross, *prad, *vturb, *bmin, *pturb; /* in/out */
    const double *accrad;  /* remapped by (K) into a scratch column, not kept */
    double *bhyd;          /* (nrhox, 6) BHYD, in/out                          */
    double *tk, *hkt, *hckt, *tkev, *tlog, *pradk;                             /* out    */
} a12_tcorr_fin_io;
int a12_atlas_tcorr3_finish(a12_ctx *ctx, const a12_tcorr_fin_in *in, a12_tcorr_fin_io *io,
                            a12_czc_state *st, double *t1, double *drhox, int32_t *nfail);

/* CONVEC(MLT_ONLY), src/atlas12_modules.f90:10323-10953 (the driver's call
 * src/atlas12c.f90:586, after ROSS(3), the ROSSTAB store, RADIAP(3) and
 * COMPUTE_HEIGHT).  CONVEC evaluates the EOS four times, at T x 1.001,
 * T x 0.999 and then (T restored) P x 1.001, P x 0.999 (:10468-10528); the
 * EOS (COMPUTE_ONE_POP) stays with the driver, which passes the module
 * EDENS and RHO after each of those evaluations: edens, rho4 (nrhox, 4),
 * depth fastest, in that order.  On the device: the finite-difference
 * thermodynamics (DLTDLP with the log-space bottom slope, HEATCP, VELSND,
 * DLRDLT, GRDADB, HSCALE), the MLT solve of each depth with its bubble-
 * opacity iteration on the context's ROSSTAB table (Picard, Steffensen,
 * cycle-mean exit; one wavefront per depth), then, unless mlt_only, the
 * gap fill, the bottom clamp (IFCORR = 1), the 1-2-1 smoothing
 * (USE_FLXCNV_SMOOTH), overshooting (OVERWT > 0) and the zeroing of the
 * top NCONV layers.  io: T, TK, HKT, HCKT, TKEV, TLOG, P in, returned as
 * CONVEC's perturb-and-restore leaves them (within a few ulps of the
 * input); FLXCNV1 in/out (kept unless overshooting); the rest out
 * (ABCONV, DELTAT: CONVEC's locals).  Host pointers, synchronous. */
typedef struct {
    int32_t nrhox, ifconv, ifcorr, nconv, use4pt, smooth, irosstab;
    double grav, mixlth, flux, overwt;
    const double *ptotal, *rho, *pradk, *tauros, *rhox, *abross, *vturb, *height, *flxrad;
    const double *edens, *rho4;
} a12_convec_in;
typedef struct {
    double *t, *tk, *hkt, *hckt, *tkev, *tlog, *p;
    double *dltdlp, *heatcp, *velsnd, *dlrdlt, *grdadb, *hscale, *vconv, *flxcnv, *flxcnv0,
        *flxcnv1, *abconv, *deltat;
} a12_convec_io;
int a12_convec(a12_ctx *ctx, const a12_convec_in *in, a12_convec_io *io, int32_t mlt_only);

/* NMOLEC's gas-phase Newton solve, src/atlas12_modules.f90:9078-9217, for
 * a batch of systems (depths of one or many models).  The molecule table
 * is READMOL's, 1-based as in the reference: locj (nummol + 1) the first
 * component of each entry in kcomps; kcomps (locj[nummol] - 1) the
 * equation of each component (2..nequa are elements, nequa + 1 is the
 * electron as a divisor); idequa (nequa) the element of each equation, 100
 * for the electron (charge) equation.  Sizes as the reference's maxeq and
 * maxmol (:162-167). */
#define A12_MOL_MAXEQ  64
#define A12_MOL_MAXMOL 400
typedef struct {
    int32_t nequa, nummol;
    const int32_t *locj, *kcomps, *idequa;
} a12_moltab;
int a12_molec_set_table(a12_ctx *ctx, const a12_moltab *tab);
/* nsys systems in nchain chains: chain c is systems chain_off[c] ..
 * chain_off[c+1]-1 and runs as NMOLEC's depth loop (:8879-8884): its first
 * system starts from xn, each later one from the previous solution times
 * pratio[s] (= P(J)/P(J-1)).  Chains of length one are independent solves
 * (pratio may then be NULL).  Per system s: xntot[s] = P/kT; xab (nsys,
 * nequa) element abundances, max(XABUND, 1e-20), 0 for the electron row
 * (entry 1 unused); equilj (nsys, nummol) the equilibrium constants as
 * NMOLEC forms them (:8897-8946; the caller's PFSAHA); xn (nsys, nequa)
 * start in (chain heads), solution out.  xnmol (nsys, nummol, may be NULL)
 * = XNMOL (:9624-9636).  iters: Newton iterations; status 0 converged, 1
 * NaN (the reference prints a warning and keeps the state), 2 no
 * convergence in 5000 iterations.  Host pointers, synchronous.  One
 * wavefront per chain, bit-identical to the serial routine. */
int a12_molec_newton_batched(a12_ctx *ctx, int32_t nsys, int32_t nchain,
                             const int32_t *chain_off, const double *xntot, const double *pratio,
                             const double *xab, const double *equilj, double *xn, double *xnmol,
                             int32_t *iters, int32_t *status);

/* MOL2 32-byte molecular record (header: int32 magic 'MOL2', int32 count,
 * 6 more int32), src/atlas12_modules.f90:19540-19556. */
typedef struct {
    float   wl;       /* vacuum wavelength [nm]                         */
    float   loggf;    /* log gf, bare (isotope correction not applied)  */
    float   e1, e2;   /* the two term energies [cm^-1]                  */
    int32_t code;     /* molecule code (822 = TiO)                      */
    int32_t iso;      /* isotope (46..50 for Ti)                        */
    int32_t gr100;    /* 100 log10(gamma_rad), 0 = classical            */
    int32_t flag;     /* upper-label flag (unused)                      */
} a12_mol2_rec;

/* SELECTLINES' filter, src/atlas12_modules.f90:19171-19781: path 1 (the
 * predicted 16-byte list, :19262-19313, records kept as read) then path 5
 * (the TiO MOL2 list, :19557-19647, skipped above TEFF 8000 K, records
 * repacked).  atm supplies TABCONT, IWAVETAB, XNFDOP, HCKT, TEFF (waveset,
 * rcoset, dopple and the depth arrays other than hckt are not read);
 * wavetab is WAVETAB (nwave1 - 1) [nm].  Both lists must ascend in
 * wavelength and lie within the continuum bins (A12_ERR_ARG otherwise).
 * Writes the kept store LINEDATA(4, *nkept) to linedata_out; if *nkept
 * exceeds cap nothing is written and A12_ERR_ARG is returned with the
 * count in *nkept.  Host pointers, synchronous. */
int a12_atlas_select_lines(a12_ctx *ctx, const a12_atlas_atm *atm, const double *wavetab,
                           const int32_t *pred, int64_t npred, const a12_mol2_rec *mol2,
                           int64_t nmol2, int32_t *linedata_out, int64_t cap, int64_t *nkept);

/* diatomic_record_t (src/mod_mklinelist.f90:71-78): what
 * read_diatomics_for_atlas (:2244-2299) returns for SELECTLINES, 48 bytes. */
typedef struct {
    double  wlvac_nm;     /* vacuum wavelength [nm]                           */
    double  gflog_dex;    /* log gf with the isotope corrections x1 + x2     */
    double  elo_cm;       /* lower term energy [cm^-1]                        */
    double  gammar_log;   /* log10 gamma_rad                                   */
    double  gammaw_log;   /* log10 gamma_w                                     */
    int32_t molcode;      /* ATLAS12 MOLCODE, one of the 46 SELECTLINES knows */
    int32_t pad;
} a12_diatomic_rec;

/* h2opokazatel.bin record, 8 bytes (src/atlas12_modules.f90:19659-19680):
 * the signs of ielo and igflog carry the isotope. */
typedef struct {
    int32_t iwl;      /* NINT-encoded ln(lambda) / ln(1 + 1/2e6)          */
    int16_t ielo;     /* +-E_lo [cm^-1]                                     */
    int16_t igflog;   /* +-(1000 log gf + 16384)                            */
} a12_h2o_rec;

/* SELECTLINES' seven lists in store order (src/atlas12_modules.f90:
 * 19262-19767); any may be empty (NULL with count 0).  The 16-byte lists
 * are LINEDATA records, int32 (n, 4). */
typedef struct {
    const int32_t *pred;              int64_t npred;     /* gfpred29dec2014.bin, :19262 */
    const int32_t *lowobs;            int64_t nlowobs;   /* lowobsat12.bin, :19316      */
    const int32_t *hilines;           int64_t nhilines;  /* hilines.bin, :19355         */
    const a12_diatomic_rec *dia;      int64_t ndia;      /* ASCII diatomics, :19417     */
    const a12_mol2_rec *mol2;         int64_t nmol2;     /* TiO MOL2, :19557            */
    const a12_h2o_rec *h2o;           int64_t nh2o;      /* h2opokazatel.bin, :19650    */
    const int32_t *h3p;               int64_t nh3p;      /* h3plus.dat, :19723          */
} a12_sel_lists;

/* SELECTLINES over all seven lists (src/atlas12_modules.f90:19171-19781),
 * the store equal to the reference's record for record:
 *   pred     kept as read; NELION outside 1..mion skipped;
 *   lowobs   kept as read; its wavelength walk goes on from pred's;
 *   hilines  kept as read;
 *   dia      in read_diatomics_for_atlas's order; sorted here by
 *            NINT(ln(lambda) / ln(1 + 1/2e6)) with SELECTLINES' heapsort
 *            (its order among equal keys), ISOX-corrected, repacked;
 *   mol2     TiO, repacked; skipped above TEFF 8000 K;
 *   h2o      isotope from the signs, width of the walk's bin, repacked;
 *            skipped above TEFF 8000 K;
 *   h3p      log gf - 1.272, repacked.
 * Lists must ascend in wavelength and lie within the continuum bins; a
 * NELION outside 1..mion in lowobs or hilines (which the reference indexes
 * unchecked) and an unknown diatomic MOLCODE (where the reference exits)
 * return A12_ERR_ARG.  Other arguments as a12_atlas_select_lines, which is
 * this call with pred and mol2 only.  Host pointers, synchronous. */
int a12_atlas_select_all(a12_ctx *ctx, const a12_atlas_atm *atm, const double *wavetab,
                         const a12_sel_lists *lists, int32_t *linedata_out, int64_t cap,
                         int64_t *nkept);

/* ---- KAPP: continuum opacity ----------------------------------------------- */
/* KAPP (src/atlas12_modules.f90:11202-11324) with every source it calls:
 * HOP (:11359), H2PLOP (:12224), HMINOP (:12360), H2MINOP (:12611),
 * H2COLLOP (:16371), HRAYOP (:12740), HE1OP (:12939), HE2OP (:13569),
 * HEMIOP (:13716), HERAOP (:13918), CONT_METAL_OPACITY_TOPBASE (:14147:
 * MBF_TOPBASE, MBF_HIGH_ION, FELO_OPACITY for Fe I and Fe II, CHOP and
 * OHOP), ELECOP (:17212), H2RAOP (:17289) and HLINOP (:17381, with
 * STARK_MMM :17854 and STARK :17528).  Replaces the 30,000 KAPP calls per
 * RUN_RT_PASS (:3918): a12_atlas_kapp_run fills the tables that
 * a12_atlas_set_kapp would upload.
 *
 * Tables: what the sources read from DATADIR or hold after their first
 * call, in the reference's units.  Arrays are C order with the last index
 * fastest; the Fortran (first-index-fastest) arrays map as noted. */
typedef struct {
    int32_t loaded;                  /* STEHLE_DATA(s)%loaded */
    int32_t n_upper_min, n_upper_max, n_dens;
    const double *dens;              /* (20) density_grid */
    const double *temp;              /* (10) temp_grid */
    const double *log_dalpha;        /* (60) log_dalpha_grid */
    const int32_t *max_dens_idx;     /* (n_trans) */
    const double *k_alpha;           /* (n_trans) */
    const double *profiles;          /* (n_trans, n_dens, 10, 60) = profiles(60,10,n_dens,n_trans) */
} a12_kapp_stark;

typedef struct {
    /* XKARZAS (:11635), REAL(4) as read from karzas_*.dat */
    const float *k_freqn;            /* (15, 29) = FREQN(29, 15) */
    const float *k_xn;               /* (15, 29) = XN(29, 15) */
    const float *k_xl;               /* (6, 6, 29) = XL(29, 6, 6): [n][l][point] */
    const float *k_ekarzas;          /* (29) */
    /* COULFF (:11995): log10 g_ff of gauntff_vanhoof.dat */
    const double *gff_logtab;        /* (146, 81) = GFF_LOGTAB(81, 146) */
    /* H2COLLOP (:16371): h2collop.dat */
    int32_t cia_npair, cia_ntmax, cia_nnu;
    double cia_nu0, cia_dnu;
    const int32_t *cia_nt;           /* (npair) */
    const double *cia_temp;          /* (npair, ntmax) */
    const double *cia_tab;           /* (npair, ntmax, nnu) = CIATAB(nnu, ntmax, npair) */
    const int32_t *cia_use;          /* (npair) USE_CIA_PAIR, may be NULL (all on) */
    /* CHOP (:16103), OHOP (:16228): crossch.dat, crossoh.dat */
    const double *cross_ch;          /* (105, 15) = CROSSCH(15, 105) */
    const double *cross_oh;          /* (130, 15) = CROSSOH(15, 130) */
    /* MBF_HIGH_ION (:14338): HI_EDGE_DATA after INIT_MBF_HIGH_ION */
    int32_t hi_nedge;
    const double *hi_edge;           /* (nedge, 7) = HI_EDGE_DATA(7, nedge) */
    /* FELO_OPACITY (:14710): Fe I, Fe II after INIT_FELO */
    int32_t felo_nlev[2], felo_xnfp_idx[2];
    const double *felo_nuth, *felo_g, *felo_e;   /* (2, 60) */
    const int32_t *felo_ifancy, *felo_npts;      /* (2, 60) */
    const double *felo_x, *felo_logsig;          /* (2, 60, 200) */
    const double *felo_seaton;                   /* (2, 60, 4) */
    /* MBF_TOPBASE (:14226): TB_SPECIES after INIT_MBF_TOPBASE, levels of all
     * species concatenated in species order, their points likewise */
    int32_t tb_nspecies, tb_nlevels, tb_npoints;
    const int32_t *tb_xnfp_idx, *tb_nlev;        /* (nspecies) */
    const double *tb_lowest_thr;                 /* (nspecies) LOWEST_THR_RY */
    const int32_t *tb_g, *tb_np;                 /* (nlevels) G_STAT, NP */
    const double *tb_ebind, *tb_eexc;            /* (nlevels) E_BIND_RY, E_EXC_RY */
    const double *tb_logemin, *tb_logestride;    /* (nlevels) */
    const double *tb_ephot, *tb_sigma;           /* (npoints) E_PHOTON_RY, SIGMA_MB */
    /* STARK_MMM (:17854): STEHLE_DATA after INIT_STARK_TABLES */
    a12_kapp_stark stark[4];
} a12_kapp_tables;

/* The depth state KAPP reads (module arrays after POPS): Fortran arrays with
 * leading dimension ld (kw); XNF/XNFP (ld, mion), BHYD (ld, 6).  equilh2 is
 * EQUILH2(T(J)) (:22437), the EOS's H2 equilibrium constant. */
typedef struct {
    int32_t nrhox, ld, mion;
    const double *t, *tkev, *tlog, *hkt, *hckt, *xne, *rho, *bmin;
    const double *bhyd, *equilh2, *xnf, *xnfp;
} a12_kapp_state;

/* Copies the tables to the device (once per run; A12_ERR_ARG on
 * inconsistent counts or a missing array). */
int a12_kapp_set_tables(a12_ctx *ctx, const a12_kapp_tables *tables);
/* Uploads the depth state and forms the per-state factors the sources keep
 * between calls (the ITEMP/TB_T_LAST caches: occupation probabilities,
 * Boltzmann factors, CIA weights, H- and H2 densities). */
int a12_kapp_set_state(a12_ctx *ctx, const a12_kapp_state *state);
/* KAPP at every frequency of the loaded ATLAS12 grid (WAVESET of
 * a12_atlas_load_lines), written as the six device tables a12_atlas_rt_run
 * reads; same layout and effect as a12_atlas_set_kapp.  The state's nrhox
 * must equal the grid's. */
int a12_atlas_kapp_run(a12_ctx *ctx);
/* Copies the six KAPP tables on the device (from a12_atlas_kapp_run or
 * a12_atlas_set_kapp) to kapp6 (6, numnu, nrhox), in a12_atlas_set_kapp's
 * order and layout. */
int a12_atlas_kapp_fetch(a12_ctx *ctx, double *kapp6);
/* KAPP at nwave wavelengths (nm), with RUN_RT_PASS's frequency set-up
 * (:3905-3915); out (nwave, 23, nrhox), host memory, per wavelength:
 * AHYD SHYD AH2P AHMIN SHMIN SIGH AHE1 AHE2 AHEMIN SIGHE AH2COLL AH2MIN
 * ACONT_METAL SIGEL SIGH2 AHLINE SHLINE ACONT SIGMAC SCONT SLINE ALINE
 * SIGMAL. */
#define A12_KAPP_NSRC 23
int a12_kapp_eval(a12_ctx *ctx, int32_t nwave, const double *wave_nm, double *out);
/* Milliseconds of the last a12_atlas_kapp_run (with a12_set_profiling on),
 * -1 when not measured. */
int a12_kapp_ms(a12_ctx *ctx, float *ms);
/* The physical constants the library's KAPP uses, in this order: HCK,
 * HOVERK, KBOL_EV, THETA_COEFF, BNU_PREFAC, SAHA_PREFAC, FREQ_RYDH,
 * RYD_ANG, LN10, COEFF_FF, SIGMA_THOMSON, CLIGHT_ANGS (mod_constants,
 * :318-407).  No GPU needed. */
int a12_kapp_constants(double *out12);

/* ---- SYNTHE complex-profile lines ----------------------------------------- */
/* nlte_line_t, src/mod_mklinelist.f90:43-62: one complex-profile line as
 * run_mklinelist keeps it for compute_line_opacity (the fort.19 record;
 * 60 bytes of fields, 64 with C's trailing alignment). */
typedef struct {
    double  wlvac;          /* vacuum wavelength [nm]                           */
    float   elo;            /* lower-level energy [cm^-1]                       */
    float   gf;             /* CGF; raw gf for types 1 and >= 4                 */
    int32_t nblo, nbup;     /* H I: lower / upper principal quantum numbers     */
    int32_t nelion;         /* species slot (nelem-1)*6 + ion stage, 1-based    */
    int32_t itype;          /* 2 coronal; 0, 3, <= -3 Voigt; -1 H I, -2 D I;
                               1 autoionising; >= 4 merged continuum (NLAST)   */
    int32_t ncon;           /* continuum edge (CONTX row, CONTH for H I)        */
    int32_t nelionx;        /* CONTX column                                     */
    float   gammar, gammas, gammaw;  /* damping / (4 pi nu); autoionising:
                                        width, Shore q, Shore w                 */
    int32_t nbuff;          /* grid index of the centre (1-based)               */
    int32_t lim;            /* wing class (not read by the opacity)             */
} a12_cplx_line;

/* Replaces compute_line_opacity (src/synthe_module.f90:3167-3702, with
 * hydrogen_line_profile :1032-1980), which SYNTHE's depth loop calls before
 * the LTE lines (src/synthe.f90:552).  atm: the atmosphere given to
 * a12_synthe_load (its window, nrhox and mw6 must match the loaded one);
 * lines in the reference's order; stark: the four Stehle series
 * (STEHLE_DATA, as for a12_kapp_set_tables); wlend: the window's WLEND;
 * cutoff: LINE_CUTOFF; ifvac: IFVAC; velshift: the call's velocity shift
 * (km/s, 0 in the reference's driver).  Copies everything to the device.
 * A12_ERR_ARG for a hydrogen line with n outside 1..99 or m outside 2..100
 * (where the reference stops), a merged hydrogen line whose NCON is outside
 * 1..15, a Voigt line with NCON > 0 outside CONTX (1..26, 1..17), or a
 * species slot outside 1..mw6.  The lines stay until a12_cplx_clear or the
 * next a12_synthe_load. */
int a12_cplx_load(a12_ctx *ctx, const a12_synthe_atm *atm, const a12_cplx_line *lines,
                  int32_t nlines, const a12_kapp_stark *stark, double wlend, double cutoff,
                  int32_t ifvac, float velshift);
/* The complex-line opacity of every depth on the device: BUFFER after the
 * call at each depth, (nrhox, length) FP32 in the context.  Once loaded,
 * every later a12_synthe_opacity_run also adds it to the opacity. */
int a12_cplx_run(a12_ctx *ctx);
/* Copies the last run's buffer (nrhox, length) and, when extents is not
 * NULL, the points each (depth, line) reached: (nrhox, nlines, 4) =
 * first/last red, first/last blue (1-based; empty when first > last; a
 * merged-continuum edge reports its range as red). */
int a12_cplx_fetch(a12_ctx *ctx, float *buffer, int32_t *extents);
/* Drops the loaded lines; later SYNTHE runs hold the LTE lines only. */
int a12_cplx_clear(a12_ctx *ctx);
/* Milliseconds of the parts of the last a12_cplx_run (with
 * a12_set_profiling on): set-up and plans, convolutions, wing extents,
 * accumulation; -1 when not measured. */
int a12_cplx_ms(a12_ctx *ctx, float ms[4]);

/* ---- model start ---------------------------------------------------------- */
/* SCALE_MODEL(TEFF_NEW, LOGG_NEW), src/atlas12_modules.f90:6967-7060, for
 * nmodel models at once (one wavefront each): regrid onto TAUSTD(i) =
 * 10**(tau1lg + (i-1) steplg), i = 1..kw, by INTEG of ABROSS and MAP1, then
 * rescale to the new Teff and log g (TEFF_NEW = 0: regrid only; no rescale
 * either when Teff changes by < 1 K and g by < 1e-3 relative).
 *   sc     (8, nmodel): TEFF, GRAV, PRADK0, PCON, PTURB0, TEFF_NEW, LOGG_NEW, -
 *   fields (nrhox, 7, nmodel): RHOX, T, P, XNE, ABROSS, PRAD, VTURB
 *   out    (kw, 9, nmodel): RHOX, T, P, XNE, ABROSS, PRAD, VTURB, PRADK,
 *          PTOTAL (0 where not rescaled: the reference leaves it untouched)
 *   so     (6, nmodel): TEFF, GRAV, GLOG, FLUX, PRADK0, PZERO (GLOG, FLUX,
 *          PZERO NaN where not rescaled)
 *   taustd (kw, optional): TAUSTD.
 * The NLTE fields BMIN/BHYD are not carried (NLTE is off by default).
 * Host pointers, synchronous. */
int a12_scale_model(a12_ctx *ctx, int32_t nmodel, int32_t nrhox, int32_t kw, double tau1lg,
                    double steplg, const double *sc, const double *fields, double *out,
                    double *so, double *taustd);

/* ---- model iteration (reduced physics) ------------------------------------ */
/* The ATLAS12 iteration loop (src/atlas12c.f90:496-606) for nmodel models in
 * one launch sequence per iteration (captured once in a hipGraph and
 * replayed), on the fixed Rosseland grid TAUSTD(j) = 10**(-6.875 +
 * 0.125 (j-1)), j = 1..A12_MODEL_NRHOX, with KAPP and the EOS replaced by
 * closed forms (H, He and one metal donor in Saha/charge balance; H-
 * bound-free and free-free, H I bound-free (n <= 8) and free-free, the
 * donor's bound-free, electron and H Rayleigh scattering; no lines) and the
 * structure put in hydrostatic equilibrium on the grid each iteration.  On
 * the device as in the ATLAS12 path above: the RUN_RT_PASS loop (JOSH(1,0)
 * with the TCORR/RADIAP/ROSS sums) on A12_MODEL_NFREQ frequencies (the
 * driver's 1e-4 dex grid from NUSTART 9599 at every 25th point, RCOSET's
 * rule), ROSS(3) and TCORR(3) without convection, then T = max(T + T1,
 * 1500 K).  No counterpart in the reference as a whole: KAPP and POPS are
 * outside this library; this is the loop that converges the models the
 * starting-structure network is trained on and compared with.
 *   par     (nmodel): Teff [K], log g, [M/H], vturb [km/s];
 *   t_start (nmodel, A12_MODEL_NRHOX): starting T on the grid;
 *   itmax:  iterations at most (1..1000);
 *   t_out   (nmodel, A12_MODEL_NRHOX): T at the end (the converged
 *           structure, or the last one);
 *   s_out   (nmodel, 5, A12_MODEL_NRHOX): P_gas, P_total, RHOX, n_e of the
 *           last hydrostatic pass and ABROSS (ROSS(3)) of the last RT pass;
 *   iters   (nmodel): the first iteration whose TCORR(3) table has
 *           max |ERROR| <= 1 %, max |DERIV| <= 10 % and max |T1| <= 1 K
 *           (the model is kept as it was at that iteration's start), 0 if
 *           none within itmax, -k if iteration k produced a non-finite T1;
 *   hist    (nmodel, itmax, 3), may be NULL: max |ERROR| %, max |DERIV| %,
 *           max |T1| K of every iteration run (untouched after convergence).
 * Host pointers, synchronous. */
#define A12_MODEL_NRHOX 72
#define A12_MODEL_NFREQ 1200
typedef struct {
    double teff, logg, mh, vturb;
} a12_model_par;
int a12_model_converge(a12_ctx *ctx, int32_t nmodel, const a12_model_par *par,
                       const double *t_start, int32_t itmax, double *t_out, double *s_out,
                       int32_t *iters, double *hist);

/* ---- starting-structure network (no counterpart in the reference) ------- */
/* A 4 -> 256 -> 256 -> 256 -> 256 -> 144 MLP (SiLU hidden layers, 235,664
 * parameters) that maps (Teff, log g, [M/H], vturb) -- standardised by the
 * caller -- to T and log Pgas on 72 depths, the start SCALE_MODEL otherwise
 * provides.  BF16 operands on MFMA, FP32 accumulation and master weights,
 * Adam (0.9, 0.999, 1e-8), mean-squared-error loss.
 *   init:   capacity `batch` samples, Glorot-uniform weights from seed;
 *   params: 235,664 floats, W1 (4 x 256, row-major: in x out), b1, W2, b2,
 *           W3, b3, W4, b4, W5 (256 x 144), b5 (set also resets Adam);
 *   set_data: the training set, x (n x 4), t (n x 144), n <= batch;
 *   grads:  loss and gradients (params order) of the set data, no update;
 *   train:  `steps` full-batch Adam steps at rate lr (one hipGraph replay
 *           per step); loss of the last step's forward pass;
 *   infer:  y (n x 144) for x (n x 4), any n;
 *   gemm:   C (m x n, FP32) = A (m x k) B (k x n), BF16 row-major operands,
 *           the MFMA tile kernel on its own (split = 1) or with k split
 *           `split` ways and the slices summed in order (as the weight
 *           gradients are), for checks. */
int a12_mlp_init(a12_ctx *ctx, int32_t batch, uint64_t seed);
int a12_mlp_get_params(a12_ctx *ctx, float *params);
int a12_mlp_set_params(a12_ctx *ctx, const float *params);
int a12_mlp_set_data(a12_ctx *ctx, const float *x, const float *t, int32_t n);
int a12_mlp_grads(a12_ctx *ctx, float *loss, float *grads);
int a12_mlp_train(a12_ctx *ctx, int32_t steps, float lr, float *loss);
int a12_mlp_infer(a12_ctx *ctx, const float *x, int32_t n, float *y);
int a12_mlp_gemm(a12_ctx *ctx, int32_t m, int32_t n, int32_t k, const uint16_t *a,
                 const uint16_t *b, float *c, int32_t split);

/* ---- measurement --------------------------------------------------------- */
/* Stage timing with HIP events on the context stream (off by default). */
int a12_set_profiling(a12_ctx *ctx, int32_t on);
/* ms[0..4]: depth prep + continuum, line set-up, class partition, tile
 * gather (of the last opacity run), SYNTHE transfer (of the last transfer
 * run); -1 where not measured.  Synchronises on the events. */
int a12_stage_ms(a12_ctx *ctx, float *ms);
/* ms[0..3] of the last ATLAS12 runs: LINOP1 line set-up, segment reach,
 * gather, and the RUN_RT_PASS frequency loop; -1 where not measured. */
int a12_atlas_stage_ms(a12_ctx *ctx, float *ms);
/* Contributions of the last opacity run: all, and those on the Voigt core
 * (centre included); the difference is the n^-2 far-wing terms. */
int a12_synthe_eval_split(a12_ctx *ctx, int64_t *total, int64_t *core);

/* ---- Voigt function, batched ------------------------------------------- */
/* voigt_profile(v, a), src/synthe_module.f90:463-537, on n points.
 * prec 0: the mixed path's FP32 evaluation; 1: FP64 (the reference's
 * arithmetic, not rounded); 2: the reference's FP32-rounded result.
 * Host pointers (synchronous); h receives the values in double. */
int a12_voigt_profile(a12_ctx *ctx, int32_t prec, int32_t n, const float *v,
                      const float *a, double *h);

#ifdef __cplusplus
}
#endif
#endif
