// a12_internal.h -- context and host-side plumbing shared by the HIP
// translation units of liba12gpu.so.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/a12gpu.h"

#define A12_NCLASS 16  // extent classes 0..14 plus "cut at every depth"
#define A12_NRMAX_LAYERS 128  // layers a context accepts (JOSH keeps them in LDS)

// Device buffer with grow-only reallocation.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    hipError_t reserve(size_t n)
    {
        if (n <= bytes) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        hipError_t e = hipMalloc(&p, n ? n : 1);
        if (e == hipSuccess) bytes = n;
        return e;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
};

struct SyntheDev {
    // dimensions
    int nrhox = 0, L = 0, nedge = 0, nf = 0, mw6 = 0, nlines = 0;
    double wbegin = 0, ratio = 0, resolu = 0;
    // atmosphere (device)
    DevBuf wledge, continall, contabs, contscat, xnfpel, dopple;
    DevBuf depth;  // 10 x nrhox doubles: t xne rho xnfh xnfhe1 xnfh2 hckt hkt hkt_atlas rhox
    // lines sorted by centre (SoA)
    DevBuf l_nbuff, l_nelion, l_cgf, l_elo, l_gr, l_gs, l_gw;
    // work arrays of the pipeline (per-depth prep, continuum, per (depth, line)
    // set-up, class partition, tile ranges, evaluation counts), one allocation
    DevBuf work;
    long long *evals_dev = nullptr;  // [all, Voigt core] of the last run, in work
    int shape = 0;                   // gather tile shape (a12_synthe_set_shape)
    uint64_t gen = 0;                // counts a12_synthe_load calls (a new atmosphere)
    // outputs
    DevBuf opacity;  // (nrhox x L) float or double
    int opac_prec = -1;
    int64_t evals = 0;
    bool loaded = false;
    bool prof_valid = false;  // ev[0..4] bracket the last opacity run
    // transfer
    DevBuf hline, hcont;
    bool transferred = false;
};

struct AtlasDev {
    int n = 0, numnu = 0, mion = 0, nl = 0, nwave1 = 0;
    double teff = 0;
    float lg2inv = 0;  // 1 / log2 of WAVESET's mean point-to-point ratio
    DevBuf depth3, depth_lo, tabcont, xnfdop, dopple, waveset, rcoset, tabs, tabs32;
    DevBuf l_nu, l_nucont, l_nel, l_const, seg;
    DevBuf state, par, reach, greach, xlines, blk, evals;
    DevBuf kapp, hnu1, acc;
    // the loaded store's grid (set_state checks a state against it) and the
    // host staging of set_state
    std::vector<double> waveset_h, stage_h;
    std::vector<int32_t> iwavetab_h;
    std::vector<double> t_h;    // T of the state set on it (a12_atlas_kapp_run checks its state)
    uint64_t grid_gen = 0;      // counts line-store loads (a new grid)
    size_t kapp_plane = 0;  // nrhox * numnu of the KAPP tables on the device
    // lines_ok: a store is loaded; loaded: a state is set on it
    bool lines_ok = false, loaded = false, xl_ready = false, kapp_ready = false, rt_done = false;
    bool prof_lo = false, prof_rt = false, prof_kapp = false;
    int shape = 0;  // LINOP1 gather tile shape (a12_atlas_set_shape / a12_atlas_autotune)
};

struct SelectDev {
    DevBuf tabcont, xnfdop, iwt, wavetab, tablog, xm, cand, flag, cnt, off, kept, mol2;
    DevBuf dia, dkey, dord, h2o, igr;   // diatomic records, keys, heap order; H2O; H2O widths
    bool tablog_ready = false;
};

struct MolDev {
    int nequa = 0, nummol = 0, nloc = 0;
    bool charge = false, ready = false;
    DevBuf tab, io;
};

// Model iteration (a12_model_converge): frequency table, per-model state
// and work arrays, sized for the largest batch run so far.
struct ModelDev {
    DevBuf ft, taustd, par, t, ot, rd, dep, part, acc, hnu, tin, tout, conv, hist, ctl;
};

// ROSSTAB's table (a12_rosstab_*): tab (3, A12_ROSSTAB_MAXTAB) TABT, TABP,
// ROSS and zp (4) ZEROT, SLOPET, ZEROP, SLOPEP on the device; nross on the
// host (every store's count is known there); io: staging of the TTAUP,
// CZ_CONSTRUCT and TCORR(3)-finish calls.
struct RossDev {
    DevBuf tab, zp, io;
    int nross = 0;
    double *pin = nullptr;  // pinned host staging of those calls (grow-only)
    size_t pin_bytes = 0;
};

struct MlpDev;    // a12_mlp.hip
struct KappDev;   // a12_kapp.hip
struct CplxDev;   // a12_cplx.hip
struct SynBatch;  // a12_synthe.hip

struct a12_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    DevBuf coefj, coefh;
    bool have_ops = false;
    SyntheDev syn;
    AtlasDev atl;
    SelectDev sel;
    MolDev mol;
    ModelDev mdl;
    RossDev ross;
    MlpDev *mlp = nullptr;
    KappDev *kapp = nullptr;
    CplxDev *cplx = nullptr;
    SynBatch *sbatch = nullptr;
    DevBuf rt_part;
    // scratch for a12_josh_batch and a12_solvit_batched
    DevBuf jb_in, jb_out, sv_a, sv_b;
    // stage timing (a12_set_profiling): events on the context stream
    bool prof = false;
    hipEvent_t ev[16] = {};  // 0-6 SYNTHE stages, 8-11 LINOP1 stages, 12-13 RT pass, 14-15 KAPP
};

#define A12_HIP(call)                                                              \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess) {                                                    \
            ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);          \
            return e_ == hipErrorOutOfMemory ? A12_ERR_OOM : A12_ERR_HIP;          \
        }                                                                          \
    } while (0)

#define A12_CHECK_LAUNCH()                                                         \
    do {                                                                           \
        hipError_t e_ = hipGetLastError();                                         \
        if (e_ != hipSuccess) {                                                    \
            ctx->err = std::string("kernel launch: ") + hipGetErrorString(e_);     \
            return A12_ERR_HIP;                                                    \
        }                                                                          \
    } while (0)

// Implemented in a12_synthe.hip
int a12i_synthe_load(a12_ctx *ctx, const a12_synthe_atm *atm, const a12_lte_line *lines,
                     int32_t nlines);
int a12i_synthe_run(a12_ctx *ctx, int prec);
int a12i_voigt(a12_ctx *ctx, int prec, int n, const float *v, const float *a, double *h);
int a12i_synthe_shapes(int idx, int32_t *pts_chunk);
int a12i_synthe_set_shape(a12_ctx *ctx, int shape);
int a12i_synthe_batch_load(a12_ctx *ctx, int32_t nmodel, const a12_synthe_atm *atms,
                           const int32_t *nb_shift, const int64_t *line_lo,
                           const int64_t *line_hi, const a12_lte_line *lines, int64_t nlines);
int a12i_synthe_batch_set_budget(a12_ctx *ctx, int64_t bytes);
int a12i_synthe_batch_run(a12_ctx *ctx, int prec, int transfer, int use_graph);
int a12i_synthe_batch_fetch(a12_ctx *ctx, int32_t model, void *opacity_out, double *hline,
                            double *hcont, int64_t *evals);
int a12i_synthe_batch_info(a12_ctx *ctx, int64_t *info);
void a12i_synthe_batch_free(a12_ctx *ctx);
// The per-depth prep and continuum of the loaded SYNTHE state (k_depth_prep,
// k_continuum) into caller arrays: xnfdop (nrhox, mw6), dscal (nrhox, 4),
// cont32 (nrhox, length).
int a12i_synthe_prep(a12_ctx *ctx, double *xnfdop, double *dscal, float *cont32);
// Implemented in a12_cplx.hip: the complex lines of the loaded SYNTHE state
// from the prep's xnfdop and cont32, added into the opacity (float, or
// double when f64), on the context stream.
bool a12i_cplx_active(const a12_ctx *ctx);
int a12i_cplx_add(a12_ctx *ctx, const double *xnfdop, const float *cont32, void *opacity, bool f64);
void a12i_cplx_free(a12_ctx *ctx);
// Implemented in a12_atlas.hip
int a12i_atlas_load(a12_ctx *ctx, const a12_atlas_atm *a, const int32_t *linedata, int32_t nlines);
int a12i_atlas_load_lines(a12_ctx *ctx, const a12_atlas_atm *a, const int32_t *linedata,
                          int32_t nlines);
int a12i_atlas_set_state(a12_ctx *ctx, const a12_atlas_atm *a);
int a12i_atlas_linop1_run(a12_ctx *ctx, int prec);
int a12i_atlas_shapes(int idx, int32_t *pts_block_batch_flat);
int a12i_atlas_autotune(a12_ctx *ctx, int prec, float *ms, int32_t *best);
// Implemented in a12_select.hip
int a12i_atlas_select_all(a12_ctx *ctx, const a12_atlas_atm *a, const double *wavetab,
                          const a12_sel_lists *lists, int32_t *out, int64_t cap, int64_t *nkept);
int a12i_atlas_select(a12_ctx *ctx, const a12_atlas_atm *a, const double *wavetab,
                      const int32_t *pred, int64_t npred, const a12_mol2_rec *mol2, int64_t nmol2,
                      int32_t *out, int64_t cap, int64_t *nkept);
// Implemented in a12_molec.hip
int a12i_molec_set_table(a12_ctx *ctx, const a12_moltab *tb);
int a12i_molec_newton(a12_ctx *ctx, int32_t nsys, int32_t nchain, const int32_t *chain_off,
                      const double *xntot, const double *pratio, const double *xab,
                      const double *equilj, double *xn, double *xnmol, int32_t *iters,
                      int32_t *status);
// Implemented in a12_mlp.hip
int a12i_mlp_init(a12_ctx *ctx, int batch, uint64_t seed);
int a12i_mlp_params(a12_ctx *ctx, float *get, const float *set);
int a12i_mlp_set_data(a12_ctx *ctx, const float *x, const float *t, int n);
int a12i_mlp_grads(a12_ctx *ctx, float *loss, float *grads);
int a12i_mlp_train(a12_ctx *ctx, int steps, float lr, float *loss);
int a12i_mlp_infer(a12_ctx *ctx, const float *x, int n, float *y);
int a12i_mlp_gemm_check(a12_ctx *ctx, int m, int n, int k, const uint16_t *a, const uint16_t *b,
                        float *c, int split);
void a12i_mlp_free(a12_ctx *ctx);
void a12i_kapp_free(a12_ctx *ctx);
// Implemented in a12_josh.hip
int a12i_atlas_rt_pass(a12_ctx *ctx, int n, int nf, double teff, const double *d_wave,
                       const double *d_rcoset, const double *d_depth3, const double *d_kapp,
                       const double *d_xlines, double *d_hnu1, double *d_acc);
// The RT pass with its chunk partials in the library's fixed device array
// (no allocation); a12i_rt_scratch_init fetches that array's address on
// the current device (once per device, from a12_ctx_create).
int a12i_rt_scratch_init(int device);
int a12i_atlas_rt_pass_noalloc(a12_ctx *ctx, int n, int nf, double teff, const double *d_wave,
                               const double *d_rcoset, const double *d_depth3,
                               const double *d_kapp, const double *d_xlines, double *d_hnu1,
                               double *d_acc);
int a12i_synthe_transfer(a12_ctx *ctx);
int a12i_synthe_transfer_launch(a12_ctx *ctx, hipStream_t st, bool f64, int L, int n,
                                double wbegin, double ratio, int nedge, const double *wledge,
                                int nf, const double *contabs, const double *contscat,
                                const double *depth, const void *opacity, double *hline,
                                double *hcont);
int a12i_atlas_tcorr3(a12_ctx *ctx, const a12_tcorr_in *in, double *rdabh, double *oldt1,
                      double *out);
int a12i_rosstab_reset(a12_ctx *ctx);
int a12i_rosstab_store(a12_ctx *ctx, int n, const double *t, const double *p,
                       const double *abross);
int a12i_rosstab_load(a12_ctx *ctx, int nross, const double *zero, const double *tab);
int a12i_rosstab_fetch(a12_ctx *ctx, int32_t *nross, double *zero, double *tab);
int a12i_rosstab_lookup(a12_ctx *ctx, int mode, int nq, const double *temp,
                        const double *pressure, double *kappa);
int a12i_ttaup(a12_ctx *ctx, int mode, int n, const double *t, const double *tau,
               const double *prad, const double *pturb, double grav, double *abstd,
               double *ptotal, double *p, int32_t *nfail);
int a12i_cz_construct(a12_ctx *ctx, const a12_czc_in *in, a12_czc_state *st, double *t1,
                      int32_t *engaged);
int a12i_tcorr3_finish(a12_ctx *ctx, const a12_tcorr_fin_in *in, a12_tcorr_fin_io *io,
                       a12_czc_state *st, double *t1, double *drhox, int32_t *nfail);
int a12i_convec(a12_ctx *ctx, const a12_convec_in *in, a12_convec_io *io, int32_t mlt_only);
int a12i_solvit_batched(a12_ctx *ctx, int n, int nsys, const double *a, const double *b,
                        double *x);
int a12i_model_converge(a12_ctx *ctx, int nmodel, const a12_model_par *par,
                        const double *t_start, int itmax, double *t_out, double *s_out,
                        int32_t *iters, double *hist);
int a12i_scale_model(a12_ctx *ctx, int nmodel, int n, int kw, double tau1lg, double steplg,
                     const double *sc, const double *fields, double *out, double *so,
                     double *taustd_out);
int a12i_josh_batch(a12_ctx *ctx, int ifscat, int ifsurf, int nrhox, int nbatch,
                    const double *rhox, const double *const in[7], double *hnu1,
                    double *knu1, double *const out[5]);
