// a12_capi.hip -- the extern "C" entry points declared in include/a12gpu.h.
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>

#include "a12_internal.h"

namespace {
bool ctx_ok(const a12_ctx *ctx) { return ctx != nullptr; }
}  // namespace

extern "C" {

int a12_ctx_create(int32_t device, a12_ctx **out)
{
    if (!out) return A12_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return A12_ERR_HIP;
    if (device < 0 || device >= ndev) return A12_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return A12_ERR_HIP;
    a12_ctx *c = new (std::nothrow) a12_ctx();
    if (!c) return A12_ERR_OOM;
    c->device = device;
    // the no-allocation RT pass's fixed scratch array: its address on this
    // device, fetched once per device with the device current
    if (a12i_rt_scratch_init(device) != A12_OK) {
        delete c;
        return A12_ERR_HIP;
    }
    // default priority stream
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return A12_ERR_HIP;
    }
    *out = c;
    return A12_OK;
}

int a12_ctx_destroy(a12_ctx *ctx)
{
    if (!ctx) return A12_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    SyntheDev &S = ctx->syn;
    DevBuf *bufs[] = {&S.wledge, &S.continall, &S.contabs, &S.contscat, &S.xnfpel, &S.dopple,
                      &S.depth, &S.l_nbuff, &S.l_nelion, &S.l_cgf, &S.l_elo, &S.l_gr, &S.l_gs,
                      &S.l_gw, &S.work, &S.opacity, &S.hline, &S.hcont, &ctx->coefj, &ctx->coefh,
                      &ctx->jb_in, &ctx->jb_out, &ctx->sv_a, &ctx->sv_b, &ctx->rt_part};
    AtlasDev &A = ctx->atl;
    DevBuf *abufs[] = {&A.depth3, &A.depth_lo, &A.tabcont, &A.xnfdop, &A.dopple, &A.waveset,
                       &A.rcoset, &A.tabs, &A.tabs32, &A.l_nu, &A.l_nucont, &A.l_nel, &A.l_const,
                       &A.seg, &A.state, &A.par, &A.reach, &A.xlines, &A.blk, &A.evals, &A.kapp,
                       &A.hnu1, &A.acc};
    SelectDev &L = ctx->sel;
    DevBuf *sbufs[] = {&L.tabcont, &L.xnfdop, &L.iwt, &L.wavetab, &L.tablog, &L.xm,
                       &L.cand, &L.flag, &L.cnt, &L.off, &L.kept, &L.mol2};
    for (DevBuf *b : sbufs) b->release();
    ModelDev &M = ctx->mdl;
    DevBuf *mbufs[] = {&M.ft, &M.taustd, &M.par, &M.t, &M.ot, &M.rd, &M.dep, &M.part,
                       &M.acc, &M.hnu, &M.tin, &M.tout, &M.conv, &M.hist, &M.ctl};
    for (DevBuf *b : mbufs) b->release();
    ctx->mol.tab.release();
    ctx->mol.io.release();
    ctx->ross.tab.release();
    ctx->ross.zp.release();
    ctx->ross.io.release();
    if (ctx->ross.pin) (void)hipHostFree(ctx->ross.pin);
    for (DevBuf *b : abufs) b->release();
    for (DevBuf *b : bufs) b->release();
    a12i_mlp_free(ctx);
    a12i_kapp_free(ctx);
    a12i_cplx_free(ctx);
    a12i_synthe_batch_free(ctx);
    for (hipEvent_t e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return A12_OK;
}

const char *a12_last_error(const a12_ctx *ctx)
{
    if (!ctx) return "null context";
    return ctx->err.c_str();
}

void *a12_stream(a12_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int a12_set_operators(a12_ctx *ctx, const double *coefj, const double *coefh)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    if (!coefj || !coefh) {
        ctx->err = "a12_set_operators: null matrix";
        return A12_ERR_ARG;
    }
    A12_HIP(hipSetDevice(ctx->device));
    A12_HIP(ctx->coefj.reserve(sizeof(double) * 51 * 51));
    A12_HIP(ctx->coefh.reserve(sizeof(double) * 51 * 51));
    A12_HIP(hipMemcpyAsync(ctx->coefj.p, coefj, sizeof(double) * 51 * 51, hipMemcpyHostToDevice,
                           ctx->stream));
    A12_HIP(hipMemcpyAsync(ctx->coefh.p, coefh, sizeof(double) * 51 * 51, hipMemcpyHostToDevice,
                           ctx->stream));
    A12_HIP(hipStreamSynchronize(ctx->stream));
    ctx->have_ops = true;
    return A12_OK;
}

int a12_synthe_load(a12_ctx *ctx, const a12_synthe_atm *atm, const a12_lte_line *lines,
                    int32_t nlines)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_synthe_load(ctx, atm, lines, nlines);
}

int a12_synthe_opacity_run(a12_ctx *ctx, int32_t prec)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_synthe_run(ctx, prec);
}

int a12_synthe_opacity_fetch(a12_ctx *ctx, void *opacity_out)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    SyntheDev &S = ctx->syn;
    if (S.opac_prec < 0) {
        ctx->err = "a12_synthe_opacity_fetch: nothing computed";
        return A12_ERR_STATE;
    }
    if (!opacity_out) {
        ctx->err = "a12_synthe_opacity_fetch: null output";
        return A12_ERR_ARG;
    }
    const size_t rb = S.opac_prec == A12_PREC_FP64 ? sizeof(double) : sizeof(float);
    A12_HIP(hipMemcpyAsync(opacity_out, S.opacity.p, rb * (size_t)S.L * S.nrhox,
                           hipMemcpyDeviceToHost, ctx->stream));
    long long ev = 0;
    A12_HIP(hipMemcpyAsync(&ev, S.evals_dev, sizeof(long long), hipMemcpyDeviceToHost,
                           ctx->stream));
    A12_HIP(hipStreamSynchronize(ctx->stream));
    S.evals = ev;
    return A12_OK;
}

const void *a12_synthe_opacity_device(a12_ctx *ctx, int32_t *prec_out)
{
    if (!ctx) return nullptr;
    if (prec_out) *prec_out = ctx->syn.opac_prec;
    return ctx->syn.opacity.p;
}

int64_t a12_synthe_evals(a12_ctx *ctx)
{
    if (!ctx || ctx->syn.opac_prec < 0) return -1;
    long long ev = 0;
    if (hipMemcpyAsync(&ev, ctx->syn.evals_dev, sizeof(long long), hipMemcpyDeviceToHost,
                       ctx->stream) != hipSuccess)
        return -1;
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return -1;
    ctx->syn.evals = ev;
    return ev;
}

int a12_synthe_line_opacity(a12_ctx *ctx, const a12_synthe_atm *atm,
                            const a12_lte_line *lines, int32_t nlines, int32_t prec,
                            void *opacity_out, int64_t *evals_out)
{
    int rc = a12_synthe_load(ctx, atm, lines, nlines);
    if (rc) return rc;
    rc = a12_synthe_opacity_run(ctx, prec);
    if (rc) return rc;
    rc = a12_synthe_opacity_fetch(ctx, opacity_out);
    if (rc) return rc;
    if (evals_out) *evals_out = ctx->syn.evals;
    return A12_OK;
}

int a12_synthe_transfer_run(a12_ctx *ctx)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_synthe_transfer(ctx);
}

int a12_synthe_transfer_fetch(a12_ctx *ctx, double *hnu_line, double *hnu_cont)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    SyntheDev &S = ctx->syn;
    if (!S.transferred) {
        ctx->err = "a12_synthe_transfer_fetch: no transfer result";
        return A12_ERR_STATE;
    }
    if (!hnu_line || !hnu_cont) {
        ctx->err = "a12_synthe_transfer_fetch: null output";
        return A12_ERR_ARG;
    }
    A12_HIP(hipMemcpyAsync(hnu_line, S.hline.p, sizeof(double) * S.L, hipMemcpyDeviceToHost,
                           ctx->stream));
    A12_HIP(hipMemcpyAsync(hnu_cont, S.hcont.p, sizeof(double) * S.L, hipMemcpyDeviceToHost,
                           ctx->stream));
    A12_HIP(hipStreamSynchronize(ctx->stream));
    return A12_OK;
}

int a12_synthe_transfer(a12_ctx *ctx, const a12_synthe_atm *atm, const void *opacity,
                        int32_t prec, double *hnu_line, double *hnu_cont)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    if (!opacity || (prec != A12_PREC_MIXED && prec != A12_PREC_FP64)) {
        ctx->err = "a12_synthe_transfer: null opacity or bad prec";
        return A12_ERR_ARG;
    }
    int rc = a12_synthe_load(ctx, atm, nullptr, 0);
    if (rc) return rc;
    SyntheDev &S = ctx->syn;
    const size_t rb = prec == A12_PREC_FP64 ? sizeof(double) : sizeof(float);
    A12_HIP(S.opacity.reserve(rb * (size_t)S.L * S.nrhox));
    A12_HIP(hipMemcpyAsync(S.opacity.p, opacity, rb * (size_t)S.L * S.nrhox,
                           hipMemcpyHostToDevice, ctx->stream));
    S.opac_prec = prec;
    rc = a12_synthe_transfer_run(ctx);
    if (rc) return rc;
    return a12_synthe_transfer_fetch(ctx, hnu_line, hnu_cont);
}

int a12_josh_batch(a12_ctx *ctx, int32_t ifscat, int32_t ifsurf, int32_t nrhox, int32_t nbatch,
                   const double *rhox, const double *acont, const double *aline,
                   const double *sigmac, const double *sigmal, const double *scont,
                   const double *sline, const double *bnu, double *hnu1, double *knu1,
                   double *taunu, double *snu, double *jnu, double *hnu, double *jmins)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    if (nrhox < 4 || nrhox > A12_NRMAX_LAYERS || nbatch < 0 || (ifscat != 0 && ifscat != 1) ||
        (ifsurf != 0 && ifsurf != 1) || !rhox || !acont || !aline || !sigmac || !sigmal ||
        !scont || !sline || !bnu) {
        ctx->err = "a12_josh_batch: bad arguments";
        return A12_ERR_ARG;
    }
    if (nbatch == 0) return A12_OK;
    A12_HIP(hipSetDevice(ctx->device));
    const double *in[7] = {acont, aline, sigmac, sigmal, scont, sline, bnu};
    double *out[5] = {taunu, snu, jnu, hnu, jmins};
    return a12i_josh_batch(ctx, ifscat, ifsurf, nrhox, nbatch, rhox, in, hnu1, knu1, out);
}

int a12_solvit_batched(a12_ctx *ctx, int32_t n, int32_t nsys, const double *a, const double *b,
                       double *x)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    if (n < 1 || n > 64 || nsys < 0 || (nsys > 0 && (!a || !b || !x))) {
        ctx->err = "a12_solvit_batched: bad arguments (1 <= n <= 64, nsys >= 0)";
        return A12_ERR_ARG;
    }
    if (nsys == 0) return A12_OK;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_solvit_batched(ctx, n, nsys, a, b, x);
}

int a12_atlas_load(a12_ctx *ctx, const a12_atlas_atm *atm, const int32_t *linedata,
                   int32_t nlines)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_atlas_load(ctx, atm, linedata, nlines);
}

int a12_atlas_load_lines(a12_ctx *ctx, const a12_atlas_atm *atm, const int32_t *linedata,
                         int32_t nlines)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_atlas_load_lines(ctx, atm, linedata, nlines);
}

int a12_atlas_set_state(a12_ctx *ctx, const a12_atlas_atm *atm)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_atlas_set_state(ctx, atm);
}

int a12_atlas_select_lines(a12_ctx *ctx, const a12_atlas_atm *atm, const double *wavetab,
                           const int32_t *pred, int64_t npred, const a12_mol2_rec *mol2,
                           int64_t nmol2, int32_t *linedata_out, int64_t cap, int64_t *nkept)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_atlas_select(ctx, atm, wavetab, pred, npred, mol2, nmol2, linedata_out, cap,
                             nkept);
}

int a12_atlas_select_all(a12_ctx *ctx, const a12_atlas_atm *atm, const double *wavetab,
                         const a12_sel_lists *lists, int32_t *linedata_out, int64_t cap,
                         int64_t *nkept)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_atlas_select_all(ctx, atm, wavetab, lists, linedata_out, cap, nkept);
}

int a12_atlas_linop1_run(a12_ctx *ctx, int32_t prec)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_atlas_linop1_run(ctx, prec);
}

int a12_atlas_shape_info(int32_t idx, int32_t *pts_block_batch)
{
    return a12i_atlas_shapes(idx, pts_block_batch);
}

int a12_atlas_set_shape(a12_ctx *ctx, int32_t shape)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    if (shape < 0 || shape >= a12i_atlas_shapes(-1, nullptr)) {
        ctx->err = "a12_atlas_set_shape: no such shape";
        return A12_ERR_ARG;
    }
    ctx->atl.shape = shape;
    return A12_OK;
}

int a12_atlas_autotune(a12_ctx *ctx, int32_t prec, float *ms, int32_t *best)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_atlas_autotune(ctx, prec, ms, best);
}

int a12_atlas_linop1_fetch(a12_ctx *ctx, double *xlines, int64_t *evals)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    AtlasDev &A = ctx->atl;
    if (!A.xl_ready) {
        ctx->err = "a12_atlas_linop1_fetch: no XLINES computed";
        return A12_ERR_STATE;
    }
    if (xlines)
        A12_HIP(hipMemcpyAsync(xlines, A.xlines.p, sizeof(double) * (size_t)A.n * A.numnu,
                               hipMemcpyDeviceToHost, ctx->stream));
    long long ev = 0;
    A12_HIP(hipMemcpyAsync(&ev, A.evals.p, sizeof(ev), hipMemcpyDeviceToHost, ctx->stream));
    A12_HIP(hipStreamSynchronize(ctx->stream));
    if (evals) *evals = ev;
    return A12_OK;
}

int a12_atlas_set_kapp(a12_ctx *ctx, const double *acont, const double *sigmac,
                       const double *scont, const double *sline, const double *aline,
                       const double *sigmal)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    AtlasDev &A = ctx->atl;
    if (!A.loaded) {
        ctx->err = "a12_atlas_set_kapp: call a12_atlas_set_state (or a12_atlas_load) first";
        return A12_ERR_STATE;
    }
    const double *src[6] = {acont, sigmac, scont, sline, aline, sigmal};
    for (const double *p : src)
        if (!p) {
            ctx->err = "a12_atlas_set_kapp: null table";
            return A12_ERR_ARG;
        }
    A12_HIP(hipSetDevice(ctx->device));
    const size_t plane = (size_t)A.n * A.numnu;
    A12_HIP(A.kapp.reserve(sizeof(double) * 6 * plane));
    for (int k = 0; k < 6; ++k)
        A12_HIP(hipMemcpyAsync(A.kapp.as<double>() + k * plane, src[k], sizeof(double) * plane,
                               hipMemcpyHostToDevice, ctx->stream));
    A12_HIP(hipStreamSynchronize(ctx->stream));
    A.kapp_ready = true;
    A.kapp_plane = plane;
    return A12_OK;
}

int a12_atlas_rt_run(a12_ctx *ctx, const double *xlines)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    AtlasDev &A = ctx->atl;
    A12_HIP(hipSetDevice(ctx->device));
    if (!A.kapp_ready) {
        ctx->err = "a12_atlas_rt_run: call a12_atlas_set_kapp first";
        return A12_ERR_STATE;
    }
    const size_t plane = (size_t)A.n * A.numnu;
    if (A.kapp_plane != plane) {
        ctx->err = "a12_atlas_rt_run: KAPP tables were set for another grid";
        return A12_ERR_STATE;
    }
    if (xlines) {
        A12_HIP(A.xlines.reserve(sizeof(double) * plane));
        A12_HIP(hipMemcpyAsync(A.xlines.p, xlines, sizeof(double) * plane, hipMemcpyHostToDevice,
                               ctx->stream));
        A.xl_ready = true;
    }
    if (!A.xl_ready) {
        ctx->err = "a12_atlas_rt_run: no XLINES (run linop1 or pass xlines)";
        return A12_ERR_STATE;
    }
    A12_HIP(A.hnu1.reserve(sizeof(double) * A.numnu));
    A12_HIP(A.acc.reserve(sizeof(double) * 9 * A.n));
    int rc = a12i_atlas_rt_pass(ctx, A.n, A.numnu, A.teff, A.waveset.as<double>(),
                                A.rcoset.as<double>(), A.depth3.as<double>(), A.kapp.as<double>(),
                                A.xlines.as<double>(), A.hnu1.as<double>(), A.acc.as<double>());
    if (rc) return rc;
    A.rt_done = true;
    A.prof_rt = ctx->prof;
    return A12_OK;
}

int a12_atlas_rt_fetch(a12_ctx *ctx, double *acc, double *hnu1)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    AtlasDev &A = ctx->atl;
    if (!A.rt_done) {
        ctx->err = "a12_atlas_rt_fetch: no RT pass computed";
        return A12_ERR_STATE;
    }
    if (acc)
        A12_HIP(hipMemcpyAsync(acc, A.acc.p, sizeof(double) * 9 * A.n, hipMemcpyDeviceToHost,
                               ctx->stream));
    if (hnu1)
        A12_HIP(hipMemcpyAsync(hnu1, A.hnu1.p, sizeof(double) * A.numnu, hipMemcpyDeviceToHost,
                               ctx->stream));
    A12_HIP(hipStreamSynchronize(ctx->stream));
    return A12_OK;
}

int a12_atlas_rt_pass_noalloc(a12_ctx *ctx, int32_t nrhox, int32_t numnu, double teff,
                              const double *d_waveset, const double *d_rcoset,
                              const double *d_depth3, const double *d_kapp,
                              const double *d_xlines, double *d_hnu1, double *d_acc)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_atlas_rt_pass_noalloc(ctx, nrhox, numnu, teff, d_waveset, d_rcoset, d_depth3,
                                      d_kapp, d_xlines, d_hnu1, d_acc);
}

int a12_model_converge(a12_ctx *ctx, int32_t nmodel, const a12_model_par *par,
                       const double *t_start, int32_t itmax, double *t_out, double *s_out,
                       int32_t *iters, double *hist)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_model_converge(ctx, nmodel, par, t_start, itmax, t_out, s_out, iters, hist);
}

int a12_scale_model(a12_ctx *ctx, int32_t nmodel, int32_t nrhox, int32_t kw, double tau1lg,
                    double steplg, const double *sc, const double *fields, double *out,
                    double *so, double *taustd)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_scale_model(ctx, nmodel, nrhox, kw, tau1lg, steplg, sc, fields, out, so, taustd);
}

int a12_mlp_init(a12_ctx *ctx, int32_t batch, uint64_t seed)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_mlp_init(ctx, batch, seed);
}

int a12_mlp_get_params(a12_ctx *ctx, float *params)
{
    if (!ctx_ok(ctx) || !params) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_mlp_params(ctx, params, nullptr);
}

int a12_mlp_set_params(a12_ctx *ctx, const float *params)
{
    if (!ctx_ok(ctx) || !params) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_mlp_params(ctx, nullptr, params);
}

int a12_mlp_set_data(a12_ctx *ctx, const float *x, const float *t, int32_t n)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_mlp_set_data(ctx, x, t, n);
}

int a12_mlp_grads(a12_ctx *ctx, float *loss, float *grads)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_mlp_grads(ctx, loss, grads);
}

int a12_mlp_train(a12_ctx *ctx, int32_t steps, float lr, float *loss)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_mlp_train(ctx, steps, lr, loss);
}

int a12_mlp_infer(a12_ctx *ctx, const float *x, int32_t n, float *y)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_mlp_infer(ctx, x, n, y);
}

int a12_mlp_gemm(a12_ctx *ctx, int32_t m, int32_t n, int32_t k, const uint16_t *a,
                 const uint16_t *b, float *c, int32_t split)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_mlp_gemm_check(ctx, m, n, k, a, b, c, split);
}

int a12_atlas_tcorr3(a12_ctx *ctx, const a12_tcorr_in *in, double *rdabh, double *oldt1,
                     double *out)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_atlas_tcorr3(ctx, in, rdabh, oldt1, out);
}

int a12_rosstab_reset(a12_ctx *ctx)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    return a12i_rosstab_reset(ctx);
}

int a12_rosstab_store(a12_ctx *ctx, int32_t nrhox, const double *t, const double *p,
                      const double *abross)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_rosstab_store(ctx, nrhox, t, p, abross);
}

int a12_rosstab_load(a12_ctx *ctx, int32_t nross, const double *zero, const double *tab)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_rosstab_load(ctx, nross, zero, tab);
}

int a12_rosstab_fetch(a12_ctx *ctx, int32_t *nross, double *zero, double *tab)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_rosstab_fetch(ctx, nross, zero, tab);
}

int a12_rosstab_lookup(a12_ctx *ctx, int32_t mode, int32_t nq, const double *temp,
                       const double *pressure, double *kappa)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_rosstab_lookup(ctx, mode, nq, temp, pressure, kappa);
}

int a12_ttaup(a12_ctx *ctx, int32_t irosstab, int32_t n, const double *t, const double *tau,
              const double *prad, const double *pturb, double grav, double *abstd,
              double *ptotal, double *p, int32_t *nfail)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_ttaup(ctx, irosstab, n, t, tau, prad, pturb, grav, abstd, ptotal, p, nfail);
}

int a12_cz_construct(a12_ctx *ctx, const a12_czc_in *in, a12_czc_state *st, double *t1,
                     int32_t *engaged)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_cz_construct(ctx, in, st, t1, engaged);
}

int a12_atlas_tcorr3_finish(a12_ctx *ctx, const a12_tcorr_fin_in *in, a12_tcorr_fin_io *io,
                            a12_czc_state *st, double *t1, double *drhox, int32_t *nfail)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_tcorr3_finish(ctx, in, io, st, t1, drhox, nfail);
}

int a12_convec(a12_ctx *ctx, const a12_convec_in *in, a12_convec_io *io, int32_t mlt_only)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_convec(ctx, in, io, mlt_only);
}

int a12_molec_set_table(a12_ctx *ctx, const a12_moltab *tab)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_molec_set_table(ctx, tab);
}

int a12_molec_newton_batched(a12_ctx *ctx, int32_t nsys, int32_t nchain,
                             const int32_t *chain_off, const double *xntot, const double *pratio,
                             const double *xab, const double *equilj, double *xn, double *xnmol,
                             int32_t *iters, int32_t *status)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_molec_newton(ctx, nsys, nchain, chain_off, xntot, pratio, xab, equilj, xn, xnmol,
                             iters, status);
}

int a12_synthe_shape_info(int32_t idx, int32_t pts_chunk[2])
{
    return a12i_synthe_shapes(idx, pts_chunk);
}

int a12_synthe_set_shape(a12_ctx *ctx, int32_t shape)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    return a12i_synthe_set_shape(ctx, shape);
}

int a12_synthe_batch_load(a12_ctx *ctx, int32_t nmodel, const a12_synthe_atm *atms,
                          const int32_t *nb_shift, const int64_t *line_lo,
                          const int64_t *line_hi, const a12_lte_line *lines, int64_t nlines)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_synthe_batch_load(ctx, nmodel, atms, nb_shift, line_lo, line_hi, lines, nlines);
}

int a12_synthe_batch_set_budget(a12_ctx *ctx, int64_t bytes)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    return a12i_synthe_batch_set_budget(ctx, bytes);
}

int a12_synthe_batch_run(a12_ctx *ctx, int32_t prec, int32_t transfer, int32_t use_graph)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_synthe_batch_run(ctx, prec, transfer, use_graph);
}

int a12_synthe_batch_fetch(a12_ctx *ctx, int32_t model, void *opacity_out, double *hnu_line,
                           double *hnu_cont, int64_t evals[2])
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_synthe_batch_fetch(ctx, model, opacity_out, hnu_line, hnu_cont, evals);
}

int a12_synthe_batch_info(a12_ctx *ctx, int64_t info[5])
{
    if (!ctx_ok(ctx) || !info) return A12_ERR_ARG;
    return a12i_synthe_batch_info(ctx, info);
}

int a12_set_profiling(a12_ctx *ctx, int32_t on)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    if (on && !ctx->ev[0])
        for (hipEvent_t &e : ctx->ev) A12_HIP(hipEventCreate(&e));
    ctx->prof = on != 0;
    return A12_OK;
}

int a12_stage_ms(a12_ctx *ctx, float *ms)
{
    if (!ctx_ok(ctx) || !ms) return A12_ERR_ARG;
    for (int i = 0; i < 5; ++i) ms[i] = -1.0f;
    if (ctx->syn.prof_valid) {
        A12_HIP(hipEventSynchronize(ctx->ev[4]));
        A12_HIP(hipEventElapsedTime(&ms[0], ctx->ev[0], ctx->ev[1]));
        A12_HIP(hipEventElapsedTime(&ms[1], ctx->ev[1], ctx->ev[2]));
        A12_HIP(hipEventElapsedTime(&ms[2], ctx->ev[2], ctx->ev[3]));
        A12_HIP(hipEventElapsedTime(&ms[3], ctx->ev[3], ctx->ev[4]));
    }
    if (ctx->prof && ctx->syn.transferred) {
        A12_HIP(hipEventSynchronize(ctx->ev[6]));
        A12_HIP(hipEventElapsedTime(&ms[4], ctx->ev[5], ctx->ev[6]));
    }
    return A12_OK;
}

int a12_atlas_stage_ms(a12_ctx *ctx, float *ms)
{
    if (!ctx_ok(ctx) || !ms) return A12_ERR_ARG;
    for (int i = 0; i < 4; ++i) ms[i] = -1.0f;
    if (ctx->atl.prof_lo && ctx->atl.xl_ready) {
        A12_HIP(hipEventSynchronize(ctx->ev[11]));
        A12_HIP(hipEventElapsedTime(&ms[0], ctx->ev[8], ctx->ev[9]));
        A12_HIP(hipEventElapsedTime(&ms[1], ctx->ev[9], ctx->ev[10]));
        A12_HIP(hipEventElapsedTime(&ms[2], ctx->ev[10], ctx->ev[11]));
    }
    if (ctx->atl.prof_rt && ctx->atl.rt_done) {
        A12_HIP(hipEventSynchronize(ctx->ev[13]));
        A12_HIP(hipEventElapsedTime(&ms[3], ctx->ev[12], ctx->ev[13]));
    }
    return A12_OK;
}

int a12_synthe_eval_split(a12_ctx *ctx, int64_t *total, int64_t *core)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    if (ctx->syn.opac_prec < 0) {
        ctx->err = "a12_synthe_eval_split: nothing computed";
        return A12_ERR_STATE;
    }
    long long ev[2] = {0, 0};
    A12_HIP(hipMemcpyAsync(ev, ctx->syn.evals_dev, sizeof(ev), hipMemcpyDeviceToHost,
                           ctx->stream));
    A12_HIP(hipStreamSynchronize(ctx->stream));
    if (total) *total = ev[0];
    if (core) *core = ev[1];
    return A12_OK;
}

int a12_voigt_profile(a12_ctx *ctx, int32_t prec, int32_t n, const float *v, const float *a,
                      double *h)
{
    if (!ctx_ok(ctx)) return A12_ERR_ARG;
    A12_HIP(hipSetDevice(ctx->device));
    return a12i_voigt(ctx, prec, n, v, a, h);
}

}  // extern "C"
