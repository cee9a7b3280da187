// a12_cplx.hip -- SYNTHE's complex-profile lines on the device:
// compute_line_opacity (src/synthe_module.f90:3167-3702) with
// hydrogen_line_profile (:1032-1980) for every depth of the loaded SYNTHE
// atmosphere.  The arithmetic is a12_cplx_core.h; the kernels:
//   k_cl_chain   (at load) per (line, wing): the running wavelength (or
//                frequency) product of the wing walk, kept every CK points
//   k_cl_plan    per (line, depth): the set-up before the wing loops
//   k_cl_owner   per depth: which earlier line's Stehle convolution each H
//                line reads (the reference's per-transition cache)
//   k_cl_conv    per (H line, depth) that builds one: one wavefront builds
//                the Stehle (x) Voigt convolution in LDS
//   k_cl_extent  per (line, depth): one wavefront walks both wings 64 points
//                at a time and finds where each walk stops
//   k_cl_accum   per (depth, point): the sum over lines in line order of
//                the kappas that reach the point (FP32, the reference's
//                order of additions into BUFFER)
//   k_cl_add     opacity += buffer, after SYNTHE's LTE gather
#pragma clang fp contract(off)

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "a12_internal.h"
#include "a12_cplx_core.h"

using namespace cplx;

struct CplxDev {
    int nlines = 0, nh = 0, n = 0, L = 0, mw6 = 0;
    uint64_t syn_gen = 0;
    Grid G{};
    Stark S{};  // device pointers
    DevBuf lines, hl, hline_of, hidx, depth, sd, si, plan, owner, conv, ck, ckoff, ext, buf, prep;
    bool loaded = false, ran = false, prof_valid = false;
    hipEvent_t ev[5] = {};
};

namespace {

constexpr int kWave = 64;

CplxDev &cdev(a12_ctx *ctx)
{
    if (!ctx->cplx) ctx->cplx = new CplxDev();
    return *ctx->cplx;
}

template <class T> hipError_t upload(DevBuf &b, const T *src, size_t n, hipStream_t st)
{
    hipError_t e = b.reserve(sizeof(T) * std::max<size_t>(n, 1));
    if (e != hipSuccess || n == 0) return e;
    return hipMemcpyAsync(b.p, src, sizeof(T) * n, hipMemcpyHostToDevice, st);
}

__device__ __forceinline__ double chain_val(const Grid &G, const double *ck, int aut, int red, int k)
{
    double v = ck[k / CK];
    for (int i = k % CK; i > 0; --i) v = chain_next(G, aut, red, v);
    return v;
}

__global__ void k_cl_chain(int nlines, Grid G, const a12_cplx_line *__restrict__ lines,
                           const int *__restrict__ ckoff, double *__restrict__ ck)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int l = t >> 1, red = (t & 1) == 0;
    if (l >= nlines) return;
    const int off = ckoff[l * 4 + (red ? 0 : 2)], cnt = ckoff[l * 4 + (red ? 1 : 3)];
    if (cnt <= 0) return;
    const int nb = lines[l].nbuff;
    const int minred = nb > 1 ? nb : 1, ib0 = nb < G.length + 1 ? nb : G.length + 1;
    const int aut = lines[l].itype == 1;
    double v = chain_start(G, aut, red, minred, ib0);
    for (int k = 0; k < cnt; ++k) {
        if (k % CK == 0) ck[off + k / CK] = v;
        v = chain_next(G, aut, red, v);
    }
}

__global__ void k_cl_plan(int nlines, int mw6, Grid G, const a12_cplx_line *__restrict__ lines,
                          const Depth *__restrict__ depth, const double *__restrict__ xnfdop,
                          const double *__restrict__ dopple, const double *__restrict__ xnfpel,
                          const float *__restrict__ cont32, Plan *__restrict__ plan)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (l >= nlines) return;
    const a12_cplx_line ln = lines[l];
    const Depth D = depth[j];
    const double xp =
        (ln.nelion >= 1 && ln.nelion <= mw6) ? xnfpel_of(xnfpel[(size_t)j * mw6 + ln.nelion - 1], D.rho) : 0.0;
    Plan P;
    plan_line(ln, G, D, xnfdop + (size_t)j * mw6, dopple + (size_t)j * mw6, xp, cont32 + (size_t)j * G.length, P);
    plan[(size_t)j * nlines + l] = P;
}

__global__ void k_cl_owner(int nlines, int n, Grid G, const a12_cplx_line *__restrict__ lines,
                           const Plan *__restrict__ plan, int *__restrict__ owner)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int last = -1;
    for (int l = 0; l < nlines; ++l) {
        const size_t q = (size_t)j * nlines + l;
        owner[q] = l;
        const Plan &P = plan[q];
        if (P.kind != K_HISO && P.kind != K_HBETA && P.kind != K_HMERGE) continue;
        const double wr0 = chain_start(G, 0, 1, P.minred, P.ib0);
        const double wb0 = chain_start(G, 0, 0, P.minred, P.ib0);
        if (!calls_profile(P, G, wr0, wb0)) continue;
        if (last >= 0 && lines[last].nblo == lines[l].nblo && lines[last].nbup == lines[l].nbup)
            owner[q] = owner[(size_t)j * nlines + last];
        last = l;
    }
}

__global__ __launch_bounds__(kWave) void k_cl_conv(int nlines, int nh, Stark S, const int *__restrict__ hline_of,
                                                   const HLine *__restrict__ hl, const Depth *__restrict__ depth,
                                                   const Plan *__restrict__ plan, const int *__restrict__ owner,
                                                   HConv *__restrict__ conv)
{
    const int h = blockIdx.x, j = blockIdx.y, tid = threadIdx.x;
    const int l = hline_of[h];
    const size_t q = (size_t)j * nlines + l;
    const Plan &P = plan[q];
    if (P.kind != K_HISO && P.kind != K_HBETA && P.kind != K_HMERGE) return;
    if (owner[q] != l) return;
    const HLine &H = hl[h];
    const Depth &D = depth[j];
    if (!use_stehle(S, H.n, H.m, D.xne)) return;
    __shared__ ConvSetup C;
    __shared__ double da[NCONV_MAX], pr[NCONV_MAX], st[NCONV_MAX];
    if (tid == 0) conv_params(H, D, S, P.hwdop, C);
    __syncthreads();
    for (int ia = tid; ia < NDA; ia += kWave) conv_fill(C, ia);
    __syncthreads();
    if (tid == 0) conv_sizes(C);
    __syncthreads();
    const int ic = C.ic;
    for (int c = tid; c < ic; c += kWave) conv_grid_point(C, c, da[c], pr[c]);
    __syncthreads();
    for (int c = tid; c < ic; c += kWave) st[c] = conv_step(da, c, ic);
    __syncthreads();
    HConv &cv = conv[(size_t)j * nh + h];
    if (tid < NDA) {
        const double v = conv_output(C, da, pr, st, tid);
        cv.prof[tid] = v;
        cv.lprof[tid] = v > 0.0 ? log(v) : 0.0;
    }
    if (tid == 0) {
        cv.F0 = C.F0;
        cv.lambda0 = C.lambda0;
        cv.dop_dalpha = C.dop_dalpha;
        cv.voigt_a = C.voigt_a;
        cv.pad = 0;
    }
}

// The line's profile inputs at depth j: H line constants, Stehle flag and
// the convolution its cache holds.
struct LineCtx {
    const HLine *H;
    const HConv *cv;
    bool stehle;
};
__device__ __forceinline__ LineCtx line_ctx(const Stark &S, int nlines, int nh, int j, int l, const Plan &P,
                                            const int *hidx, const HLine *hl, const Depth &D, const int *owner,
                                            const HConv *conv)
{
    LineCtx c{nullptr, nullptr, false};
    if (P.kind != K_HISO && P.kind != K_HBETA && P.kind != K_HMERGE) return c;
    const int h = hidx[l];
    c.H = &hl[h];
    c.stehle = use_stehle(S, c.H->n, c.H->m, D.xne);
    if (c.stehle) c.cv = &conv[(size_t)j * nh + hidx[owner[(size_t)j * nlines + l]]];
    return c;
}

__global__ __launch_bounds__(kWave) void k_cl_extent(int nlines, int nh, Grid G, Stark S,
                                                     const a12_cplx_line *__restrict__ lines,
                                                     const int *__restrict__ hidx, const HLine *__restrict__ hl,
                                                     const Depth *__restrict__ depth, const Plan *__restrict__ plan,
                                                     const int *__restrict__ owner, const HConv *__restrict__ conv,
                                                     const float *__restrict__ cont32, const int *__restrict__ ckoff,
                                                     const double *__restrict__ ck, int4 *__restrict__ ext)
{
    const int l = blockIdx.x, j = blockIdx.y, lane = threadIdx.x;
    const int L = G.length;
    const size_t q = (size_t)j * nlines + l;
    const Plan P = plan[q];
    int4 e = make_int4(1, 0, 1, 0);
    if (P.kind == K_CONT) {
        e.x = P.nb1;
        e.y = P.nb3 < L ? P.nb3 : L;
    }
    if (P.kind == K_SKIP || P.kind == K_CONT) {
        if (lane == 0) ext[q] = e;
        return;
    }
    const a12_cplx_line ln = lines[l];
    const Depth D = depth[j];
    const LineCtx lc = line_ctx(S, nlines, nh, j, l, P, hidx, hl, D, owner, conv);
    const float *cont = cont32 + (size_t)j * L;
    const int aut = P.kind == K_AUT;
    int rlo = INT_MAX, rhi = INT_MIN, blo = INT_MAX, bhi = INT_MIN;
    int stopped = 0;
    if (P.red && P.minred <= L) {
        const int nred = L - P.minred + 1;
        const double *ckr = ck + ckoff[l * 4 + 0];
        const double thr0 = (double)cont[P.minred - 1] * G.cutoff;
        for (int k0 = 0; k0 < nred; k0 += kWave) {
            const int k = k0 + lane;
            const bool valid = k < nred;
            Step s{0.0f, 0, 0};
            if (valid) {
                const int p = P.minred + k;
                const double v = chain_val(G, ckr, aut, 1, k);
                s = red_step(P, ln, lc.H, D, S, lc.stehle, lc.cv, k, v, (double)cont[p - 1] * G.cutoff, thr0);
            }
            const unsigned long long stopm = __ballot(valid && s.stop);
            const int lim = stopm ? __ffsll((long long)stopm) - 1 : kWave - 1;
            const unsigned long long keep = lim == 63 ? ~0ull : ((1ull << (lim + 1)) - 1);
            const unsigned long long addm = __ballot(valid && s.add) & keep;
            if (addm) {
                const int first = __ffsll((long long)addm) - 1, last = 63 - __clzll((long long)addm);
                rlo = min(rlo, P.minred + k0 + first);
                rhi = max(rhi, P.minred + k0 + last);
            }
            if (stopm) {
                stopped = 1;
                break;
            }
        }
    }
    if (stopped ? P.blue_if_red_stop : P.blue_if_red_done) {
        const int nblue = P.ib0 - 1;
        const double *ckb = ck + ckoff[l * 4 + 2];
        for (int k0 = 0; k0 < nblue; k0 += kWave) {
            const int k = k0 + lane;
            const bool valid = k < nblue;
            Step s{0.0f, 0, 0};
            if (valid) {
                const int p = P.ib0 - 1 - k;
                const double v = chain_val(G, ckb, aut, 0, k);
                s = blue_step(P, ln, lc.H, D, S, lc.stehle, lc.cv, v, (double)cont[p - 1] * G.cutoff);
            }
            const unsigned long long stopm = __ballot(valid && s.stop);
            const int lim = stopm ? __ffsll((long long)stopm) - 1 : kWave - 1;
            const unsigned long long keep = lim == 63 ? ~0ull : ((1ull << (lim + 1)) - 1);
            const unsigned long long addm = __ballot(valid && s.add) & keep;
            if (addm) {
                const int first = __ffsll((long long)addm) - 1, last = 63 - __clzll((long long)addm);
                bhi = max(bhi, P.ib0 - 1 - (k0 + first));
                blo = min(blo, P.ib0 - 1 - (k0 + last));
            }
            if (stopm) break;
        }
    }
    if (rlo <= rhi) {
        e.x = rlo;
        e.y = rhi;
    }
    if (blo <= bhi) {
        e.z = blo;
        e.w = bhi;
    }
    if (lane == 0) ext[q] = e;
}

constexpr int kAccBlock = 256;

__global__ __launch_bounds__(kAccBlock) void k_cl_accum(int nlines, int nh, Grid G, Stark S,
                                                        const a12_cplx_line *__restrict__ lines,
                                                        const int *__restrict__ hidx, const HLine *__restrict__ hl,
                                                        const Depth *__restrict__ depth,
                                                        const Plan *__restrict__ plan, const int *__restrict__ owner,
                                                        const HConv *__restrict__ conv, const int *__restrict__ ckoff,
                                                        const double *__restrict__ ck, const int4 *__restrict__ ext,
                                                        float *__restrict__ buf)
{
    const int L = G.length, j = blockIdx.y;
    const int p0 = blockIdx.x * kAccBlock + 1, p1 = min(L, p0 + kAccBlock - 1);
    const int p = p0 + threadIdx.x;
    const Depth D = depth[j];
    float acc = 0.0f;
    for (int l = 0; l < nlines; ++l) {
        const size_t q = (size_t)j * nlines + l;
        const int4 e = ext[q];
        const bool rr = e.x <= e.y && e.x <= p1 && e.y >= p0;
        const bool bb = e.z <= e.w && e.z <= p1 && e.w >= p0;
        if (!rr && !bb) continue;
        const bool in_red = p >= e.x && p <= e.y, in_blue = p >= e.z && p <= e.w;
        if (!in_red && !in_blue) continue;
        const Plan &P = plan[q];
        if (P.kind == K_CONT) {
            acc = acc + cont_kappa(P, p);
            continue;
        }
        const a12_cplx_line &ln = lines[l];
        const LineCtx lc = line_ctx(S, nlines, nh, j, l, P, hidx, hl, D, owner, conv);
        const int aut = P.kind == K_AUT;
        const int red = in_red ? 1 : 0;
        const int k = red ? p - P.minred : P.ib0 - 1 - p;
        const double v = chain_val(G, ck + ckoff[l * 4 + (red ? 0 : 2)], aut, red, k);
        acc = acc + point_kappa(P, ln, lc.H, D, S, lc.stehle, lc.cv, v);
    }
    if (p <= L) buf[(size_t)j * L + p - 1] = acc;
}

__global__ void k_cl_add(size_t total, const float *__restrict__ buf, void *opacity, int f64)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    if (f64) static_cast<double *>(opacity)[i] += (double)buf[i];
    else static_cast<float *>(opacity)[i] += buf[i];
}

// The pipeline for one SYNTHE state: plans, owners, convolutions, extents,
// accumulation into C.buf, on the context stream.
int launch_pipeline(a12_ctx *ctx, CplxDev &C, const double *xnfdop, const float *cont32)
{
    SyntheDev &Y = ctx->syn;
    hipStream_t st = ctx->stream;
    const int n = C.n, N = C.nlines;
    const bool prof = ctx->prof;
    if (prof && !C.ev[0])
        for (hipEvent_t &e : C.ev) A12_HIP(hipEventCreate(&e));
    if (prof) A12_HIP(hipEventRecord(C.ev[0], st));
    if (N > 0) {
        hipLaunchKernelGGL(k_cl_plan, dim3((N + 255) / 256, n), dim3(256), 0, st, N, C.mw6, C.G,
                           C.lines.as<a12_cplx_line>(), C.depth.as<Depth>(), xnfdop, Y.dopple.as<double>(),
                           Y.xnfpel.as<double>(), cont32, C.plan.as<Plan>());
        A12_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_cl_owner, dim3((n + 63) / 64), dim3(64), 0, st, N, n, C.G,
                           C.lines.as<a12_cplx_line>(), C.plan.as<Plan>(), C.owner.as<int>());
        A12_CHECK_LAUNCH();
    }
    if (prof) A12_HIP(hipEventRecord(C.ev[1], st));
    if (C.nh > 0) {
        hipLaunchKernelGGL(k_cl_conv, dim3(C.nh, n), dim3(kWave), 0, st, N, C.nh, C.S, C.hline_of.as<int>(),
                           C.hl.as<HLine>(), C.depth.as<Depth>(), C.plan.as<Plan>(), C.owner.as<int>(),
                           C.conv.as<HConv>());
        A12_CHECK_LAUNCH();
    }
    if (prof) A12_HIP(hipEventRecord(C.ev[2], st));
    if (N > 0) {
        hipLaunchKernelGGL(k_cl_extent, dim3(N, n), dim3(kWave), 0, st, N, C.nh, C.G, C.S,
                           C.lines.as<a12_cplx_line>(), C.hidx.as<int>(), C.hl.as<HLine>(), C.depth.as<Depth>(),
                           C.plan.as<Plan>(), C.owner.as<int>(), C.conv.as<HConv>(), cont32,
                           C.ckoff.as<int>(), C.ck.as<double>(), C.ext.as<int4>());
        A12_CHECK_LAUNCH();
    }
    if (prof) A12_HIP(hipEventRecord(C.ev[3], st));
    hipLaunchKernelGGL(k_cl_accum, dim3((C.L + kAccBlock - 1) / kAccBlock, n), dim3(kAccBlock), 0, st, N, C.nh,
                       C.G, C.S, C.lines.as<a12_cplx_line>(), C.hidx.as<int>(), C.hl.as<HLine>(),
                       C.depth.as<Depth>(), C.plan.as<Plan>(), C.owner.as<int>(), C.conv.as<HConv>(),
                       C.ckoff.as<int>(), C.ck.as<double>(), C.ext.as<int4>(), C.buf.as<float>());
    A12_CHECK_LAUNCH();
    if (prof) A12_HIP(hipEventRecord(C.ev[4], st));
    C.prof_valid = prof;
    C.ran = true;
    return A12_OK;
}

int check_state(a12_ctx *ctx, const CplxDev &C, const char *who)
{
    if (!C.loaded) {
        ctx->err = std::string(who) + ": call a12_cplx_load first";
        return A12_ERR_STATE;
    }
    if (!ctx->syn.loaded || ctx->syn.gen != C.syn_gen) {
        ctx->err = std::string(who) + ": the SYNTHE atmosphere changed since a12_cplx_load";
        return A12_ERR_STATE;
    }
    return A12_OK;
}

std::string check_stark(const a12_kapp_stark *st)
{
    for (int s = 0; s < 4; ++s) {
        const a12_kapp_stark &x = st[s];
        if (!x.loaded) continue;
        if (x.n_dens < 2 || x.n_dens > 20 || x.n_upper_min < s + 2 || x.n_upper_max < x.n_upper_min ||
            !x.dens || !x.temp || !x.log_dalpha || !x.max_dens_idx || !x.k_alpha || !x.profiles)
            return "Stehle series " + std::to_string(s + 1) + " is inconsistent";
        const int nt = x.n_upper_max - x.n_upper_min + 1;
        for (int q = 0; q < nt; ++q)
            if (x.max_dens_idx[q] < 1 || x.max_dens_idx[q] > x.n_dens)
                return "Stehle series " + std::to_string(s + 1) + ": density index out of range";
    }
    return "";
}

std::string check_line(const a12_cplx_line &ln, int mw6)
{
    const int t = ln.itype;
    if (t == 2) return "";
    if (t == -1 || t == -2) {
        if (ln.nblo < 1 || ln.nblo > 99 || ln.nbup < 2 || ln.nbup > 100 || ln.nbup <= ln.nblo)
            return "hydrogen line with n, m out of range";
        if (ln.ncon != 0 && ln.nbup != ln.nblo + 1 && ln.nbup != ln.nblo + 2 && (ln.ncon < 1 || ln.ncon > 15))
            return "merged hydrogen line with NCON outside 1..15";
        return "";
    }
    if (ln.nelion < 1 || ln.nelion > mw6) return "species slot out of range";
    if ((t == 0 || t == 3 || t <= -3) && ln.ncon > 0 && (ln.ncon > 26 || ln.nelionx < 1 || ln.nelionx > 17))
        return "NCON / NELIONX outside CONTX";
    return "";
}

}  // namespace

bool a12i_cplx_active(const a12_ctx *ctx) { return ctx->cplx && ctx->cplx->loaded; }

int a12i_cplx_add(a12_ctx *ctx, const double *xnfdop, const float *cont32, void *opacity, bool f64)
{
    CplxDev &C = *ctx->cplx;
    int rc = check_state(ctx, C, "a12_synthe_opacity_run (complex lines)");
    if (rc) return rc;
    rc = launch_pipeline(ctx, C, xnfdop, cont32);
    if (rc) return rc;
    const size_t total = (size_t)C.n * C.L;
    hipLaunchKernelGGL(k_cl_add, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, ctx->stream, total,
                       C.buf.as<float>(), opacity, f64 ? 1 : 0);
    A12_CHECK_LAUNCH();
    return A12_OK;
}

void a12i_cplx_free(a12_ctx *ctx)
{
    if (!ctx->cplx) return;
    CplxDev &C = *ctx->cplx;
    DevBuf *bufs[] = {&C.lines, &C.hl, &C.hline_of, &C.hidx, &C.depth, &C.sd, &C.si, &C.plan, &C.owner,
                      &C.conv, &C.ck, &C.ckoff, &C.ext, &C.buf, &C.prep};
    for (DevBuf *b : bufs) b->release();
    for (hipEvent_t e : C.ev)
        if (e) (void)hipEventDestroy(e);
    delete ctx->cplx;
    ctx->cplx = nullptr;
}

extern "C" {

int a12_cplx_load(a12_ctx *ctx, const a12_synthe_atm *atm, const a12_cplx_line *lines, int32_t nlines,
                  const a12_kapp_stark *stark, double wlend, double cutoff, int32_t ifvac, float velshift)
{
    if (!ctx) return A12_ERR_ARG;
    if (!atm || !stark || nlines < 0 || (nlines > 0 && !lines)) {
        ctx->err = "a12_cplx_load: null argument";
        return A12_ERR_ARG;
    }
    SyntheDev &Y = ctx->syn;
    if (!Y.loaded) {
        ctx->err = "a12_cplx_load: call a12_synthe_load first";
        return A12_ERR_STATE;
    }
    if (atm->nrhox != Y.nrhox || atm->length != Y.L || atm->wbegin != Y.wbegin || atm->ratio != Y.ratio ||
        atm->mw6 != Y.mw6 || !atm->t || !atm->xne || !atm->hckt || !atm->rho || !atm->xnfh || !atm->xnfhe1 ||
        !atm->xnfh2 || !atm->xnfpel || !atm->dopple) {
        ctx->err = "a12_cplx_load: atm is not the atmosphere given to a12_synthe_load";
        return A12_ERR_ARG;
    }
    if (!(cutoff >= 0.0) || !(wlend > 0.0)) {
        ctx->err = "a12_cplx_load: bad cutoff or wlend";
        return A12_ERR_ARG;
    }
    std::string e = check_stark(stark);
    if (!e.empty()) {
        ctx->err = "a12_cplx_load: " + e;
        return A12_ERR_ARG;
    }
    for (int l = 0; l < nlines; ++l) {
        e = check_line(lines[l], atm->mw6);
        if (!e.empty()) {
            ctx->err = "a12_cplx_load: line " + std::to_string(l) + ": " + e;
            return A12_ERR_ARG;
        }
    }
    // the window: RATIOLG and IXWLBEG as the driver forms them (src/synthe.f90:316-324)
    Grid G;
    G.length = atm->length;
    G.wbegin = atm->wbegin;
    G.ratio = atm->ratio;
    G.ratiolg = log(atm->ratio);
    G.ixwlbeg = (int)std::lround(log(atm->wbegin) / G.ratiolg);
    if (exp((double)G.ixwlbeg * G.ratiolg) != atm->wbegin) {
        ctx->err = "a12_cplx_load: wbegin is not EXP(IXWLBEG * LOG(ratio))";
        return A12_ERR_ARG;
    }
    G.wlend = wlend;
    G.cutoff = cutoff;
    G.ifvac = ifvac;
    G.dopratio = 1.0 + (double)velshift / CLIGHT_KMS;
    A12_HIP(hipSetDevice(ctx->device));
    CplxDev &C = cdev(ctx);
    C.loaded = false;
    C.ran = false;
    C.prof_valid = false;
    const int n = atm->nrhox, L = atm->length, mw6 = atm->mw6;
    hipStream_t st = ctx->stream;
    // per depth
    std::vector<Depth> dep(n);
    for (int j = 0; j < n; ++j)
        depth_setup(atm->t[j], atm->xne[j], atm->hckt[j], atm->rho[j], atm->xnfh[j], atm->xnfhe1[j], atm->xnfh2[j],
                    atm->xnfpel[(size_t)j * mw6 + 0], atm->xnfpel[(size_t)j * mw6 + 1], dep[j]);
    // hydrogen lines
    std::vector<int> hidx(std::max(nlines, 1), -1), hof;
    std::vector<HLine> hl;
    for (int l = 0; l < nlines; ++l)
        if (lines[l].itype == -1 || lines[l].itype == -2) {
            hidx[l] = (int)hof.size();
            hof.push_back(l);
            HLine H;
            hline_setup(lines[l].nblo, lines[l].nbup, H);
            hl.push_back(H);
        }
    const int nh = (int)hof.size();
    // wing chains: checkpoints every CK points of each wing
    std::vector<int> ckoff((size_t)std::max(nlines, 1) * 4, 0);
    size_t nck = 0;
    for (int l = 0; l < nlines; ++l) {
        const int t = lines[l].itype;
        if (t == 2 || t >= 4) continue;
        const int nb = lines[l].nbuff;
        const int minred = nb > 1 ? nb : 1, ib0 = nb < L + 1 ? nb : L + 1;
        const int nred = std::max(0, L - minred + 1), nblue = std::max(0, ib0 - 1);
        ckoff[l * 4 + 0] = (int)nck;
        ckoff[l * 4 + 1] = nred;
        nck += (nred + CK - 1) / CK;
        ckoff[l * 4 + 2] = (int)nck;
        ckoff[l * 4 + 3] = nblue;
        nck += (nblue + CK - 1) / CK;
    }
    if (nck > (size_t)INT_MAX) {
        ctx->err = "a12_cplx_load: too many lines for the chain table";
        return A12_ERR_ARG;
    }
    // Stehle tables: one double and one int pack, pointers into them
    std::vector<double> sd;
    std::vector<int> si;
    std::vector<double> dg(4 * NDA);
    stark_pow10(stark, dg.data());
    struct Off { size_t dens, temp, lg, dg, ka, prof, maxd; } off[4] = {};
    for (int s = 0; s < 4; ++s) {
        const a12_kapp_stark &x = stark[s];
        if (!x.loaded) continue;
        const int nt = x.n_upper_max - x.n_upper_min + 1;
        off[s].dens = sd.size();
        sd.insert(sd.end(), x.dens, x.dens + x.n_dens);
        off[s].temp = sd.size();
        sd.insert(sd.end(), x.temp, x.temp + NTEMP);
        off[s].lg = sd.size();
        sd.insert(sd.end(), x.log_dalpha, x.log_dalpha + NDA);
        off[s].dg = sd.size();
        sd.insert(sd.end(), dg.begin() + s * NDA, dg.begin() + (s + 1) * NDA);
        off[s].ka = sd.size();
        sd.insert(sd.end(), x.k_alpha, x.k_alpha + nt);
        off[s].prof = sd.size();
        sd.insert(sd.end(), x.profiles, x.profiles + (size_t)nt * x.n_dens * NTEMP * NDA);
        off[s].maxd = si.size();
        si.insert(si.end(), x.max_dens_idx, x.max_dens_idx + nt);
    }
    A12_HIP(upload(C.sd, sd.data(), sd.size(), st));
    A12_HIP(upload(C.si, si.data(), si.size(), st));
    Stark S{};
    for (int s = 0; s < 4; ++s) {
        StarkSeries &z = S.s[s];
        const a12_kapp_stark &x = stark[s];
        z.loaded = x.loaded;
        z.nmin = x.n_upper_min;
        z.nmax = x.n_upper_max;
        z.nd = x.n_dens;
        if (!x.loaded) continue;
        const double *b = C.sd.as<double>();
        z.dens = b + off[s].dens;
        z.temp = b + off[s].temp;
        z.lg = b + off[s].lg;
        z.dg = b + off[s].dg;
        z.kalpha = b + off[s].ka;
        z.prof = b + off[s].prof;
        z.maxd = C.si.as<int>() + off[s].maxd;
    }
    A12_HIP(upload(C.lines, lines, (size_t)nlines, st));
    A12_HIP(upload(C.depth, dep.data(), dep.size(), st));
    A12_HIP(upload(C.hl, hl.data(), hl.size(), st));
    A12_HIP(upload(C.hline_of, hof.data(), hof.size(), st));
    A12_HIP(upload(C.hidx, hidx.data(), (size_t)nlines, st));
    A12_HIP(upload(C.ckoff, ckoff.data(), (size_t)nlines * 4, st));
    A12_HIP(C.ck.reserve(sizeof(double) * std::max<size_t>(nck, 1)));
    const size_t nq = (size_t)n * std::max(nlines, 1);
    A12_HIP(C.plan.reserve(sizeof(Plan) * nq));
    A12_HIP(C.owner.reserve(sizeof(int) * nq));
    A12_HIP(C.ext.reserve(sizeof(int4) * nq));
    A12_HIP(C.conv.reserve(sizeof(HConv) * (size_t)n * std::max(nh, 1)));
    A12_HIP(C.buf.reserve(sizeof(float) * (size_t)n * L));
    if (nlines > 0) {
        hipLaunchKernelGGL(k_cl_chain, dim3((2 * nlines + 63) / 64), dim3(64), 0, st, nlines, G,
                           C.lines.as<a12_cplx_line>(), C.ckoff.as<int>(), C.ck.as<double>());
        A12_CHECK_LAUNCH();
    }
    A12_HIP(hipStreamSynchronize(st));
    C.nlines = nlines;
    C.nh = nh;
    C.n = n;
    C.L = L;
    C.mw6 = mw6;
    C.G = G;
    C.S = S;
    C.syn_gen = Y.gen;
    C.loaded = true;
    return A12_OK;
}

int a12_cplx_run(a12_ctx *ctx)
{
    if (!ctx) return A12_ERR_ARG;
    if (!ctx->cplx) {
        ctx->err = "a12_cplx_run: call a12_cplx_load first";
        return A12_ERR_STATE;
    }
    CplxDev &C = *ctx->cplx;
    int rc = check_state(ctx, C, "a12_cplx_run");
    if (rc) return rc;
    A12_HIP(hipSetDevice(ctx->device));
    // the SYNTHE prep this run reads: xnfdop (n, mw6), dscal (n, 4), cont32 (n, L)
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t b1 = al(sizeof(double) * (size_t)C.n * C.mw6), b2 = al(sizeof(double) * 4 * (size_t)C.n);
    A12_HIP(C.prep.reserve(b1 + b2 + sizeof(float) * (size_t)C.n * C.L));
    char *base = C.prep.as<char>();
    double *xnfdop = reinterpret_cast<double *>(base);
    double *dscal = reinterpret_cast<double *>(base + b1);
    float *cont32 = reinterpret_cast<float *>(base + b1 + b2);
    rc = a12i_synthe_prep(ctx, xnfdop, dscal, cont32);
    if (rc) return rc;
    rc = launch_pipeline(ctx, C, xnfdop, cont32);
    if (rc) return rc;
    A12_HIP(hipStreamSynchronize(ctx->stream));
    return A12_OK;
}

int a12_cplx_fetch(a12_ctx *ctx, float *buffer, int32_t *extents)
{
    if (!ctx) return A12_ERR_ARG;
    if (!ctx->cplx || !ctx->cplx->ran) {
        ctx->err = "a12_cplx_fetch: nothing computed";
        return A12_ERR_STATE;
    }
    CplxDev &C = *ctx->cplx;
    A12_HIP(hipSetDevice(ctx->device));
    if (buffer)
        A12_HIP(hipMemcpyAsync(buffer, C.buf.p, sizeof(float) * (size_t)C.n * C.L, hipMemcpyDeviceToHost,
                               ctx->stream));
    if (extents && C.nlines > 0)
        A12_HIP(hipMemcpyAsync(extents, C.ext.p, sizeof(int4) * (size_t)C.n * C.nlines, hipMemcpyDeviceToHost,
                               ctx->stream));
    A12_HIP(hipStreamSynchronize(ctx->stream));
    return A12_OK;
}

int a12_cplx_clear(a12_ctx *ctx)
{
    if (!ctx) return A12_ERR_ARG;
    if (ctx->cplx) {
        ctx->cplx->loaded = false;
        ctx->cplx->ran = false;
    }
    return A12_OK;
}

int a12_cplx_ms(a12_ctx *ctx, float ms[4])
{
    if (!ctx || !ms) return A12_ERR_ARG;
    for (int i = 0; i < 4; ++i) ms[i] = -1.0f;
    if (!ctx->cplx || !ctx->cplx->prof_valid) return A12_OK;
    CplxDev &C = *ctx->cplx;
    A12_HIP(hipEventSynchronize(C.ev[4]));
    for (int i = 0; i < 4; ++i) A12_HIP(hipEventElapsedTime(&ms[i], C.ev[i], C.ev[i + 1]));
    return A12_OK;
}

}  // extern "C"
