// a12_kapp.hip -- KAPP, the continuum opacity (src/atlas12_modules.f90:11202-11324
// and every source it calls), on gfx950.
//
// Work split (a12_kapp_core.h):
//   k_kapp_freq  one thread per frequency: the frequency-only row (Karzas &
//                Latter cross sections of 80 H levels, He I levels and
//                edges, H- tables, Rayleigh fits, CH/OH rows, Fe I/II cross
//                sections, HLINOP's levels).  Rows of the ATLAS12 grid are
//                formed once per line-store grid and kept.
//   k_kapp_state one thread per depth: the per-state factors the sources
//                cache between calls; k_kapp_tb: TOPbase Boltzmann factors,
//                one thread per (level, depth).
//   k_kapp_cells one thread per (frequency, depth), consecutive threads
//                walk the depths of a frequency: every source and KAPP's
//                totals, written as the six (nrhox, numnu) tables the RT
//                pass reads (or all 23 per-source arrays for a12_kapp_eval).
// The per-cell work is VALU FP64 (no HBM-heavy streams: a frequency row is
// read by the 72 threads of its depths from cache); TOPbase's ~2,000 level
// loop dominates it.
#include "a12_internal.h"

#pragma clang fp contract(off)

#include "a12_kapp_core.h"
#include "a12_kapp_host.h"

#include <cstring>
#include <vector>

using namespace kapp;

struct KappDev {
    KPack pack;
    DevBuf td, ti, tf;           // packed tables
    bool tables = false;
    int n = 0, mion = 0;
    std::vector<double> t_h;     // T of the KAPP state
    DevBuf st, planes;           // packed state, per-state planes
    bool state = false;
    DevBuf fq;                   // frequency rows of the ATLAS12 grid
    uint64_t fq_gen = 0;
    int fq_numnu = 0;
    DevBuf ew, eout, erow;       // a12_kapp_eval scratch
    DevBuf tbsum;                // MBF_TOPBASE's sum per cell
    DevBuf cbuf;                 // the 17 per-source values per cell between the kernels
    DevBuf tbs, etbs;            // TOPbase cross sections (nlev, nw) of the grid / of an eval
    std::vector<double> stage;
};

namespace {

KTab dev_tab(const KappDev &K)
{
    return kapp_tab(K.pack, K.td.as<double>(), K.ti.as<int>(), K.tf.as<float>());
}

__global__ void k_kapp_freq(KTab T, int nw, const double *__restrict__ waves, double *__restrict__ rows)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nw) return;
    kp_freq(T, waves[w], rows + (size_t)w * KF_N);
}

__global__ void k_kapp_state(KTab T, KState S, double *__restrict__ planes)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= S.n) return;
    kp_state(T, S, Q_HOLTSMARK, j, planes);
}

__global__ void k_kapp_tb(KTab T, KState S, double *__restrict__ planes)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= T.tb_nlev_tot * S.n) return;
    const int l = idx / S.n, j = idx - l * S.n;
    planes[(size_t)(KS_TB + l) * S.n + j] = kp_state_tb(T, S, l, j);
}

// SIGMA_NU of every TOPbase level at every wavelength: sig (nlev, nw)
__global__ void k_kapp_tbsig(KTab T, int nw, const double *__restrict__ waves, double *__restrict__ sig)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    const int l = blockIdx.y;
    if (w >= nw) return;
    sig[(size_t)l * nw + w] = kp_tb_sigma(T, l, waves[w]);
}

// MBF_TOPBASE's sum, the first part of ACONT_METAL: one lane per wavelength,
// a workgroup per (64 wavelengths, 8 depths), so a level's weights for the
// 8 depths are the same for every lane; each depth's sum runs over the
// levels in the reference's order (kp_cell_tb).  The depth group is the
// grid's fastest index: the workgroups of one wavelength tile run together
// and share its cross sections through the caches.
constexpr int TB_DEPTHS = 8;
__global__ void __launch_bounds__(64) k_kapp_tbsum(KTab T, KState S, const double *__restrict__ planes,
                                                   const double *__restrict__ sig,
                                                   const double *__restrict__ waves, int nw,
                                                   double *__restrict__ tbsum)
{
    const int w = blockIdx.y * 64 + threadIdx.x;
    const int n = S.n, j0 = blockIdx.x * TB_DEPTHS;
    if (w >= nw) return;
    const double freq = CLIGHT_NMS / waves[w];
    double sor[TB_DEPTHS], acc[TB_DEPTHS];
#pragma unroll
    for (int d = 0; d < TB_DEPTHS; ++d) {
        const int j = j0 + d < n ? j0 + d : n - 1;
        sor[d] = (1.0 - exp(-freq * S.hkt[j])) / S.rho[j];
        acc[d] = 0.0;
    }
    const double *W = planes + (size_t)KS_TB * n + j0;
    const int nd = n - j0 < TB_DEPTHS ? n - j0 : TB_DEPTHS;
    auto add = [&](int l, double sg) {
        if (sg == 0.0) return;
        const double *wl = W + (size_t)l * n;
        if (nd == TB_DEPTHS) {
#pragma unroll
            for (int d = 0; d < TB_DEPTHS; ++d) acc[d] = acc[d] + wl[d] * sg * sor[d];
        } else {
            for (int d = 0; d < nd; ++d) acc[d] = acc[d] + wl[d] * sg * sor[d];
        }
    };
    // eight levels' cross sections loaded together, then added in level order
    int l = 0;
    for (; l + 8 <= T.tb_nlev_tot; l += 8) {
        double sg[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) sg[u] = sig[(size_t)(l + u) * nw + w];
#pragma unroll
        for (int u = 0; u < 8; ++u) add(l + u, sg[u]);
    }
    for (; l < T.tb_nlev_tot; ++l) add(l, sig[(size_t)l * nw + w]);
#pragma unroll
    for (int d = 0; d < TB_DEPTHS; ++d)
        if (d < nd) tbsum[(size_t)w * n + j0 + d] = acc[d];
}

// The three parts of the rest of KAPP, one thread per cell; o (17, cells)
// holds the per-source values between them.
__global__ void __launch_bounds__(256) k_kapp_ca(KTab T, KState S, const double *__restrict__ planes,
                                                 const double *__restrict__ rows, const double *__restrict__ waves,
                                                 int nw, double *__restrict__ o)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int n = S.n;
    const long cells = (long)nw * n;
    if (idx >= cells) return;
    const int w = (int)(idx / n), j = (int)(idx - (long)w * n);
    double v[17];
    kp_cell_a(T, S, planes, rows + (size_t)w * KF_N, waves[w], j, v);
    const int ks[8] = {0, 1, 2, 3, 4, 5, 10, 11};
#pragma unroll
    for (int q = 0; q < 8; ++q) o[ks[q] * cells + idx] = v[ks[q]];
}

__global__ void __launch_bounds__(256) k_kapp_cb(KTab T, KState S, const double *__restrict__ planes,
                                                 const double *__restrict__ rows, const double *__restrict__ waves,
                                                 int nw, const double *__restrict__ tbsum, double *__restrict__ o)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int n = S.n;
    const long cells = (long)nw * n;
    if (idx >= cells) return;
    const int w = (int)(idx / n), j = (int)(idx - (long)w * n);
    double v[17];
    kp_cell_b(T, S, planes, rows + (size_t)w * KF_N, waves[w], j, tbsum[idx], v);
#pragma unroll
    for (int q = 6; q <= 14; ++q)
        if (q <= 9 || q >= 12) o[q * cells + idx] = v[q];
}

__global__ void __launch_bounds__(256) k_kapp_cc(KTab T, KState S, const double *__restrict__ planes,
                                                 const double *__restrict__ rows, const double *__restrict__ waves,
                                                 int nw, double *__restrict__ o)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int n = S.n;
    const long cells = (long)nw * n;
    if (idx >= cells) return;
    const int w = (int)(idx / n), j = (int)(idx - (long)w * n);
    double v[17];
    v[15] = v[16] = 0.0;
    if (rows[(size_t)w * KF_N + KF_HLN] > 0.0) kp_cell_c(T, S, planes, rows + (size_t)w * KF_N, waves[w], j, v);
    o[15 * cells + idx] = v[15];
    o[16 * cells + idx] = v[16];
}

// KAPP's totals.  mode 0: out = six tables (nrhox, nw) in the order ACONT
// SIGMAC SCONT SLINE ALINE SIGMAL; mode 1: out (nw, 23, nrhox)
__global__ void __launch_bounds__(256) k_kapp_tot(KState S, const double *__restrict__ waves, int nw,
                                                  const double *__restrict__ o, int mode,
                                                  double *__restrict__ out)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int n = S.n;
    const long cells = (long)nw * n;
    if (idx >= cells) return;
    const int w = (int)(idx / n), j = (int)(idx - (long)w * n);
    double v[17], cell[A12_KAPP_NSRC];
#pragma unroll
    for (int q = 0; q < 17; ++q) v[q] = o[q * cells + idx];
    kp_totals(v, waves[w], S.hkt[j], cell);
    if (mode == 0) {
#pragma unroll
        for (int k = 0; k < 6; ++k) out[k * cells + idx] = cell[17 + k];
    } else {
#pragma unroll
        for (int k = 0; k < A12_KAPP_NSRC; ++k) out[((size_t)w * A12_KAPP_NSRC + k) * n + j] = cell[k];
    }
}

KappDev &kdev(a12_ctx *ctx)
{
    if (!ctx->kapp) ctx->kapp = new KappDev();
    return *ctx->kapp;
}

int launch_rows(a12_ctx *ctx, KappDev &K, int nw, const double *d_waves, double *d_rows)
{
    k_kapp_freq<<<(nw + 63) / 64, 64, 0, ctx->stream>>>(dev_tab(K), nw, d_waves, d_rows);
    A12_HIP(hipGetLastError());
    return A12_OK;
}

int launch_tbsig(a12_ctx *ctx, KappDev &K, int nw, const double *d_waves, double *d_sig)
{
    const KTab T = dev_tab(K);
    if (T.tb_nlev_tot < 1) return A12_OK;
    k_kapp_tbsig<<<dim3((nw + 255) / 256, T.tb_nlev_tot), 256, 0, ctx->stream>>>(T, nw, d_waves, d_sig);
    A12_HIP(hipGetLastError());
    return A12_OK;
}

// d_sig: the TOPbase cross sections of these wavelengths (launch_tbsig)
int launch_cells(a12_ctx *ctx, KappDev &K, int nw, const double *d_waves, const double *d_rows,
                 const double *d_sig, int mode, double *d_out)
{
    const long cells = (long)nw * K.n;
    const KState S = kapp_state_view(K.st.as<double>(), K.n, K.mion);
    A12_HIP(K.tbsum.reserve(sizeof(double) * cells));
    const unsigned nb = (unsigned)((cells + 255) / 256);
    k_kapp_tbsum<<<dim3((K.n + TB_DEPTHS - 1) / TB_DEPTHS, (nw + 63) / 64), 64, 0, ctx->stream>>>(
        dev_tab(K), S, K.planes.as<double>(), d_sig, d_waves, nw, K.tbsum.as<double>());
    A12_HIP(hipGetLastError());
    A12_HIP(K.cbuf.reserve(sizeof(double) * 17 * cells));
    const KTab T = dev_tab(K);
    double *o = K.cbuf.as<double>();
    k_kapp_ca<<<nb, 256, 0, ctx->stream>>>(T, S, K.planes.as<double>(), d_rows, d_waves, nw, o);
    A12_HIP(hipGetLastError());
    k_kapp_cb<<<nb, 256, 0, ctx->stream>>>(T, S, K.planes.as<double>(), d_rows, d_waves, nw,
                                           K.tbsum.as<double>(), o);
    A12_HIP(hipGetLastError());
    k_kapp_cc<<<nb, 256, 0, ctx->stream>>>(T, S, K.planes.as<double>(), d_rows, d_waves, nw, o);
    A12_HIP(hipGetLastError());
    k_kapp_tot<<<nb, 256, 0, ctx->stream>>>(S, d_waves, nw, o, mode, d_out);
    A12_HIP(hipGetLastError());
    return A12_OK;
}

}  // namespace

void a12i_kapp_free(a12_ctx *ctx)
{
    if (!ctx->kapp) return;
    KappDev &K = *ctx->kapp;
    DevBuf *b[] = {&K.td, &K.ti, &K.tf, &K.st, &K.planes, &K.fq, &K.ew, &K.eout, &K.erow, &K.tbsum, &K.tbs, &K.etbs, &K.cbuf};
    for (DevBuf *p : b) p->release();
    delete ctx->kapp;
    ctx->kapp = nullptr;
}

extern "C" {

int a12_kapp_constants(double *out12)
{
    if (!out12) return A12_ERR_ARG;
    const double v[12] = {HCK, HOVERK, KBOL_EV, THETA_COEFF, BNU_PREFAC, saha_prefac(), FREQ_RYDH,
                          RYD_ANG, LN10, COEFF_FF, SIGMA_THOMSON, CLIGHT_ANGS};
    std::memcpy(out12, v, sizeof(v));
    return A12_OK;
}

int a12_kapp_set_tables(a12_ctx *ctx, const a12_kapp_tables *tables)
{
    if (!ctx) return A12_ERR_ARG;
    KappDev &K = kdev(ctx);
    KPack P;
    const std::string e = kapp_pack(tables, P);
    if (!e.empty()) {
        ctx->err = "a12_kapp_set_tables: " + e;
        return A12_ERR_ARG;
    }
    A12_HIP(hipSetDevice(ctx->device));
    K.tables = K.state = false;
    K.fq_gen = 0;
    K.pack = std::move(P);
    A12_HIP(K.td.reserve(sizeof(double) * K.pack.d.size()));
    A12_HIP(K.ti.reserve(sizeof(int) * K.pack.i.size()));
    A12_HIP(K.tf.reserve(sizeof(float) * K.pack.f.size()));
    A12_HIP(hipMemcpyAsync(K.td.p, K.pack.d.data(), sizeof(double) * K.pack.d.size(), hipMemcpyHostToDevice,
                           ctx->stream));
    A12_HIP(hipMemcpyAsync(K.ti.p, K.pack.i.data(), sizeof(int) * K.pack.i.size(), hipMemcpyHostToDevice,
                           ctx->stream));
    A12_HIP(hipMemcpyAsync(K.tf.p, K.pack.f.data(), sizeof(float) * K.pack.f.size(), hipMemcpyHostToDevice,
                           ctx->stream));
    A12_HIP(hipStreamSynchronize(ctx->stream));
    K.tables = true;
    return A12_OK;
}

int a12_kapp_set_state(a12_ctx *ctx, const a12_kapp_state *state)
{
    if (!ctx) return A12_ERR_ARG;
    KappDev &K = kdev(ctx);
    if (!K.tables) {
        ctx->err = "a12_kapp_set_state: call a12_kapp_set_tables first";
        return A12_ERR_STATE;
    }
    const std::string e = kapp_check_state(state);
    if (!e.empty()) {
        ctx->err = "a12_kapp_set_state: " + e;
        return A12_ERR_ARG;
    }
    if (state->mion < kapp_max_species(K.pack)) {
        ctx->err = "a12_kapp_set_state: mion is smaller than a species index the tables use";
        return A12_ERR_ARG;
    }
    A12_HIP(hipSetDevice(ctx->device));
    K.state = false;
    K.n = state->nrhox;
    K.mion = state->mion;
    K.t_h.assign(state->t, state->t + K.n);
    kapp_pack_state(state, K.stage);
    A12_HIP(K.st.reserve(sizeof(double) * K.stage.size()));
    A12_HIP(hipMemcpyAsync(K.st.p, K.stage.data(), sizeof(double) * K.stage.size(), hipMemcpyHostToDevice,
                           ctx->stream));
    const KTab T = dev_tab(K);
    const size_t np = (size_t)kapp_nplanes(T) * K.n;
    A12_HIP(K.planes.reserve(sizeof(double) * np));
    const KState S = kapp_state_view(K.st.as<double>(), K.n, K.mion);
    k_kapp_state<<<1, 128, 0, ctx->stream>>>(T, S, K.planes.as<double>());
    A12_HIP(hipGetLastError());
    const int ntb = T.tb_nlev_tot * K.n;
    if (ntb > 0) {
        k_kapp_tb<<<(ntb + 255) / 256, 256, 0, ctx->stream>>>(T, S, K.planes.as<double>());
        A12_HIP(hipGetLastError());
    }
    A12_HIP(hipStreamSynchronize(ctx->stream));
    K.state = true;
    return A12_OK;
}

int a12_kapp_eval(a12_ctx *ctx, int32_t nwave, const double *wave_nm, double *out)
{
    if (!ctx) return A12_ERR_ARG;
    KappDev &K = kdev(ctx);
    if (!K.tables || !K.state) {
        ctx->err = "a12_kapp_eval: call a12_kapp_set_tables and a12_kapp_set_state first";
        return A12_ERR_STATE;
    }
    if (nwave < 1 || nwave > (1 << 22) || !wave_nm || !out) {
        ctx->err = "a12_kapp_eval: bad nwave or null array";
        return A12_ERR_ARG;
    }
    for (int w = 0; w < nwave; ++w)
        if (!(wave_nm[w] > 0.0) || !std::isfinite(wave_nm[w])) {
            ctx->err = "a12_kapp_eval: wavelengths must be positive and finite";
            return A12_ERR_ARG;
        }
    A12_HIP(hipSetDevice(ctx->device));
    const size_t nout = (size_t)nwave * A12_KAPP_NSRC * K.n;
    A12_HIP(K.ew.reserve(sizeof(double) * nwave));
    A12_HIP(K.erow.reserve(sizeof(double) * (size_t)nwave * KF_N));
    A12_HIP(K.eout.reserve(sizeof(double) * nout));
    A12_HIP(hipMemcpyAsync(K.ew.p, wave_nm, sizeof(double) * nwave, hipMemcpyHostToDevice, ctx->stream));
    int rc = launch_rows(ctx, K, nwave, K.ew.as<double>(), K.erow.as<double>());
    if (rc) return rc;
    A12_HIP(K.etbs.reserve(sizeof(double) * ((size_t)K.pack.scal.tb_nlev_tot * nwave + 1)));
    rc = launch_tbsig(ctx, K, nwave, K.ew.as<double>(), K.etbs.as<double>());
    if (rc) return rc;
    rc = launch_cells(ctx, K, nwave, K.ew.as<double>(), K.erow.as<double>(), K.etbs.as<double>(), 1,
                      K.eout.as<double>());
    if (rc) return rc;
    A12_HIP(hipMemcpyAsync(out, K.eout.p, sizeof(double) * nout, hipMemcpyDeviceToHost, ctx->stream));
    A12_HIP(hipStreamSynchronize(ctx->stream));
    return A12_OK;
}

int a12_atlas_kapp_run(a12_ctx *ctx)
{
    if (!ctx) return A12_ERR_ARG;
    AtlasDev &A = ctx->atl;
    KappDev &K = kdev(ctx);
    if (!A.loaded) {
        ctx->err = "a12_atlas_kapp_run: call a12_atlas_set_state (or a12_atlas_load) first";
        return A12_ERR_STATE;
    }
    if (!K.tables || !K.state) {
        ctx->err = "a12_atlas_kapp_run: call a12_kapp_set_tables and a12_kapp_set_state first";
        return A12_ERR_STATE;
    }
    if (K.n != A.n || std::memcmp(K.t_h.data(), A.t_h.data(), sizeof(double) * K.n) != 0) {
        ctx->err = "a12_atlas_kapp_run: the KAPP state is not the ATLAS12 state (NRHOX or T differ)";
        return A12_ERR_STATE;
    }
    A12_HIP(hipSetDevice(ctx->device));
    int rc;
    if (K.fq_gen != A.grid_gen || K.fq_numnu != A.numnu) {
        A12_HIP(K.fq.reserve(sizeof(double) * (size_t)A.numnu * KF_N));
        rc = launch_rows(ctx, K, A.numnu, A.waveset.as<double>(), K.fq.as<double>());
        if (rc) return rc;
        A12_HIP(K.tbs.reserve(sizeof(double) * ((size_t)K.pack.scal.tb_nlev_tot * A.numnu + 1)));
        rc = launch_tbsig(ctx, K, A.numnu, A.waveset.as<double>(), K.tbs.as<double>());
        if (rc) return rc;
        K.fq_gen = A.grid_gen;
        K.fq_numnu = A.numnu;
    }
    const size_t plane = (size_t)A.n * A.numnu;
    A12_HIP(A.kapp.reserve(sizeof(double) * 6 * plane));
    if (ctx->prof) A12_HIP(hipEventRecord(ctx->ev[14], ctx->stream));
    rc = launch_cells(ctx, K, A.numnu, A.waveset.as<double>(), K.fq.as<double>(), K.tbs.as<double>(), 0,
                      A.kapp.as<double>());
    if (rc) return rc;
    if (ctx->prof) A12_HIP(hipEventRecord(ctx->ev[15], ctx->stream));
    A12_HIP(hipStreamSynchronize(ctx->stream));
    A.kapp_ready = true;
    A.kapp_plane = plane;
    A.prof_kapp = ctx->prof;
    return A12_OK;
}

int a12_atlas_kapp_fetch(a12_ctx *ctx, double *kapp6)
{
    if (!ctx || !kapp6) return A12_ERR_ARG;
    AtlasDev &A = ctx->atl;
    if (!A.kapp_ready) {
        ctx->err = "a12_atlas_kapp_fetch: no KAPP tables on the device";
        return A12_ERR_STATE;
    }
    A12_HIP(hipSetDevice(ctx->device));
    A12_HIP(hipMemcpyAsync(kapp6, A.kapp.p, sizeof(double) * 6 * A.kapp_plane, hipMemcpyDeviceToHost,
                           ctx->stream));
    A12_HIP(hipStreamSynchronize(ctx->stream));
    return A12_OK;
}

int a12_kapp_ms(a12_ctx *ctx, float *ms)
{
    if (!ctx || !ms) return A12_ERR_ARG;
    *ms = -1.0f;
    if (ctx->atl.prof_kapp && ctx->atl.kapp_ready) {
        A12_HIP(hipEventSynchronize(ctx->ev[15]));
        A12_HIP(hipEventElapsedTime(ms, ctx->ev[14], ctx->ev[15]));
    }
    return A12_OK;
}

}  // extern "C"
