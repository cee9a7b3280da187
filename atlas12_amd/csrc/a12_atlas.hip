// a12_atlas.hip -- ATLAS12 opacity-sampling line opacity (LINOP1) on gfx950
// and the host side of the ATLAS12 frequency loop.
//
// LINOP1 (src/atlas12_modules.f90:18132-18374) adds, for every selected line
// and depth, a Kurucz-table Voigt profile into XLINES(J, NU) on the 30,000
// point opacity-sampling grid, walking at most 100 points to each side and
// stopping after the first point below the line-rejection continuum.  Here:
//   host          unpack LINEDATA, continuum bin, grid index NU, line
//                 constants (CGF, ELO, damping), stable sort by NU and the
//                 per-NU segment offsets;
//   k_lo_setup    one thread per line walks the depths in the reference's
//                 coarse (every 8th) then fine order and fixes, per depth,
//                 kappa_centre, a, the Doppler width in nm and the red/blue
//                 reach (FP64 decisions, as the reference);
//   k_lo_reach    per (NU, depth): widest red/blue reach of the segment;
//   k_lo_gather   one block per grid point, one thread per depth: the sum of
//                 every reaching line's profile, visiting segments in NU
//                 order (for a wavelength-sorted store that is the
//                 reference's own line order).  No atomics.
// MIXED evaluates the profile in FP32 and sums with Neumaier compensation;
// FP64 repeats the reference's arithmetic.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "a12_common.h"
#include "a12_internal.h"

namespace {

constexpr int kMaxWing = 100;
constexpr int kTabN = 2001;

// VOIGT(V, A), src/atlas12_modules.f90:21033-21097, tables 1-based (IV).
template <class R>
__device__ __forceinline__ R voigt_kurucz(R v, R a, const R *h0, const R *h1, const R *h2)
{
#pragma clang fp contract(off)
    int iv = (int)(v * (R)200 + (R)1.5);
    iv = iv < 1 ? 1 : (iv > kTabN ? kTabN : iv);
    const R H0 = h0[iv - 1], H1 = h1[iv - 1], H2 = h2[iv - 1];
    if (a < (R)0.2) {
        if (v <= (R)10) return (H2 * a + H1) * a + H0;
        return (R)0.5642 * a / (v * v);
    } else if (a <= (R)1.4 && a + v <= (R)3.2) {
        const R vv = v * v;
        const R hh1 = H1 + H0 * (R)1.12838;
        const R hh2 = H2 + hh1 * (R)1.12838 - H0;
        const R hh3 = ((R)1 - H2) * (R)0.37613 - hh1 * (R)0.66667 * vv + hh2 * (R)1.12838;
        const R hh4 = ((R)3 * hh3 - hh1) * (R)0.37613 + H0 * (R)0.66667 * vv * vv;
        return ((((hh4 * a + hh3) * a + hh2) * a + hh1) * a + H0) *
               ((((R)-.122727278 * a + (R).532770573) * a - (R).96284325) * a + (R).979895032);
    }
    const R aa = a * a, vv = v * v;
    const R u = (aa + vv) * (R)1.4142;
    R r = a * (R)0.79788 / u;
    if (a <= (R)100) {
        const R aau = aa / u, vvu = vv / u, uu = u * u;
        r = ((((aau - (R)10 * vvu) * aau * (R)3 + (R)15 * vvu * vvu) + (R)3 * vv - aa) / uu + (R)1) * r;
    }
    return r;
}

// One wing value of LINOP1 (:18263-18298): center * H(v, a).
template <class R>
__device__ __forceinline__ R lo_cv(R center, R v, R adamp, const R *h0, const R *h1, const R *h2)
{
#pragma clang fp contract(off)
    if (adamp > (R)0.20) return center * voigt_kurucz<R>(v, adamp, h0, h1, h2);
    if (v > (R)10) return center * (R)0.5641895835477563 * adamp / (v * v);
    const int iv = (int)(v * (R)200 + (R)1.5);
    return center * ((h2[iv - 1] * adamp + h1[iv - 1]) * adamp + h0[iv - 1]);
}

// The same wing value with the arithmetic in FP32 but every table-row and
// regime decision taken from the FP64 offset v64: a row index computed from
// a rounded v flips to the neighbouring row near integer 200 v + 1.5, a
// step of ~0.5 % in that contribution.
__device__ __forceinline__ float lo_cv_mixed(float center, double v64, float adamp, const float *h0,
                                             const float *h1, const float *h2)
{
#pragma clang fp contract(off)
    const float v = (float)v64;
    if (adamp > 0.20f) {
        int iv = (int)(v64 * 200.0 + 1.5);
        iv = iv < 1 ? 1 : (iv > kTabN ? kTabN : iv);
        const float H0 = h0[iv - 1], H1 = h1[iv - 1], H2 = h2[iv - 1];
        const float a = adamp;
        if (a <= 1.4f && (double)a + v64 <= 3.2) {
            const float vv = v * v;
            const float hh1 = H1 + H0 * 1.12838f;
            const float hh2 = H2 + hh1 * 1.12838f - H0;
            const float hh3 = (1.0f - H2) * 0.37613f - hh1 * 0.66667f * vv + hh2 * 1.12838f;
            const float hh4 = (3.0f * hh3 - hh1) * 0.37613f + H0 * 0.66667f * vv * vv;
            return center * (((((hh4 * a + hh3) * a + hh2) * a + hh1) * a + H0) *
                             (((-.122727278f * a + .532770573f) * a - .96284325f) * a + .979895032f));
        }
        const float aa = a * a, vv = v * v;
        const float u = (aa + vv) * 1.4142f;
        float r = a * 0.79788f / u;
        if (a <= 100.0f) {
            const float aau = aa / u, vvu = vv / u, uu = u * u;
            r = ((((aau - 10.0f * vvu) * aau * 3.0f + 15.0f * vvu * vvu) + 3.0f * vv - aa) / uu + 1.0f) * r;
        }
        return center * r;
    }
    if (v64 > 10.0) return center * 0.5641895835477563f * adamp / (v * v);
    const int iv = (int)(v64 * 200.0 + 1.5);
    return center * ((h2[iv - 1] * adamp + h1[iv - 1]) * adamp + h0[iv - 1]);
}

// Per line: all depths.  o_state[l*n + j]: bit 31 computed, bits 0-7 red
// reach, 8-15 blue reach (contributions to points nu.. and nu-1..).
__global__ void k_lo_setup(int nl, int n, int numnu, int mion, const int *__restrict__ l_nu,
                           const int *__restrict__ l_nucont, const int *__restrict__ l_nel,
                           const double *__restrict__ l_const, const double *__restrict__ waveset,
                           const double *__restrict__ tabcont, const double *__restrict__ xnfdop,
                           const double *__restrict__ dopple, const double *__restrict__ dep,
                           const double *__restrict__ h0, const double *__restrict__ h1,
                           const double *__restrict__ h2, int *__restrict__ o_state,
                           double *__restrict__ o_par, long long *__restrict__ blk_evals)
{
#pragma clang fp contract(off)
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    long long ev = 0;
    if (l < nl) {
        const int nu = l_nu[l], nucont = l_nucont[l], nel = l_nel[l] - 1;
        const double wlvac = l_const[5 * (size_t)l + 0], cgf = l_const[5 * (size_t)l + 1],
                     elo = l_const[5 * (size_t)l + 2], gr = l_const[5 * (size_t)l + 3],
                     gs = l_const[5 * (size_t)l + 4];
        const double gw = l_const[5 * (size_t)nl + l];
        const double *xne = dep + 2 * n, *hckt = dep + 3 * n, *txnxn = dep + 5 * n;
        for (int j = 0; j < n; ++j) o_state[(size_t)l * n + j] = 0;
        unsigned ifj = 0;  // bit k: coarse depth 8k passed (bit 0 = IFJ(1) = 0)
        for (int pass = 0; pass < 2; ++pass) {
            for (int K = 8; K <= n; K += 8) {
                int jlo, jhi;
                if (pass == 0) {
                    jlo = jhi = K;
                } else {
                    if (!((ifj >> (K / 8 - 1)) & 1u) && !((ifj >> (K / 8)) & 1u)) continue;
                    jlo = K - 7;
                    jhi = K - 1;
                }
                for (int J = jlo; J <= jhi; ++J) {
                    const int j = J - 1;
                    const size_t o = (size_t)l * n + j;
                    const double tc = tabcont[(size_t)nucont * n + j];
                    double center = cgf * xnfdop[(size_t)nel * n + j];
                    int st = 0;
                    if (!(center < tc)) {
                        center = center * exp(-elo * hckt[j]);
                        if (!(center < tc)) {
                            if (pass == 0) ifj |= 1u << (K / 8);
                            const double dop = dopple[(size_t)nel * n + j];
                            const double adamp = (gr + gs * xne[j] + gw * txnxn[j]) / dop;
                            const double dopwave = dop * wlvac;
                            int rn = 0, bn = 0;
                            const int hi = min(nu + kMaxWing, numnu - 1);
                            for (int iw = nu; iw <= hi; ++iw) {
                                ++rn;
                                if (lo_cv<double>(center, (waveset[iw] - wlvac) / dopwave, adamp, h0,
                                                  h1, h2) < tc)
                                    break;
                            }
                            for (int i = 1; i <= kMaxWing; ++i) {
                                const int iw = nu - i;
                                if (iw < 0) break;
                                ++bn;
                                if (lo_cv<double>(center, (wlvac - waveset[iw]) / dopwave, adamp, h0,
                                                  h1, h2) < tc)
                                    break;
                            }
                            st = (int)(0x80000000u | (unsigned)rn | ((unsigned)bn << 8));
                            o_par[3 * o + 0] = center;
                            o_par[3 * o + 1] = adamp;
                            o_par[3 * o + 2] = dopwave;
                            ev += rn + bn;
                        }
                    }
                    o_state[o] = st;
                }
            }
        }
    }
    __shared__ long long red[256];
    red[threadIdx.x] = ev;
    __syncthreads();
    for (int s = blockDim.x >> 1; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) blk_evals[blockIdx.x] = red[0];
}

__global__ void k_lo_sum(const long long *__restrict__ blk, int nb, long long *out)
{
    __shared__ long long red[256];
    long long s = 0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) s += blk[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = blockDim.x >> 1; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// Widest red/blue reach per (segment k, depth j); reach[(k*n + j)] packs
// red (low 8 bits) and blue (next 8).
__global__ void k_lo_reach(int nseg, int n, const int *__restrict__ seg,
                           const int *__restrict__ state, int *__restrict__ reach)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nseg * n) return;
    const int k = i / n, j = i % n;
    int r = 0, b = 0;
    for (int l = seg[k]; l < seg[k + 1]; ++l) {
        const int st = state[(size_t)l * n + j];
        if (st < 0) {
            r = max(r, st & 0xff);
            b = max(b, (st >> 8) & 0xff);
        }
    }
    reach[i] = r | (b << 8);
}

template <bool F64>
__global__ __launch_bounds__(128) void k_lo_gather(
    int n, int numnu, const int *__restrict__ seg, const int *__restrict__ reach,
    const int *__restrict__ state, const double *__restrict__ par,
    const double *__restrict__ l_const, const double *__restrict__ waveset,
    const double *__restrict__ h0, const double *__restrict__ h1, const double *__restrict__ h2,
    const float *__restrict__ f0, const float *__restrict__ f1, const float *__restrict__ f2,
    double *__restrict__ xlines)
{
#pragma clang fp contract(off)
    const int iw = blockIdx.x;
    const int j = threadIdx.x;
    if (j >= n) return;
    const double w = waveset[iw];
    double acc64 = 0.0;
    Neumaier32 acc32;
    // red contributions: segments k <= iw; blue: k > iw
    const int klo = max(0, iw - kMaxWing), khi = min(numnu, iw + kMaxWing);
    for (int k = klo; k <= khi; ++k) {
        const int rc = reach[(size_t)k * n + j];
        const bool red = k <= iw;
        const int need = red ? iw - k + 1 : k - iw;
        const int have = red ? (rc & 0xff) : ((rc >> 8) & 0xff);
        if (have < need) continue;
        for (int l = seg[k]; l < seg[k + 1]; ++l) {
            const int st = state[(size_t)l * n + j];
            if (st >= 0) continue;
            const int reachl = red ? (st & 0xff) : ((st >> 8) & 0xff);
            if (reachl < need) continue;
            const size_t o = (size_t)l * n + j;
            const double wl = l_const[5 * (size_t)l];
            if (F64) {
                const double v = (red ? (w - wl) : (wl - w)) / par[3 * o + 2];
                acc64 = acc64 + lo_cv<double>(par[3 * o], v, par[3 * o + 1], h0, h1, h2);
            } else {
                const double v = (red ? (w - wl) : (wl - w)) / par[3 * o + 2];
                acc32.add(lo_cv_mixed((float)par[3 * o], v, (float)par[3 * o + 1], f0, f1, f2));
            }
        }
    }
    xlines[(size_t)iw * n + j] = F64 ? acc64 : (double)acc32.value();
}

template <class T>
hipError_t up(DevBuf &b, const T *src, size_t n, hipStream_t s)
{
    hipError_t e = b.reserve(std::max<size_t>(n, 1) * sizeof(T));
    if (e != hipSuccess || n == 0) return e;
    return hipMemcpyAsync(b.p, src, n * sizeof(T), hipMemcpyHostToDevice, s);
}

}  // namespace

// --------------------------------------------------------------------------
// Host side
// --------------------------------------------------------------------------
int a12i_atlas_load(a12_ctx *ctx, const a12_atlas_atm *a, const int32_t *linedata, int32_t nlines)
{
    AtlasDev &A = ctx->atl;
    if (!a || (!linedata && nlines > 0) || nlines < 0) {
        ctx->err = "a12_atlas_load: null argument";
        return A12_ERR_ARG;
    }
    if (a->nrhox < 8 || a->nrhox > A12_NRMAX_LAYERS || a->numnu < 3 || a->mion < 1 ||
        a->ldion < a->nrhox || a->ldtab < a->nrhox || a->nwave1 < 2 || !a->waveset ||
        !a->rcoset || !a->tabcont || !a->iwavetab || !a->xnfdop || !a->dopple || !a->t ||
        !a->xne || !a->hckt || !a->hkt || !a->rhox || !a->xnf1 || !a->xnf3 || !a->xnf841 ||
        !a->h0tab || !a->h1tab || !a->h2tab) {
        ctx->err = "a12_atlas_load: bad dimensions or null array";
        return A12_ERR_ARG;
    }
    const int n = a->nrhox, numnu = a->numnu;
    for (int i = 1; i < numnu; ++i)
        if (!(a->waveset[i] > a->waveset[i - 1])) {
            ctx->err = "a12_atlas_load: WAVESET must ascend";
            return A12_ERR_ARG;
        }
    hipStream_t s = ctx->stream;
    A.n = n;
    A.numnu = numnu;
    A.mion = a->mion;
    A.teff = a->teff;
    // depth vectors: t hkt rhox | xne hckt txnxn (TXNXN as LINOP1 forms it, :18164)
    std::vector<double> dep((size_t)6 * n);
    for (int j = 0; j < n; ++j) {
        dep[j] = a->t[j];
        dep[n + j] = a->hkt[j];
        dep[2 * n + j] = a->rhox[j];
        dep[3 * n + j] = a->hckt[j];
        dep[4 * n + j] = a->xne[j];
        dep[5 * n + j] = (a->xnf1[j] + 0.42 * a->xnf3[j] + 0.85 * a->xnf841[j]) *
                         std::pow(a->t[j] / 10000.0, 0.3);
    }
    // k_lo_setup reads xne at dep + 2n, hckt at dep + 3n, txnxn at dep + 5n: lay
    // the copy for it out as [t hkt xne hckt . txnxn]
    std::vector<double> dset(dep);
    for (int j = 0; j < n; ++j) dset[2 * n + j] = a->xne[j];
    A12_HIP(up(A.depth3, dep.data(), (size_t)3 * n, s));  // t hkt rhox for the RT pass
    A12_HIP(up(A.depth_lo, dset.data(), dset.size(), s));
    // compact copies of the (ld, *) arrays
    std::vector<double> tmp((size_t)n * std::max(a->mion, a->nwave1));
    for (int i = 0; i < a->nwave1; ++i)
        std::memcpy(&tmp[(size_t)i * n], a->tabcont + (size_t)i * a->ldtab, sizeof(double) * n);
    A12_HIP(up(A.tabcont, tmp.data(), (size_t)a->nwave1 * n, s));
    A12_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < a->mion; ++i)
        std::memcpy(&tmp[(size_t)i * n], a->xnfdop + (size_t)i * a->ldion, sizeof(double) * n);
    A12_HIP(up(A.xnfdop, tmp.data(), (size_t)a->mion * n, s));
    A12_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < a->mion; ++i)
        std::memcpy(&tmp[(size_t)i * n], a->dopple + (size_t)i * a->ldion, sizeof(double) * n);
    A12_HIP(up(A.dopple, tmp.data(), (size_t)a->mion * n, s));
    A12_HIP(up(A.waveset, a->waveset, numnu, s));
    A12_HIP(up(A.rcoset, a->rcoset, numnu, s));
    std::vector<double> tabs((size_t)3 * kTabN);
    std::memcpy(&tabs[0], a->h0tab, sizeof(double) * kTabN);
    std::memcpy(&tabs[kTabN], a->h1tab, sizeof(double) * kTabN);
    std::memcpy(&tabs[2 * kTabN], a->h2tab, sizeof(double) * kTabN);
    std::vector<float> tabs32(tabs.begin(), tabs.end());
    A12_HIP(up(A.tabs, tabs.data(), tabs.size(), s));
    A12_HIP(up(A.tabs32, tabs32.data(), tabs32.size(), s));
    // lines: unpack and prepare as LINOP1 does per record (:18176-18236)
    const double ratiolg = std::log(1.0 + 1.0 / 2000000.0);
    const double cen_prefac4 = 0.026538 / 1.7724538509055159 / 2.99792458e17;
    const double fourpi_c_inv = 1.0 / ((4.0 * 3.141592653589793) * 2.99792458e17);
    const double start = a->waveset[0] - 1.0, stop = a->waveset[numnu - 1] + 1.0;
    struct Rec {
        int nu, nucont, nel;
        double c[6];
    };
    std::vector<Rec> recs;
    recs.reserve(nlines);
    auto tablog = [](int16_t i) { return std::pow(10.0, (double)(i - 16384) * .001); };
    for (int i = 0; i < nlines; ++i) {
        const int32_t *r = linedata + 4 * (size_t)i;
        int16_t p[6];
        std::memcpy(p, r + 1, 12);
        const int iwl = r[0];
        // first bin with IWL < IWAVETAB, clamped to the last (:18193-18200)
        const int nucont = std::min(
            (int)(std::upper_bound(a->iwavetab, a->iwavetab + a->nwave1, iwl) - a->iwavetab),
            a->nwave1 - 1);
        const int nel = std::abs(p[0] / 10);
        if (nel < 1 || nel > a->mion) continue;
        const double wl = std::exp(iwl * ratiolg);
        if (wl < start || wl > stop) continue;
        const int nu = (int)(std::upper_bound(a->waveset, a->waveset + numnu, wl) - a->waveset);
        Rec q;
        q.nu = nu;
        q.nucont = nucont;
        q.nel = nel;
        q.c[0] = wl;
        q.c[1] = cen_prefac4 * wl * tablog(p[2]);
        q.c[2] = tablog(p[1]);
        q.c[3] = tablog(p[3]) * wl * fourpi_c_inv;
        q.c[4] = tablog(p[4]) * wl * fourpi_c_inv;
        q.c[5] = tablog(p[5]) * wl * fourpi_c_inv;
        recs.push_back(q);
    }
    std::stable_sort(recs.begin(), recs.end(), [](const Rec &x, const Rec &y) { return x.nu < y.nu; });
    const int nl = (int)recs.size();
    std::vector<int> l_nu(nl), l_nucont(nl), l_nel(nl), seg(numnu + 2, 0);
    std::vector<double> l_const((size_t)6 * std::max(nl, 1));
    for (int i = 0; i < nl; ++i) {
        l_nu[i] = recs[i].nu;
        l_nucont[i] = recs[i].nucont;
        l_nel[i] = recs[i].nel;
        for (int k = 0; k < 5; ++k) l_const[5 * (size_t)i + k] = recs[i].c[k];
        l_const[5 * (size_t)nl + i] = recs[i].c[5];
    }
    // seg[k] = first sorted line with nu >= k, k = 0..numnu+1
    for (int k = 0, i = 0; k <= numnu + 1; ++k) {
        while (i < nl && l_nu[i] < k) ++i;
        seg[k] = i;
    }
    A.nl = nl;
    A12_HIP(up(A.l_nu, l_nu.data(), nl, s));
    A12_HIP(up(A.l_nucont, l_nucont.data(), nl, s));
    A12_HIP(up(A.l_nel, l_nel.data(), nl, s));
    A12_HIP(up(A.l_const, l_const.data(), (size_t)6 * nl, s));
    A12_HIP(up(A.seg, seg.data(), seg.size(), s));
    A12_HIP(hipStreamSynchronize(s));
    A.loaded = true;
    A.xl_ready = false;
    return A12_OK;
}

int a12i_atlas_linop1_run(a12_ctx *ctx, int prec)
{
    AtlasDev &A = ctx->atl;
    if (!A.loaded) {
        ctx->err = "a12_atlas_linop1_run: call a12_atlas_load first";
        return A12_ERR_STATE;
    }
    if (prec != A12_PREC_MIXED && prec != A12_PREC_FP64) {
        ctx->err = "a12_atlas_linop1_run: bad prec";
        return A12_ERR_ARG;
    }
    hipStream_t s = ctx->stream;
    const int n = A.n, numnu = A.numnu, nl = A.nl;
    const size_t nl1 = std::max(nl, 1);
    A12_HIP(A.state.reserve(sizeof(int) * nl1 * n));
    A12_HIP(A.par.reserve(sizeof(double) * 3 * nl1 * n));
    A12_HIP(A.reach.reserve(sizeof(int) * (size_t)(numnu + 1) * n));
    A12_HIP(A.xlines.reserve(sizeof(double) * (size_t)numnu * n));
    const int nb = (int)((nl1 + 255) / 256);
    A12_HIP(A.blk.reserve(sizeof(long long) * nb));
    A12_HIP(A.evals.reserve(sizeof(long long)));
    const double *tabs = A.tabs.as<double>();
    const float *tabs32 = A.tabs32.as<float>();
    if (ctx->prof) A12_HIP(hipEventRecord(ctx->ev[8], s));
    if (nl > 0) {
        hipLaunchKernelGGL(k_lo_setup, dim3(nb), dim3(256), 0, s, nl, n, numnu, A.mion,
                           A.l_nu.as<int>(), A.l_nucont.as<int>(), A.l_nel.as<int>(),
                           A.l_const.as<double>(), A.waveset.as<double>(), A.tabcont.as<double>(),
                           A.xnfdop.as<double>(), A.dopple.as<double>(), A.depth_lo.as<double>(),
                           tabs, tabs + kTabN, tabs + 2 * kTabN, A.state.as<int>(),
                           A.par.as<double>(), A.blk.as<long long>());
        A12_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_lo_sum, dim3(1), dim3(256), 0, s, A.blk.as<long long>(), nb,
                           A.evals.as<long long>());
        A12_CHECK_LAUNCH();
    } else {
        A12_HIP(hipMemsetAsync(A.evals.p, 0, sizeof(long long), s));
    }
    if (ctx->prof) A12_HIP(hipEventRecord(ctx->ev[9], s));
    const int nseg = numnu + 1;
    hipLaunchKernelGGL(k_lo_reach, dim3((nseg * n + 255) / 256), dim3(256), 0, s, nseg, n,
                       A.seg.as<int>(), A.state.as<int>(), A.reach.as<int>());
    A12_CHECK_LAUNCH();
    if (ctx->prof) A12_HIP(hipEventRecord(ctx->ev[10], s));
    auto g = prec == A12_PREC_FP64 ? k_lo_gather<true> : k_lo_gather<false>;
    hipLaunchKernelGGL(g, dim3(numnu), dim3(128), 0, s, n, numnu, A.seg.as<int>(), A.reach.as<int>(),
                       A.state.as<int>(), A.par.as<double>(), A.l_const.as<double>(),
                       A.waveset.as<double>(), tabs, tabs + kTabN, tabs + 2 * kTabN, tabs32,
                       tabs32 + kTabN, tabs32 + 2 * kTabN, A.xlines.as<double>());
    A12_CHECK_LAUNCH();
    if (ctx->prof) A12_HIP(hipEventRecord(ctx->ev[11], s));
    A.xl_ready = true;
    A.prof_lo = ctx->prof;
    return A12_OK;
}
