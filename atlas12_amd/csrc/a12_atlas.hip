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
//   k_lo_setup    one lane per (line, depth) applies the reference's
//                 coarse (every 8th) / fine depth rule and fixes kappa_centre,
//                 a, the Doppler width in nm and the red/blue reach (FP64
//                 decisions, as the reference; the Lorentz tail by bisection);
//   k_lo_gather   tiles of 8 grid points x all depths per block: the sum of
//                 every reaching line's profile, streaming the candidate
//                 lines through LDS in NU-sorted store order (for a
//                 wavelength-sorted store, the reference's own line
//                 order).  No atomics.
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

// The same wing value with the arithmetic in FP32 but every regime and
// table-row decision taken from the FP64 offset v64 and damping a64, as the
// FP64 path takes them: a row index from a rounded v flips to the
// neighbouring row near integer 200 v + 1.5 (~0.5 % of that contribution),
// and a rounded a at the 0.2 boundary switches between the table and
// VOIGT (~4 %).
__device__ __forceinline__ float lo_cv_mixed(float center, double v64, double a64, const float *h0,
                                             const float *h1, const float *h2)
{
#pragma clang fp contract(off)
    const float v = (float)v64, a = (float)a64;
    if (a64 > 0.20) {
        int iv = (int)(v64 * 200.0 + 1.5);
        iv = iv < 1 ? 1 : (iv > kTabN ? kTabN : iv);
        const float H0 = h0[iv - 1], H1 = h1[iv - 1], H2 = h2[iv - 1];
        if (a64 <= 1.4 && a64 + v64 <= 3.2) {
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
        if (a64 <= 100.0) {
            const float aau = aa / u, vvu = vv / u, uu = u * u;
            r = ((((aau - 10.0f * vvu) * aau * 3.0f + 15.0f * vvu * vvu) + 3.0f * vv - aa) / uu + 1.0f) * r;
        }
        return center * r;
    }
    if (v64 > 10.0) return center * 0.5641895835477563f * a / (v * v);
    const int iv = (int)(v64 * 200.0 + 1.5);
    return center * ((h2[iv - 1] * a + h1[iv - 1]) * a + h0[iv - 1]);
}

// Number of wing points LINOP1 adds on one side before and including the
// first value below tc: points base+sgn, base+2 sgn, ... (red: offsets
// w - wl; blue: wl - w), at most kMaxWing or up to the grid end.  Table and
// VOIGT regimes are walked point by point; once a <= 0.2 and v > 10 the
// value is center * INVSQRTPI * a / v^2, which never increases along the
// walk in FP64 (v = (w - wl) / dopwave is monotone in w under rounding and
// so is each later operation), so the first point below tc is found by
// bisection with the same predicate the walk would test.
__device__ int lo_reach(int first, int last, int step, const double *__restrict__ waveset,
                        double wlvac, double dopwave, double center, double adamp, double tc,
                        const double *h0, const double *h1, const double *h2)
{
#pragma clang fp contract(off)
    int cnt = 0;
    int iw = first;
    const int nmax = step > 0 ? last - first + 1 : first - last + 1;
    if (nmax <= 0) return 0;
    for (; cnt < nmax; ++cnt, iw += step) {
        const double v = step > 0 ? (waveset[iw] - wlvac) / dopwave : (wlvac - waveset[iw]) / dopwave;
        if (adamp <= 0.20 && v > 10.0) break;  // monotone Lorentz tail from here
        if (lo_cv<double>(center, v, adamp, h0, h1, h2) < tc) return cnt + 1;
    }
    if (cnt == nmax) return cnt;
    // bisection over walk positions [cnt, nmax): first with value < tc
    const double k = center * 0.5641895835477563 * adamp;
    int lo = cnt, hi = nmax;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const int pw = first + mid * step;
        const double v = step > 0 ? (waveset[pw] - wlvac) / dopwave : (wlvac - waveset[pw]) / dopwave;
        if (k / (v * v) < tc) hi = mid;
        else lo = mid + 1;
    }
    return lo < nmax ? lo + 1 : nmax;
}

// One lane per (line, depth).  LINOP1 computes the coarse depths J = 8, 16,
// ... first, then each fine depth J in K-7..K-1 only where coarse depth K-8
// or K passed both centre tests (IFJ, :18239-18365); the decision depends on
// nothing else, so a fine lane re-tests its two coarse neighbours itself.
// o_state[l*n + j]: bit 31 computed, bits 0-7 red reach, 8-15 blue reach
// (points nu.. and nu-1..); o_par: centre, a, Doppler width in nm.
__device__ __forceinline__ bool lo_center_ok(double cgf, double elo, double xd, double tc,
                                             double hckt, double *center)
{
#pragma clang fp contract(off)
    double c = cgf * xd;
    if (c < tc) return false;
    c = c * exp(-elo * hckt);
    if (c < tc) return false;
    *center = c;
    return true;
}

__global__ __launch_bounds__(256) void k_lo_setup(
    int nl, int n, int numnu, const int *__restrict__ l_nu, const int *__restrict__ l_nucont,
    const int *__restrict__ l_nel, const double *__restrict__ l_const,
    const double *__restrict__ waveset, const double *__restrict__ tabcont,
    const double *__restrict__ xnfdop, const double *__restrict__ dopple,
    const double *__restrict__ dep, const double *__restrict__ h0, const double *__restrict__ h1,
    const double *__restrict__ h2, int *__restrict__ o_state, double *__restrict__ o_par,
    long long *__restrict__ blk_evals)
{
#pragma clang fp contract(off)
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long ev = 0;
    if (g < (long long)nl * n) {
        const int l = (int)(g / n), j = (int)(g - (long long)l * n);
        const int J = j + 1, ncoarse = (n / 8) * 8;
        int st = 0;
        if (J <= ncoarse) {
            const int nu = l_nu[l], nucont = l_nucont[l], nel = l_nel[l] - 1;
            const double cgf = l_const[5 * (size_t)l + 1], elo = l_const[5 * (size_t)l + 2];
            const double *xd = xnfdop + (size_t)nel * n, *tcv = tabcont + (size_t)nucont * n;
            const double *hckt = dep + 3 * n;
            bool on = true;
            if (J % 8 != 0) {  // fine depth: a neighbouring coarse depth must have passed
                const int K = (J / 8 + 1) * 8;
                double c;
                on = (K - 8 >= 8 && lo_center_ok(cgf, elo, xd[K - 9], tcv[K - 9], hckt[K - 9], &c)) ||
                     lo_center_ok(cgf, elo, xd[K - 1], tcv[K - 1], hckt[K - 1], &c);
            }
            double center;
            const double tc = tcv[j];
            if (on && lo_center_ok(cgf, elo, xd[j], tc, hckt[j], &center)) {
                const double wlvac = l_const[5 * (size_t)l + 0], gr = l_const[5 * (size_t)l + 3],
                             gs = l_const[5 * (size_t)l + 4], gw = l_const[5 * (size_t)nl + l];
                const double dop = dopple[(size_t)nel * n + j];
                const double adamp = (gr + gs * dep[2 * n + j] + gw * dep[5 * n + j]) / dop;
                const double dopwave = dop * wlvac;
                // red: IW = NU .. min(NU+100, NUMNU); blue: NU-1 down, at most 100
                const int rn = lo_reach(nu, min(nu + kMaxWing, numnu - 1), 1, waveset, wlvac,
                                        dopwave, center, adamp, tc, h0, h1, h2);
                const int bn = lo_reach(nu - 1, max(nu - kMaxWing, 0), -1, waveset, wlvac,
                                        dopwave, center, adamp, tc, h0, h1, h2);
                st = (int)(0x80000000u | (unsigned)rn | ((unsigned)bn << 8));
                o_par[3 * g + 0] = center;
                o_par[3 * g + 1] = adamp;
                o_par[3 * g + 2] = dopwave;
                ev = rn + bn;
            }
        }
        o_state[g] = st;
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

// Gather: a block holds tiles of kLoPts consecutive grid points; lane t
// serves tile t / n and depth t % n.  The lines that can reach the block's
// points are one contiguous run of the NU-sorted store (segments NU =
// first - 100 .. last + 100); the block streams it in batches of kLoBatch
// lines, staging their per-depth state in LDS with coalesced loads, and
// each lane takes from every staged line the points of its tile that the
// line's red/blue reach covers at its depth.  Lines are visited in store
// order, which for every grid point is the order LINOP1 adds them, into the
// lane's own LDS column: no atomics, deterministic.
constexpr int kLoPts = 8;
constexpr int kLoBlock = 512;
constexpr int kLoBatch = 32;

template <bool F64>
__global__ __launch_bounds__(kLoBlock) void k_lo_gather(
    int n, int numnu, int tiles, const int *__restrict__ seg, const int *__restrict__ l_nu,
    const int *__restrict__ state, const double *__restrict__ par,
    const double *__restrict__ l_const, const double *__restrict__ waveset,
    const double *__restrict__ h0, const double *__restrict__ h1, const double *__restrict__ h2,
    const float *__restrict__ f0, const float *__restrict__ f1, const float *__restrict__ f2,
    double *__restrict__ xlines)
{
#pragma clang fp contract(off)
    // LDS: F64 sums, or FP32 sums + compensations (same bytes)
    __shared__ double acc_d[kLoPts * kLoBlock];
    __shared__ int s_state[kLoBatch * A12_NRMAX_LAYERS];
    __shared__ int s_nu[kLoBatch];
    __shared__ double s_wl[kLoBatch];
    float *acc_s = reinterpret_cast<float *>(acc_d);
    float *acc_c = acc_s + kLoPts * kLoBlock;
    const int per = kLoBlock / n;  // tiles per block
    const int t = threadIdx.x;
    const int lt = t / n, j = t - lt * n;
    const int tile0 = blockIdx.x * per;
    const int tile = tile0 + lt;
    const bool live = lt < per && tile < tiles;
    const int iw0 = tile * kLoPts, iw1 = min(iw0 + kLoPts, numnu) - 1;
    for (int p = 0; p < kLoPts; ++p) {
        if (F64) acc_d[p * kLoBlock + t] = 0.0;
        else acc_s[p * kLoBlock + t] = acc_c[p * kLoBlock + t] = 0.0f;
    }
    // the block's candidate lines: segments NU in [first - 100, last + 100]
    // (red from NU reaches NU..NU+100, blue NU-100..NU-1)
    const int bfirst = tile0 * kLoPts, blast = min((tile0 + per) * kLoPts, numnu) - 1;
    const int lbeg = seg[max(0, bfirst - kMaxWing)], lend = seg[min(numnu, blast + kMaxWing) + 1];
    for (int b0 = lbeg; b0 < lend; b0 += kLoBatch) {
        const int nb = min(kLoBatch, lend - b0);
        __syncthreads();
        for (int i = t; i < nb * n; i += kLoBlock) s_state[i] = state[(size_t)b0 * n + i];
        if (t < nb) {
            s_nu[t] = l_nu[b0 + t];
            s_wl[t] = l_const[5 * (size_t)(b0 + t)];
        }
        __syncthreads();
        if (!live) continue;
        for (int i = 0; i < nb; ++i) {
            const int st = s_state[i * n + j];
            if (st >= 0) continue;
            const int k = s_nu[i], rn = st & 0xff, bn = (st >> 8) & 0xff;
            const int ra = max(k, iw0), rb = min(k + rn - 1, iw1);
            const int ba = max(k - bn, iw0), bb = min(k - 1, iw1);
            if (ra > rb && ba > bb) continue;
            const size_t o = (size_t)(b0 + i) * n + j;
            const double wl = s_wl[i];
            const double center = par[3 * o], adamp = par[3 * o + 1], dopwave = par[3 * o + 2];
            // a point lies on one side of the line only: the ranges are disjoint
            for (int side = 0; side < 2; ++side) {
                const int a = side ? ba : ra, b = side ? bb : rb;
                for (int iw = a; iw <= b; ++iw) {
                    const double v = (side ? (wl - waveset[iw]) : (waveset[iw] - wl)) / dopwave;
                    const int q = (iw - iw0) * kLoBlock + t;
                    if (F64) {
                        acc_d[q] = acc_d[q] + lo_cv<double>(center, v, adamp, h0, h1, h2);
                    } else {
                        const float x = lo_cv_mixed((float)center, v, adamp, f0, f1, f2);
                        const float s0 = acc_s[q], tt = s0 + x;
                        acc_c[q] += fabsf(s0) >= fabsf(x) ? (s0 - tt) + x : (x - tt) + s0;
                        acc_s[q] = tt;
                    }
                }
            }
        }
    }
    if (live)
        for (int iw = iw0; iw <= iw1; ++iw) {
            const int q = (iw - iw0) * kLoBlock + t;
            xlines[(size_t)iw * n + j] = F64 ? acc_d[q] : (double)acc_s[q] + (double)acc_c[q];
        }
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
    const int nb = (int)((nl1 * n + 255) / 256);  // one lane per (line, depth)
    A12_HIP(A.blk.reserve(sizeof(long long) * nb));
    A12_HIP(A.evals.reserve(sizeof(long long)));
    const double *tabs = A.tabs.as<double>();
    const float *tabs32 = A.tabs32.as<float>();
    if (ctx->prof) A12_HIP(hipEventRecord(ctx->ev[8], s));
    if (nl > 0) {
        hipLaunchKernelGGL(k_lo_setup, dim3(nb), dim3(256), 0, s, nl, n, numnu,
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
    if (ctx->prof) A12_HIP(hipEventRecord(ctx->ev[10], s));
    auto g = prec == A12_PREC_FP64 ? k_lo_gather<true> : k_lo_gather<false>;
    const int tiles = (numnu + kLoPts - 1) / kLoPts, per = kLoBlock / n;
    hipLaunchKernelGGL(g, dim3((tiles + per - 1) / per), dim3(kLoBlock), 0, s, n, numnu, tiles,
                       A.seg.as<int>(), A.l_nu.as<int>(),
                       A.state.as<int>(), A.par.as<double>(), A.l_const.as<double>(),
                       A.waveset.as<double>(), tabs, tabs + kTabN, tabs + 2 * kTabN, tabs32,
                       tabs32 + kTabN, tabs32 + 2 * kTabN, A.xlines.as<double>());
    A12_CHECK_LAUNCH();
    if (ctx->prof) A12_HIP(hipEventRecord(ctx->ev[11], s));
    A.xl_ready = true;
    A.prof_lo = ctx->prof;
    return A12_OK;
}
