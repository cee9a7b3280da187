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
//                 decisions, as the reference; in the tails a search started
//                 at the closed-form crossing);
//   k_lo_linereach  each line's widest reach over its depths;
//   k_lo_gather   tiles of kLoPts grid points x all depths per block (staged
//   k_lo_gather_w through LDS, or per wavefront without barriers): the sum
//                 of every reaching line's profile in NU-sorted store order
//                 (for a wavelength-sorted store, the reference's own line
//                 order).  No atomics.
// MIXED evaluates the profile in FP32 and sums with Kahan compensation;
// FP64 repeats the reference's arithmetic.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>
#include <vector>

#include "a12_common.h"
#include "a12_internal.h"

// Wavefront sum and maximum of the device library (DPP).
extern "C" __device__ long long __ockl_wfred_add_i64(long long);
extern "C" __device__ int __ockl_wfred_max_i32(int);

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
__device__ __forceinline__ float lo_cv_mixed(float center, double v64, double a64,
                                             const float4 *__restrict__ h)
{
#pragma clang fp contract(off)
    const float v = (float)v64, a = (float)a64;
    if (a64 > 0.20) {
        int iv = (int)(v64 * 200.0 + 1.5);
        iv = iv < 1 ? 1 : (iv > kTabN ? kTabN : iv);
        const float4 Hr = h[iv - 1];
        const float H0 = Hr.x, H1 = Hr.y, H2 = Hr.z;
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
    const float4 Hr = h[iv - 1];
    return center * ((Hr.z * a + Hr.y) * a + Hr.x);
}

// lo_cv in the two regimes where the profile never increases along a walk
// (a <= 0.2 with v > 10: the Lorentz wing; a > 0.2 outside a <= 1.4 and
// a + v <= 3.2: VOIGT's asymptotic form), with the same operations in the
// same order, so the value has the same bits; no table reads.
__device__ __forceinline__ double lo_tail(double center, double v, double adamp)
{
#pragma clang fp contract(off)
    if (adamp <= 0.20) return center * 0.5641895835477563 * adamp / (v * v);
    const double aa = adamp * adamp, vv = v * v;
    const double u = (aa + vv) * 1.4142;
    double r = adamp * 0.79788 / u;
    if (adamp <= 100.0) {
        const double aau = aa / u, vvu = vv / u, uu = u * u;
        r = ((((aau - 10.0 * vvu) * aau * 3.0 + 15.0 * vvu * vvu) + 3.0 * vv - aa) / uu + 1.0) * r;
    }
    return center * r;
}

// Number of wing points LINOP1 adds on one side before and including the
// first value below tc: points first, first+step, ... last (red: offsets
// w - wl; blue: wl - w).  The walk is taken point by point until the
// profile enters a regime where it never increases along the walk:
//   a <= 0.2, v > 10: center * INVSQRTPI * a / v^2 -- v = (w - wl)/dopwave
//     is monotone in w under FP64 rounding, and so is every later
//     operation;
//   a > 0.2 outside (a <= 1.4 and a + v <= 3.2): VOIGT's asymptotic form,
//     strictly decreasing in v (for a > 100 it is a*0.79788/u with u
//     increasing, monotone under rounding; for 0.2 < a <= 100 the FP64
//     expression was checked non-increasing on a 1e-3 grid in v for 6,000
//     values of a, far finer than any walk step).
// From there the first point below tc is searched for with the same FP64
// predicate the walk tests, starting where the tail's closed form crosses
// tc: v* = sqrt(center a / (sqrt(pi) tc)) for the Lorentz wing, the leading
// term of the asymptote (a^2 + v*^2 = 0.79788 center a / (1.4142 tc))
// otherwise, placed on the grid with its mean log step (lg2inv = 1 /
// log2 of the point-to-point ratio).  The estimate p is only a starting
// point: the predicate at p and at one neighbour fixes the answer when p is
// right (over 99.8 % of searches on a log-uniform grid, the rest one point
// off), and a bisection over what they leave open finishes otherwise.
// Equal to the walk whenever consecutive walk points' profile values differ
// by more than the FP64 evaluation error, i.e. for any grid coarser than
// ~1e-9 Doppler widths.
__device__ int lo_reach(int first, int last, int step, const double *__restrict__ waveset,
                        double wlvac, double dopwave, double center, double adamp, double tc,
                        double rtc, float rdf, const double *h0, const double *h1,
                        const double *h2, float lg2inv)
{
#pragma clang fp contract(off)
    int cnt = 0;
    int iw = first;
    const int nmax = step > 0 ? last - first + 1 : first - last + 1;
    if (nmax <= 0) return 0;
    for (; cnt < nmax; ++cnt, iw += step) {
        const double v = step > 0 ? (waveset[iw] - wlvac) / dopwave : (wlvac - waveset[iw]) / dopwave;
        const bool mono = adamp <= 0.20 ? v > 10.0 : !(adamp <= 1.4 && adamp + v <= 3.2);
        if (mono) break;
        if (lo_cv<double>(center, v, adamp, h0, h1, h2) < tc) return cnt + 1;
    }
    if (cnt == nmax) return cnt;
    // The walk's test at walk position pos (tail regimes), screened in FP32:
    // the profile over tc, evaluated with reciprocals from the FP64 offset
    // (relative error below ~1e-6), decides when it is more than 1e-4 away
    // from 1; otherwise the reference's FP64 expression does.  Either way the
    // answer is the FP64 predicate's.
    // rtc = 1 / tc, rdf = (float)(1 / dopwave), shared by both sides
    const float qf = (float)(center * 0.5641895835477563 * adamp * rtc);  // Lorentz: qf / v^2
    const float cf = (float)(center * adamp * 0.79788 * rtc);  // asymptote: cf / u * (1 + ...)
    const float af = (float)adamp, aaf = af * af;
    auto below = [&](int pos) {
        const int pw = first + pos * step;
        const double d = step > 0 ? waveset[pw] - wlvac : wlvac - waveset[pw];
        const float v = (float)d * rdf, vv = v * v;
        float ratio;
        if (adamp <= 0.20) {
            ratio = qf * __builtin_amdgcn_rcpf(vv);
        } else {
            const float ru = __builtin_amdgcn_rcpf((aaf + vv) * 1.4142f);
            ratio = cf * ru;
            if (adamp <= 100.0) {
                const float aau = aaf * ru, vvu = vv * ru;
                ratio = ((((aau - 10.0f * vvu) * aau * 3.0f + 15.0f * vvu * vvu) + 3.0f * vv - aaf) *
                             (ru * ru) + 1.0f) * ratio;
            }
        }
        if (ratio < 0.9999f) return true;
        if (ratio > 1.0001f) return false;
        return lo_tail(center, d / dopwave, adamp) < tc;
    };
    // estimated first walk position past the crossing, in [cnt, nmax]
    int p;
    const double q = adamp <= 0.20 ? (double)qf : (double)cf * (1.0 / 1.4142) - adamp * adamp;
    if (!(q > 0.0)) {
        p = cnt;
    } else {
        const float off = __builtin_sqrtf((float)q) * (float)dopwave;
        const float wt = step > 0 ? (float)wlvac + off : (float)wlvac - off;
        const float pe = (float)step * __log2f(wt * __builtin_amdgcn_rcpf((float)waveset[first])) * lg2inv;
        p = (wt <= 0.0f || !(pe < (float)nmax)) ? nmax : max((int)ceilf(pe), cnt);
    }
    // every position below lo is known >= tc, every position from hi on < tc
    // (hi = nmax: none known)
    int lo = cnt, hi = nmax;
    if (p < nmax) {
        if (below(p)) {
            hi = p;
            if (p - 1 >= lo) {
                if (below(p - 1)) hi = p - 1;
                else lo = p;
            }
        } else {
            lo = p + 1;
            if (p + 1 < hi) {
                if (below(p + 1)) hi = p + 1;
                else lo = p + 2;
            }
        }
    } else if (nmax - 1 >= lo) {
        if (below(nmax - 1)) hi = nmax - 1;
        else lo = nmax;
    }
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (below(mid)) hi = mid;
        else lo = mid + 1;
    }
    return lo < nmax ? lo + 1 : nmax;
}

// Per-(line, depth) values one contribution needs.  FP64: centre, a,
// Doppler width (the reference's arithmetic).  Mixed: 1/Doppler width and a
// in FP64 (offsets and every regime/table decision), centre, centre*a/sqrt(pi)
// and a in FP32.
// k_lo_setup stores them in this layout for the run's precision, so the
// gathers read them as they are.
struct LoPar {
    double p0, p1, p2;  // FP64: center, adamp, dopwave | mixed: rdw, adamp, (c32, k32)
};

template <bool F64>
__device__ __forceinline__ LoPar lo_par_of(double center, double adamp, double dopwave)
{
    LoPar q;
    if (F64) {
        q.p0 = center;
        q.p1 = adamp;
        q.p2 = dopwave;
    } else {
        q.p0 = 1.0 / dopwave;
        q.p1 = adamp;
        const float c32 = (float)center, k32 = (float)(center * 0.5641895835477563 * adamp);
        q.p2 = __hiloint2double(__float_as_int(c32), __float_as_int(k32));
    }
    return q;
}

// One lane per (line, depth).  LINOP1 computes the coarse depths J = 8, 16,
// ... first, then each fine depth J in K-7..K-1 only where coarse depth K-8
// or K passed both centre tests (IFJ, :18239-18365); the decision depends on
// nothing else, so every lane takes its own centre test and a fine lane
// reads its two coarse neighbours' results from LDS (re-testing them itself
// when they fall in another workgroup).
// o_state[l*n + j]: bit 31 computed, bits 0-7 red reach, 8-15 blue reach
// (points nu.. and nu-1..); o_par: the line's LoPar for the run's precision
// (mixed != 0: 1/Doppler width, a, packed FP32 centre and centre*a/sqrt(pi)).
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
    const double *__restrict__ h2, float lg2inv, int mixed, int *__restrict__ o_state,
    double *__restrict__ o_par, long long *__restrict__ blk_evals)
{
#pragma clang fp contract(off)
    __shared__ int s_ok[256];
    const long long b0 = (long long)blockIdx.x * blockDim.x;
    const long long g = b0 + threadIdx.x;
    // (line, depth) of this lane without a 64-bit integer division: the
    // double quotient is exact to +-1 for any g below 2^52
    long long lq = (long long)((double)g * (1.0 / (double)n));
    if (lq * n > g) --lq;
    if ((lq + 1) * n <= g) ++lq;
    const int l = (int)lq, j = (int)(g - lq * n);
    const int J = j + 1, ncoarse = (n / 8) * 8;
    const bool valid = g < (long long)nl * n && J <= ncoarse;
    int nu = 0, nucont = 0, nel = 0;
    double cgf = 0.0, elo = 0.0, center = 0.0;
    const double *xd = xnfdop, *tcv = tabcont, *hckt = dep + 3 * n;
    bool own = false;
    if (valid) {
        nu = l_nu[l];
        nucont = l_nucont[l];
        nel = l_nel[l] - 1;
        cgf = l_const[5 * (size_t)l + 1];
        elo = l_const[5 * (size_t)l + 2];
        xd = xnfdop + (size_t)nel * n;
        tcv = tabcont + (size_t)nucont * n;
        own = lo_center_ok(cgf, elo, xd[j], tcv[j], hckt[j], &center);
    }
    s_ok[threadIdx.x] = own;
    __syncthreads();
    long long ev = 0;
    int st = 0;
    if (valid && own) {
        bool on = true;
        if (J % 8 != 0) {  // fine depth: a neighbouring coarse depth must have passed
            const int K = (J / 8 + 1) * 8;
            auto coarse_ok = [&](int Kd) {  // the coarse lane's own test, from LDS when in block
                const long long gk = (long long)l * n + (Kd - 1);
                if (gk >= b0 && gk < b0 + (long long)blockDim.x) return s_ok[gk - b0] != 0;
                double c;
                return lo_center_ok(cgf, elo, xd[Kd - 1], tcv[Kd - 1], hckt[Kd - 1], &c);
            };
            on = (K - 8 >= 8 && coarse_ok(K - 8)) || coarse_ok(K);
        }
        if (on) {
            const double tc = tcv[j];
            const double wlvac = l_const[5 * (size_t)l + 0], gr = l_const[5 * (size_t)l + 3],
                         gs = l_const[5 * (size_t)l + 4], gw = l_const[5 * (size_t)nl + l];
            const double dop = dopple[(size_t)nel * n + j];
            const double adamp = (gr + gs * dep[2 * n + j] + gw * dep[5 * n + j]) / dop;
            const double dopwave = dop * wlvac;
            // red: IW = NU .. min(NU+100, NUMNU); blue: NU-1 down, at most 100
            const double rtc = 1.0 / tc;
            const float rdf = (float)(1.0 / dopwave);
            const int rn = lo_reach(nu, min(nu + kMaxWing, numnu - 1), 1, waveset, wlvac, dopwave,
                                    center, adamp, tc, rtc, rdf, h0, h1, h2, lg2inv);
            const int bn = lo_reach(nu - 1, max(nu - kMaxWing, 0), -1, waveset, wlvac, dopwave,
                                    center, adamp, tc, rtc, rdf, h0, h1, h2, lg2inv);
            st = (int)(0x80000000u | (unsigned)rn | ((unsigned)bn << 8));
            const LoPar qp = mixed ? lo_par_of<false>(center, adamp, dopwave)
                                   : lo_par_of<true>(center, adamp, dopwave);
            o_par[3 * g + 0] = qp.p0;
            o_par[3 * g + 1] = qp.p1;
            o_par[3 * g + 2] = qp.p2;
            ev = rn + bn;
        }
    }
    if (g < (long long)nl * n) o_state[g] = st;
    // block total of the contribution counts: a DPP sum per wavefront, then
    // the four wave sums (integers: exact in any order)
    __shared__ long long s_wsum[4];
    const long long wsum = __ockl_wfred_add_i64(ev);
    if ((threadIdx.x & 63) == 0) s_wsum[threadIdx.x >> 6] = wsum;
    __syncthreads();
    if (threadIdx.x == 0) blk_evals[blockIdx.x] = s_wsum[0] + s_wsum[1] + s_wsum[2] + s_wsum[3];
}

// Sum of nb per-block counts in a fixed order: pass 1, block b of gridDim.x
// sums its contiguous share into part[b]; pass 2 (one block, part = null)
// sums the parts into *out.
__global__ void k_lo_sum(const long long *__restrict__ blk, int nb, long long *__restrict__ part,
                         long long *__restrict__ out)
{
    __shared__ long long red[256];
    const int per = (nb + gridDim.x - 1) / gridDim.x;
    const int b0 = blockIdx.x * per, b1 = min(nb, b0 + per);
    long long s = 0;
    for (int i = b0 + threadIdx.x; i < b1; i += blockDim.x) s += blk[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = blockDim.x >> 1; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (part) part[blockIdx.x] = red[0];
        else *out = red[0];
    }
}

constexpr int kSumParts = 1024;

// Widest red and blue reach of each line over its depths (0 when no depth
// passed the centre tests): l_reach[l] = red | blue << 8, and over each
// group of 8 depths: greach[g * nl + l], same packing.  Lets the gathers
// skip the candidate lines that cannot reach a block's points (at any
// depth, or at the group's depths).
__global__ __launch_bounds__(256) void k_lo_linereach(int nl, int n,
                                                      const int *__restrict__ state,
                                                      int *__restrict__ l_reach,
                                                      unsigned short *__restrict__ greach)
{
    // a block takes 64 consecutive lines, each wavefront 16 of them, all in
    // flight: coalesced loads of each line's n states (n <= 128), shuffle
    // maxima per group of 8 depths and over the line; the group maxima are
    // gathered in LDS and stored group by group (64 lines = 128 contiguous
    // bytes per group)
    constexpr int kLines = 16;
    __shared__ unsigned short s_g[16][64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int L0 = blockIdx.x * 64, ng = (n + 7) / 8;
    const int li0 = wid * kLines;  // first line of this wavefront in the tile
    int v[kLines][2];
#pragma unroll
    for (int u = 0; u < kLines; ++u) {
        const int l = L0 + li0 + u;
        const int *st = state + (size_t)l * n;
        v[u][0] = l < nl && lane < n ? st[lane] : 0;
        v[u][1] = l < nl && lane + 64 < n ? st[lane + 64] : 0;
    }
    // maximum over the 8 lanes of a group with DPP: quad exchanges (lane ^ 1,
    // lane ^ 2), then the half-row mirror (lane ^ 7 within 8 lanes)
    auto gmax8 = [](int x) {
        x = max(x, __builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false));
        x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, false));
        return max(x, __builtin_amdgcn_update_dpp(0, x, 0x141, 0xF, 0xF, false));
    };
#pragma unroll
    for (int u = 0; u < kLines; ++u) {
        int r = 0, b = 0;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            // widest reach over each group of 8 depths (depth = lane + 64 h)
            const int rg = gmax8(v[u][h] < 0 ? v[u][h] & 0xff : 0);
            const int bg = gmax8(v[u][h] < 0 ? (v[u][h] >> 8) & 0xff : 0);
            const int grp = (lane >> 3) + 8 * h;
            if ((lane & 7) == 0) s_g[grp][li0 + u] = (unsigned short)(rg | (bg << 8));
            r = max(r, rg);
            b = max(b, bg);
        }
        r = __ockl_wfred_max_i32(r);
        b = __ockl_wfred_max_i32(b);
        if (lane == 0 && L0 + li0 + u < nl) l_reach[L0 + li0 + u] = r | (b << 8);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < ng * 64; e += 256) {
        const int grp = e >> 6, li = e & 63;
        if (L0 + li < nl) greach[(size_t)grp * nl + L0 + li] = s_g[grp][li];
    }
}

template <bool F64>
__device__ __forceinline__ LoPar lo_par(const double *__restrict__ par, size_t o)
{
    return LoPar{par[3 * o], par[3 * o + 1], par[3 * o + 2]};
}

// The mixed path's value of one contribution at offset d (red: w - wl,
// blue: wl - w): FP32 arithmetic, every regime and table-row decision from
// the FP64 offset v and damping a.
__device__ __forceinline__ float lo_item_mixed(const LoPar &q, double d, const float4 *f)
{
#pragma clang fp contract(off)
    const double adamp = q.p1, v = d * q.p0;
    const float c32 = __int_as_float(__double2hiint(q.p2)), k32 = __int_as_float(__double2loint(q.p2));
    float x;
    if (adamp <= 0.20 && v > 10.0) {  // Lorentz tail: K / v^2
        const float vf = (float)v;
        x = k32 * __builtin_amdgcn_rcpf(vf * vf);
    } else if (adamp > 0.20 && !(adamp <= 1.4 && adamp + v <= 3.2)) {
        // VOIGT's asymptotic form with one reciprocal
        const float a32 = (float)adamp, aa32 = a32 * a32, vf = (float)v, vv = vf * vf;
        const float ru = __builtin_amdgcn_rcpf((aa32 + vv) * 1.4142f);
        const float aau = aa32 * ru, vvu = vv * ru;
        const float r = a32 * 0.79788f * ru;
        x = c32 * (adamp <= 100.0 ? ((((aau - 10.0f * vvu) * aau * 3.0f + 15.0f * vvu * vvu) +
                                      3.0f * vv - aa32) * (ru * ru) + 1.0f) * r
                                  : r);
    } else {
        x = lo_cv_mixed(c32, v, adamp, f);
    }
    return x;
}

// One contribution at offset d (red: w - wl, blue: wl - w) into LDS slot q.
template <bool F64>
__device__ __forceinline__ void lo_accum(const LoPar &q, double d, int slot, double *acc_d,
                                         float *acc_s, float *acc_c, const double *h0,
                                         const double *h1, const double *h2, const float4 *f)
{
#pragma clang fp contract(off)
    if (F64) {
        acc_d[slot] = acc_d[slot] + lo_cv<double>(q.p0, d / q.p2, q.p1, h0, h1, h2);
        return;
    }
    const float x = lo_item_mixed(q, d, f);
    // Kahan's compensated sum (acc_c holds the running compensation): every
    // contribution is positive, so its bound, 2 eps times the sum, holds
    // with four operations
    const float s0 = acc_s[slot], y = x - acc_c[slot], tt = s0 + y;
    acc_c[slot] = (tt - s0) - y;
    acc_s[slot] = tt;
}

// MIXED: the items of one side of one line for this lane, walked away from
// the centre (red: iw = i0, i0 + 1, ..., i1; blue: iw = i0, i0 - 1, ..., i1),
// so that v grows along the walk.  Once v is in a tail regime (a <= 0.2 and
// v > 10: the Lorentz wing; a > 0.2 outside a <= 1.4 and a + v <= 3.2:
// VOIGT's asymptote) every later item is too, and the rest of the side runs
// without regime tests.  Each item is the value lo_accum adds, with the same
// operations: the sums are those of lo_accum's loop.
__device__ __forceinline__ void lo_side_mixed(const LoPar &q, const double *w_t, double wl,
                                              int i0, int i1, int dir, int iw0, int stride, int t,
                                              float *acc_s, float *acc_c, const float4 *f)
{
#pragma clang fp contract(off)
    const double adamp = q.p1;
    const float c32 = __int_as_float(__double2hiint(q.p2)), k32 = __int_as_float(__double2loint(q.p2));
    const int n_items = dir > 0 ? i1 - i0 + 1 : i0 - i1 + 1;
    int k = 0;
    for (; k < n_items; ++k) {
        const int iw = i0 + dir * k;
        const double d = dir > 0 ? w_t[iw] - wl : wl - w_t[iw], v = d * q.p0;
        if (adamp <= 0.20 ? v > 10.0 : !(adamp <= 1.4 && adamp + v <= 3.2)) break;
        const float x = lo_cv_mixed(c32, v, adamp, f);
        const int slot = (iw - iw0) * stride + t;
        const float s0 = acc_s[slot], y = x - acc_c[slot], tt = s0 + y;
        acc_c[slot] = (tt - s0) - y;
        acc_s[slot] = tt;
    }
    if (adamp <= 0.20) {
        for (; k < n_items; ++k) {  // Lorentz tail: K / v^2
            const int iw = i0 + dir * k;
            const double d = dir > 0 ? w_t[iw] - wl : wl - w_t[iw], v = d * q.p0;
            const float vf = (float)v;
            const float x = k32 * __builtin_amdgcn_rcpf(vf * vf);
            const int slot = (iw - iw0) * stride + t;
            const float s0 = acc_s[slot], y = x - acc_c[slot], tt = s0 + y;
            acc_c[slot] = (tt - s0) - y;
            acc_s[slot] = tt;
        }
    } else {
        const float a32 = (float)adamp, aa32 = a32 * a32;
        for (; k < n_items; ++k) {  // VOIGT's asymptotic form with one reciprocal
            const int iw = i0 + dir * k;
            const double d = dir > 0 ? w_t[iw] - wl : wl - w_t[iw], v = d * q.p0;
            const float vf = (float)v, vv = vf * vf;
            const float ru = __builtin_amdgcn_rcpf((aa32 + vv) * 1.4142f);
            const float aau = aa32 * ru, vvu = vv * ru;
            const float r = a32 * 0.79788f * ru;
            const float x = c32 * (adamp <= 100.0 ? ((((aau - 10.0f * vvu) * aau * 3.0f + 15.0f * vvu * vvu) +
                                                      3.0f * vv - aa32) * (ru * ru) + 1.0f) * r
                                                  : r);
            const int slot = (iw - iw0) * stride + t;
            const float s0 = acc_s[slot], y = x - acc_c[slot], tt = s0 + y;
            acc_c[slot] = (tt - s0) - y;
            acc_s[slot] = tt;
        }
    }
}

__device__ __forceinline__ LoPar lo_shfl(const LoPar &q, int src)
{
    LoPar r;
    r.p0 = __shfl(q.p0, src, 64);
    r.p1 = __shfl(q.p1, src, 64);
    r.p2 = __shfl(q.p2, src, 64);
    return r;
}

// Gather: a block holds tiles of kLoPts consecutive grid points; lane t
// serves tile t / n and depth t % n.  The lines that can reach the block's
// points are one contiguous run of the NU-sorted store (segments NU =
// first - 100 .. last + 100); the block streams it in batches of kLoBatch
// lines, staging their per-depth state in LDS with coalesced loads.  For
// every staged line, each lane's items are the points of its tile that the
// line's red/blue reach covers at its depth.
//   kFlat = false: each lane walks its own items;
//   kFlat = true:  a wave scans the item counts of its 64 lanes and deals
//                  all of the line's items out over the 64 lanes (owner by
//                  binary search over the prefix sums, parameters by
//                  shuffle), so lanes with long ranges no longer hold up
//                  the others.
// A line's items target distinct (point, depth) slots and lines are taken
// one after another in store order, which for every grid point is the
// order LINOP1 adds them: the sum in each LDS slot is the same in both
// forms and every tile shape.  No atomics, deterministic.
template <bool F64, int kLoPts, int kLoBlock, int kLoBatch, bool kFlat>
__global__ __launch_bounds__(kLoBlock) void k_lo_gather(
    int n, int numnu, int tiles, const int *__restrict__ seg, const int *__restrict__ l_nu,
    const int *__restrict__ l_reach, const int *__restrict__ state, const double *__restrict__ par,
    const double *__restrict__ l_const, const double *__restrict__ waveset,
    const double *__restrict__ h0, const double *__restrict__ h1, const double *__restrict__ h2,
    const float4 *__restrict__ f,
    double *__restrict__ xlines)
{
#pragma clang fp contract(off)
    static_assert(kLoPts <= 32 && kLoBlock % 64 == 0, "tile shape");
    // LDS: F64 sums, or FP32 sums + compensations (same bytes)
    __shared__ double acc_d[kLoPts * kLoBlock];
    __shared__ int s_state[kLoBatch * A12_NRMAX_LAYERS];
    __shared__ int s_nu[kLoBatch], s_row[kLoBatch];
    __shared__ int s_idx[kLoBlock], s_wcnt[kLoBlock / 64], s_nact;
    __shared__ double s_wl[kLoBatch];
    __shared__ double s_w[kLoBlock / 8 * kLoPts];  // WAVESET of the block's points
    float *acc_s = reinterpret_cast<float *>(acc_d);
    float *acc_c = acc_s + kLoPts * kLoBlock;
    const int per = kLoBlock / n;  // tiles per block
    const int t = threadIdx.x, lane = t & 63;
    const int lt = t / n, j = t - lt * n;
    const int tile0 = blockIdx.x * per;
    const int tile = tile0 + lt;
    const bool live = lt < per && tile < tiles;
    const int iw0 = tile * kLoPts, iw1 = min(iw0 + kLoPts, numnu) - 1;
    for (int p = 0; p < kLoPts; ++p) {
        if (F64) acc_d[p * kLoBlock + t] = 0.0;
        else acc_s[p * kLoBlock + t] = acc_c[p * kLoBlock + t] = 0.0f;
    }
    if (t < per * kLoPts) s_w[t] = waveset[min(tile0 * kLoPts + t, numnu - 1)];
    // the block's candidate lines: segments NU in [first - 100, last + 100]
    // (red from NU reaches NU..NU+100, blue NU-100..NU-1)
    const int bfirst = tile0 * kLoPts, blast = min((tile0 + per) * kLoPts, numnu) - 1;
    const int lbeg = seg[max(0, bfirst - kMaxWing)], lend = seg[min(numnu, blast + kMaxWing) + 1];
    // candidates are scanned kLoBlock at a time; those whose widest reach
    // meets [bfirst, blast] are compacted in store order into s_idx, then
    // staged and processed kLoBatch at a time
    for (int c0 = lbeg; c0 < lend; c0 += kLoBlock) {
        {
            const int q = c0 + t;
            bool act = false;
            if (q < lend) {
                const int k = l_nu[q], rr = l_reach[q], rmax = rr & 0xff, bmax = (rr >> 8) & 0xff;
                act = (rmax > 0 && k <= blast && k + rmax - 1 >= bfirst) ||
                      (bmax > 0 && k - 1 >= bfirst && k - bmax <= blast);
            }
            const unsigned long long m = __ballot(act);
            __syncthreads();  // the previous chunk is done with s_idx
            if (lane == 0) s_wcnt[t >> 6] = __popcll(m);
            __syncthreads();
            if (act) {
                int pos = __popcll(m & ((1ull << lane) - 1ull));
                for (int w = 0; w < (t >> 6); ++w) pos += s_wcnt[w];
                s_idx[pos] = q;
            }
            if (t == 0) {
                int sm = 0;
                for (int w = 0; w < kLoBlock / 64; ++w) sm += s_wcnt[w];
                s_nact = sm;
            }
        }
        __syncthreads();
        const int nact = s_nact;
        for (int b0 = 0; b0 < nact; b0 += kLoBatch) {
            const int nb = min(kLoBatch, nact - b0);
            __syncthreads();
            for (int i = t; i < nb * n; i += kLoBlock) {
                const int li = i / n;
                s_state[i] = state[(size_t)s_idx[b0 + li] * n + (i - li * n)];
            }
            if (t < nb) {
                const int lq = s_idx[b0 + t];
                s_nu[t] = l_nu[lq];
                s_wl[t] = l_const[5 * (size_t)lq];
                s_row[t] = lq;
            }
            __syncthreads();
            if (!kFlat && !live) continue;
            for (int i = 0; i < nb; ++i) {
                // this lane's items: tile points in [ra, rb] (red) and [ba, bb] (blue)
                int ra = 1, rb = 0, ba = 1, bb = 0;
                if (live) {
                    const int st = s_state[i * n + j];
                    if (st < 0) {
                        const int k = s_nu[i], rn = st & 0xff, bn = (st >> 8) & 0xff;
                        ra = max(k, iw0);
                        rb = min(k + rn - 1, iw1);
                        ba = max(k - bn, iw0);
                        bb = min(k - 1, iw1);
                    }
                }
                const int nred = max(0, rb - ra + 1), nblue = max(0, bb - ba + 1), c = nred + nblue;
                const double wl = s_wl[i];
                const size_t orow = (size_t)s_row[i] * n;
                if (!kFlat) {
                    if (c == 0) continue;
                    const LoPar q = lo_par<F64>(par, orow + j);
                    const double *w_t = s_w + lt * kLoPts - iw0;
                    for (int iw = ra; iw <= rb; ++iw)
                        lo_accum<F64>(q, w_t[iw] - wl, (iw - iw0) * kLoBlock + t, acc_d, acc_s, acc_c,
                                      h0, h1, h2, f);
                    for (int iw = ba; iw <= bb; ++iw)
                        lo_accum<F64>(q, wl - w_t[iw], (iw - iw0) * kLoBlock + t, acc_d, acc_s, acc_c,
                                      h0, h1, h2, f);
                } else {
                    // wave-wide inclusive scan of the item counts
                    int incl = c;
                    for (int dd = 1; dd < 64; dd <<= 1) {
                        const int y = __shfl_up(incl, dd, 64);
                        if (lane >= dd) incl += y;
                    }
                    const int total = __shfl(incl, 63, 64);
                    if (total == 0) continue;
                    const LoPar q = c > 0 ? lo_par<F64>(par, orow + j) : LoPar{1.0, 0.0, 1.0};
                    // owner's item layout: red offset, red count, blue offset, tile base in s_w
                    // (an empty range packs as offset 0: its bounds may lie outside the tile)
                    const int pk = (nred ? ra - iw0 : 0) | (nred << 6) | ((nblue ? ba - iw0 : 0) << 12) |
                                   ((lt * kLoPts) << 18);
                    const int excl = incl - c;
                    for (int base = 0; base < total; base += 64) {
                        const int item = base + lane;
                        int lo = 0, hi = 63;
    #pragma unroll
                        for (int s = 0; s < 6; ++s) {
                            const int mid = (lo + hi) >> 1;
                            if (__shfl(incl, mid, 64) > item) hi = mid;
                            else lo = mid + 1;
                        }
                        const int o = min(lo, 63);
                        const int ex = __shfl(excl, o, 64), opk = __shfl(pk, o, 64);
                        const LoPar qo = lo_shfl(q, o);
                        if (item < total) {
                            const int kk = item - ex, onred = (opk >> 6) & 63;
                            const bool blue = kk >= onred;
                            const int pt = blue ? ((opk >> 12) & 63) + (kk - onred) : (opk & 63) + kk;
                            const double w = s_w[(opk >> 18) + pt];
                            const int to = t - lane + o;  // owner's lane in the block
                            lo_accum<F64>(qo, blue ? wl - w : w - wl, pt * kLoBlock + to, acc_d, acc_s,
                                          acc_c, h0, h1, h2, f);
                        }
                    }
                }
            }
        }
    }
    if (live)
        for (int iw = iw0; iw <= iw1; ++iw) {
            const int q = (iw - iw0) * kLoBlock + t;
            xlines[(size_t)iw * n + j] = F64 ? acc_d[q] : (double)acc_s[q] - (double)acc_c[q];
        }
}

// Barrier-free form of the gather: after the block's first barrier every
// wavefront works alone.  A wave takes the candidate lines that can reach
// its own points (its lanes' tiles), 64 at a time, keeps those whose widest
// reach meets them (ballot) and walks them in store order (lowest bit
// first): per line its NU and wavelength are wave-uniform loads, each lane
// reads its depth's state and parameters and adds its items into its own
// LDS slots.  The same items in the same per-slot order as k_lo_gather.
template <bool F64, int kLoPts, int kLoBlock>
__global__ __launch_bounds__(kLoBlock) void k_lo_gather_w(
    int n, int numnu, int tiles, const int *__restrict__ seg, const int *__restrict__ l_nu,
    const int *__restrict__ l_reach, const int *__restrict__ state, const double *__restrict__ par,
    const double *__restrict__ l_const, const double *__restrict__ waveset,
    const double *__restrict__ h0, const double *__restrict__ h1, const double *__restrict__ h2,
    const float4 *__restrict__ f,
    double *__restrict__ xlines)
{
#pragma clang fp contract(off)
    static_assert(kLoPts <= 32 && kLoBlock % 64 == 0, "tile shape");
    __shared__ double acc_d[kLoPts * kLoBlock];
    __shared__ double s_w[kLoBlock / 8 * kLoPts];
    float *acc_s = reinterpret_cast<float *>(acc_d);
    float *acc_c = acc_s + kLoPts * kLoBlock;
    const int per = kLoBlock / n;
    const int t = threadIdx.x, lane = t & 63;
    const int lt = t / n, j = t - lt * n;
    const int tile0 = blockIdx.x * per;
    const int tile = tile0 + lt;
    const bool live = lt < per && tile < tiles;
    const int iw0 = tile * kLoPts, iw1 = min(iw0 + kLoPts, numnu) - 1;
    for (int p = 0; p < kLoPts; ++p) {
        if (F64) acc_d[p * kLoBlock + t] = 0.0;
        else acc_s[p * kLoBlock + t] = acc_c[p * kLoBlock + t] = 0.0f;
    }
    for (int i = t; i < per * kLoPts; i += kLoBlock) s_w[i] = waveset[min(tile0 * kLoPts + i, numnu - 1)];
    __syncthreads();
    const int t0 = t - lane;  // this wave's first lane
    const int wlt0 = t0 / n, wlt1 = min((t0 + 63) / n, per - 1);
    if (wlt0 < per && tile0 + wlt0 < tiles) {  // uniform over the wave
        const int wfirst = (tile0 + wlt0) * kLoPts;
        const int wlast = min((tile0 + wlt1 + 1) * kLoPts, numnu) - 1;
        const int lbeg = seg[max(0, wfirst - kMaxWing)];
        const int lend = seg[min(numnu, wlast + kMaxWing) + 1];
        const double *w_t = s_w + lt * kLoPts - iw0;
        for (int c0 = lbeg; c0 < lend; c0 += 64) {
            const int q = c0 + lane;
            bool act = false;
            if (q < lend) {
                const int k = l_nu[q], rr = l_reach[q], rmax = rr & 0xff, bmax = (rr >> 8) & 0xff;
                act = (rmax > 0 && k <= wlast && k + rmax - 1 >= wfirst) ||
                      (bmax > 0 && k - 1 >= wfirst && k - bmax <= wlast);
            }
            unsigned long long m = __ballot(act);
            if (!m) continue;
            // one line ahead: the next line's state, parameters, NU and
            // wavelength are loaded while the current one is added
            int row = c0 + __builtin_ctzll(m);  // lowest first: store order
            m &= m - 1;
            int st = live ? state[(size_t)row * n + j] : 0;
            double p0 = par[3 * ((size_t)row * n + j)], p1 = par[3 * ((size_t)row * n + j) + 1],
                   p2 = par[3 * ((size_t)row * n + j) + 2];
            int k = l_nu[row];
            double wl = l_const[5 * (size_t)row];
            for (;;) {
                const bool more = m != 0;
                int st_n = 0, k_n = 0, row_n = 0;
                double q0 = 0.0, q1 = 0.0, q2 = 0.0, wl_n = 0.0;
                if (more) {
                    row_n = c0 + __builtin_ctzll(m);
                    m &= m - 1;
                    const size_t o = (size_t)row_n * n + j;
                    st_n = live ? state[o] : 0;
                    q0 = par[3 * o];
                    q1 = par[3 * o + 1];
                    q2 = par[3 * o + 2];
                    k_n = l_nu[row_n];
                    wl_n = l_const[5 * (size_t)row_n];
                }
                if (st < 0) {
                    const int rn = st & 0xff, bn = (st >> 8) & 0xff;
                    const int ra = max(k, iw0), rb = min(k + rn - 1, iw1);
                    const int ba = max(k - bn, iw0), bb = min(k - 1, iw1);
                    if (ra <= rb || ba <= bb) {
                        const LoPar qp = {p0, p1, p2};
                        if (F64) {
                            for (int iw = ra; iw <= rb; ++iw)
                                lo_accum<F64>(qp, w_t[iw] - wl, (iw - iw0) * kLoBlock + t, acc_d,
                                              acc_s, acc_c, h0, h1, h2, f);
                            for (int iw = ba; iw <= bb; ++iw)
                                lo_accum<F64>(qp, wl - w_t[iw], (iw - iw0) * kLoBlock + t, acc_d,
                                              acc_s, acc_c, h0, h1, h2, f);
                        } else {
                            lo_side_mixed(qp, w_t, wl, ra, rb, 1, iw0, kLoBlock, t, acc_s, acc_c,
                                          f);
                            lo_side_mixed(qp, w_t, wl, bb, ba, -1, iw0, kLoBlock, t, acc_s,
                                          acc_c, f);
                        }
                    }
                }
                if (!more) break;
                row = row_n;
                st = st_n;
                p0 = q0;
                p1 = q1;
                p2 = q2;
                k = k_n;
                wl = wl_n;
            }
        }
    }
    if (live)
        for (int iw = iw0; iw <= iw1; ++iw) {
            const int q = (iw - iw0) * kLoBlock + t;
            xlines[(size_t)iw * n + j] = F64 ? acc_d[q] : (double)acc_s[q] - (double)acc_c[q];
        }
}

// Point form of the gather: one lane per (depth, kPP output points).  A
// wavefront covers 8 kPP grid points x 8 depths: lane = 8 * depth offset +
// point offset p, its points p, p + 8, ...; wavefront w of the grid takes
// point block w % npb and depth group w / npb.  It ballots the candidate
// lines (NU within 100 points of its block) on their widest reach over its
// 8 depths (greach, from k_lo_linereach) and walks the kept ones in store
// order, one line ahead: the line's NU and wavelength are wave-uniform
// scalar loads, each lane reads its depth's state and parameters (8
// distinct records per line) and adds the items of its points that lie in
// the reach into register sums.  Each element sees the same lines in the
// same order with the same item values as in the other forms, so XLINES
// has the same bits.  kRun > 0: workgroups are renumbered so that each XCD
// (workgroup b runs on XCD b % 8) takes runs of kRun consecutive ones,
// which puts neighbouring point blocks, and the line records they share,
// in one L2.
template <bool F64, int kPP, int kRun>
__global__ __launch_bounds__(256) void k_lo_gather_p(
    int n, int numnu, int npb, int ngroups, int nl, const int *__restrict__ seg,
    const int *__restrict__ l_nu, const unsigned short *__restrict__ greach,
    const int *__restrict__ state, const double *__restrict__ par,
    const double *__restrict__ l_const, const double *__restrict__ waveset,
    const double *__restrict__ h0, const double *__restrict__ h1, const double *__restrict__ h2,
    const float4 *__restrict__ f, double *__restrict__ xlines)
{
#pragma clang fp contract(off)
    int b = blockIdx.x;
    if (kRun > 0) {
        const int full = ((int)gridDim.x / (8 * kRun)) * (8 * kRun);
        if (b < full) {
            const int xcd = b & 7, r = b >> 3;
            b = ((r / kRun) * 8 + xcd) * kRun + r % kRun;
        }
    }
    const int lane = threadIdx.x & 63;
    // wave-uniform values in scalar registers: the line loop's NU and
    // wavelength loads and its record addresses are then scalar
    const int wv = __builtin_amdgcn_readfirstlane(b * 4 + (int)(threadIdx.x >> 6));
    if (wv >= npb * ngroups) return;  // whole wavefronts
    const int pb = wv % npb, g = wv / npb;
    const int wfirst = pb * 8 * kPP, wlast = min(wfirst + 8 * kPP, numnu) - 1;
    const int iw = wfirst + (lane & 7), j = g * 8 + (lane >> 3);
    const bool jlive = j < n;
    const int jl = min(j, n - 1);
    double w[kPP], sd[kPP];
    float s[kPP], c[kPP];
    bool live[kPP];
#pragma unroll
    for (int u = 0; u < kPP; ++u) {
        live[u] = jlive && iw + 8 * u < numnu;
        w[u] = waveset[min(iw + 8 * u, numnu - 1)];
        sd[u] = 0.0;
        s[u] = c[u] = 0.0f;
    }
    const int lbeg = __builtin_amdgcn_readfirstlane(seg[max(0, wfirst - kMaxWing)]);
    const int lend = __builtin_amdgcn_readfirstlane(seg[min(numnu, wlast + kMaxWing) + 1]);
    const unsigned short *gr = greach + (size_t)g * nl;
    for (int c0 = lbeg; c0 < lend; c0 += 64) {
        const int q = c0 + lane;
        bool act = false;
        if (q < lend) {
            const int k = l_nu[q], rr = gr[q], rmax = rr & 0xff, bmax = rr >> 8;
            act = (rmax > 0 && k <= wlast && k + rmax - 1 >= wfirst) ||
                  (bmax > 0 && k - 1 >= wfirst && k - bmax <= wlast);
        }
        unsigned long long m = __ballot(act);
        if (!m) continue;
        // one line ahead: the next line's state, parameters, NU and
        // wavelength are loaded while the current one is added
        int row = __builtin_amdgcn_readfirstlane(c0 + __builtin_ctzll(m));  // lowest first
        m &= m - 1;
        const int *st_r = state + (size_t)row * n;
        const double *pr = par + (size_t)3 * row * n;
        int st = st_r[jl];
        double p0 = pr[3 * jl], p1 = pr[3 * jl + 1], p2 = pr[3 * jl + 2];
        int k = l_nu[row];
        double wl = l_const[5 * (size_t)row];
        for (;;) {
            const bool more = m != 0;
            int st_n = 0, k_n = 0;
            double q0 = 0.0, q1 = 0.0, q2 = 0.0, wl_n = 0.0;
            if (more) {
                const int row_n = __builtin_amdgcn_readfirstlane(c0 + __builtin_ctzll(m));
                m &= m - 1;
                const int *st_rn = state + (size_t)row_n * n;
                const double *prn = par + (size_t)3 * row_n * n;
                st_n = st_rn[jl];
                q0 = prn[3 * jl];
                q1 = prn[3 * jl + 1];
                q2 = prn[3 * jl + 2];
                k_n = l_nu[row_n];
                wl_n = l_const[5 * (size_t)row_n];
            }
            if (st < 0) {
                // an item where -bn <= iw - NU < rn (red 0 .. rn - 1, blue
                // -1 .. -bn)
                const int rn = st & 0xff, bn = (st >> 8) & 0xff;
                const LoPar qp = {p0, p1, p2};
#pragma unroll
                for (int u = 0; u < kPP; ++u) {
                    if (live[u] && (unsigned)(iw + 8 * u - k + bn) < (unsigned)(rn + bn)) {
                        // red: w - wl > 0; blue: wl - w = -(w - wl) exactly
                        const double d = fabs(w[u] - wl);
                        if (F64) {
                            sd[u] = sd[u] + lo_cv<double>(qp.p0, d / qp.p2, qp.p1, h0, h1, h2);
                        } else {
                            const float x = lo_item_mixed(qp, d, f);
                            const float y = x - c[u], tt = s[u] + y;
                            c[u] = (tt - s[u]) - y;
                            s[u] = tt;
                        }
                    }
                }
            }
            if (!more) break;
            st = st_n;
            p0 = q0;
            p1 = q1;
            p2 = q2;
            k = k_n;
            wl = wl_n;
        }
    }
#pragma unroll
    for (int u = 0; u < kPP; ++u)
        if (live[u])
            xlines[(size_t)(iw + 8 * u) * n + j] = F64 ? sd[u] : (double)s[u] - (double)c[u];
}

// Tile shapes (points per tile, lanes per block, lines per LDS batch, form:
// 0 staged, 1 flat, 2 barrier-free per wave) the autotuner chooses from;
// LDS = 8 B x points x lanes (+ 4 B x 128 x batch when staged).
using LoTileFn = void (*)(int, int, int, const int *, const int *, const int *, const int *,
                         const double *, const double *, const double *, const double *,
                         const double *, const double *, const float4 *, double *);
using LoPointFn = void (*)(int, int, int, int, int, const int *, const int *,
                           const unsigned short *, const int *, const double *, const double *,
                           const double *, const double *, const double *, const double *,
                           const float4 *, double *);
struct LoShape {
    int pts, block, batch;
    int flat;  // form: 0 staged, 1 flat, 2 per wave, 3 point (batch: XCD run length)
    LoTileFn mixed, fp64;
    LoPointFn pmixed, pfp64;
};
#define A12_LO_SHAPE(P, B, T, F) \
    {P, B, T, F ? 1 : 0, k_lo_gather<false, P, B, T, F>, k_lo_gather<true, P, B, T, F>, nullptr, nullptr}
#define A12_LO_WAVE(P, B) \
    {P, B, 0, 2, k_lo_gather_w<false, P, B>, k_lo_gather_w<true, P, B>, nullptr, nullptr}
#define A12_LO_POINT(P, R) \
    {8 * P, 256, R, 3, nullptr, nullptr, k_lo_gather_p<false, P, R>, k_lo_gather_p<true, P, R>}
const LoShape kLoShapes[] = {
    A12_LO_SHAPE(8, 512, 32, false),  A12_LO_SHAPE(4, 512, 64, false),
    A12_LO_SHAPE(4, 256, 32, false),  A12_LO_SHAPE(8, 256, 32, false),
    A12_LO_SHAPE(4, 256, 8, false),   A12_LO_SHAPE(8, 512, 32, true),
    A12_LO_SHAPE(2, 256, 32, false),  A12_LO_SHAPE(4, 256, 16, false),
    A12_LO_SHAPE(2, 256, 16, false),  A12_LO_SHAPE(8, 256, 16, false),
    A12_LO_WAVE(4, 256),              A12_LO_WAVE(2, 256),
    A12_LO_WAVE(8, 256),              A12_LO_WAVE(4, 512),
    A12_LO_POINT(1, 0),               A12_LO_POINT(1, 16),
    A12_LO_POINT(2, 0),               A12_LO_POINT(2, 16),
    A12_LO_POINT(4, 0),
};
#undef A12_LO_SHAPE
#undef A12_LO_WAVE
#undef A12_LO_POINT
constexpr int kNumLoShapes = sizeof(kLoShapes) / sizeof(kLoShapes[0]);

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
// The line store, once per store (src/atlas12_modules.f90:18176-18236 per
// record: TABLOG decoding, continuum bin NUCONT, vacuum wavelength, grid
// index NU, CGF and the three damping constants).  None of it depends on
// the atmosphere, so an iteration re-reads nothing here.  The grid fields
// of a (numnu, waveset, nwave1, iwavetab, mion) are the only ones read.
int a12i_atlas_load_lines(a12_ctx *ctx, const a12_atlas_atm *a, const int32_t *linedata,
                          int32_t nlines)
{
    AtlasDev &A = ctx->atl;
    if (!a || (!linedata && nlines > 0) || nlines < 0) {
        ctx->err = "a12_atlas_load_lines: null argument";
        return A12_ERR_ARG;
    }
    if (a->numnu < 3 || a->mion < 1 || a->nwave1 < 2 || !a->waveset || !a->iwavetab) {
        ctx->err = "a12_atlas_load_lines: bad grid dimensions or null array";
        return A12_ERR_ARG;
    }
    const int numnu = a->numnu;
    for (int i = 1; i < numnu; ++i)
        if (!(a->waveset[i] > a->waveset[i - 1])) {
            ctx->err = "a12_atlas_load_lines: WAVESET must ascend";
            return A12_ERR_ARG;
        }
    // a new store invalidates everything computed from the previous one
    A.lines_ok = A.loaded = A.xl_ready = A.kapp_ready = A.rt_done = false;
    A.prof_lo = A.prof_rt = false;
    hipStream_t s = ctx->stream;
    A.numnu = numnu;
    A.mion = a->mion;
    A.nwave1 = a->nwave1;
    A.waveset_h.assign(a->waveset, a->waveset + numnu);
    A.iwavetab_h.assign(a->iwavetab, a->iwavetab + a->nwave1);
    // mean points per octave of the grid: places the reach search's first
    // probe (a starting point only; any ascending grid gives exact reaches)
    A.lg2inv = (float)((numnu - 1) / std::log2(a->waveset[numnu - 1] / a->waveset[0]));
    if (!std::isfinite(A.lg2inv)) A.lg2inv = 0.0f;
    A12_HIP(up(A.waveset, a->waveset, numnu, s));
    // TABLOG (:2083-2086): 10**((I - 16384) * 0.001) for every int16 code,
    // the same expression the per-record decode would evaluate
    std::vector<double> tablog(65536);
    for (int i = 0; i < 65536; ++i)
        tablog[i] = std::pow(10.0, (double)((int)(int16_t)(uint16_t)i - 16384) * .001);
    auto tl = [&](int16_t c) { return tablog[(uint16_t)c]; };
    const double ratiolg = std::log(1.0 + 1.0 / 2000000.0);
    const double cen_prefac4 = 0.026538 / 1.7724538509055159 / 2.99792458e17;
    const double fourpi_c_inv = 1.0 / ((4.0 * 3.141592653589793) * 2.99792458e17);
    const double start = a->waveset[0] - 1.0, stop = a->waveset[numnu - 1] + 1.0;
    struct Rec {
        int nu, nucont, nel;
        double c[6];
    };
    std::vector<Rec> recs((size_t)nlines);
    std::vector<char> keep((size_t)nlines, 0);
    // records are independent: decode them on a few host threads
    const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    auto decode = [&](int64_t i0, int64_t i1) {
        for (int64_t i = i0; i < i1; ++i) {
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
            Rec &q = recs[i];
            q.nu = nu;
            q.nucont = nucont;
            q.nel = nel;
            q.c[0] = wl;
            q.c[1] = cen_prefac4 * wl * tl(p[2]);
            q.c[2] = tl(p[1]);
            q.c[3] = tl(p[3]) * wl * fourpi_c_inv;
            q.c[4] = tl(p[4]) * wl * fourpi_c_inv;
            q.c[5] = tl(p[5]) * wl * fourpi_c_inv;
            keep[i] = 1;
        }
    };
    {
        std::vector<std::thread> th;
        const int64_t per = ((int64_t)nlines + nth - 1) / nth;
        for (int k = 0; k < nth; ++k) {
            const int64_t i0 = k * per, i1 = std::min<int64_t>(nlines, i0 + per);
            if (i0 < i1) th.emplace_back(decode, i0, i1);
        }
        for (auto &t : th) t.join();
    }
    // stable counting sort by NU (0..numnu): kept records in store order
    // within each NU, as the reference meets them along the store
    std::vector<int> seg(numnu + 2, 0);
    for (int64_t i = 0; i < nlines; ++i)
        if (keep[i]) ++seg[recs[i].nu + 1];
    for (int k = 1; k <= numnu + 1; ++k) seg[k] += seg[k - 1];
    const int nl = seg[numnu + 1];
    std::vector<int> l_nu(nl), l_nucont(nl), l_nel(nl);
    std::vector<double> l_const((size_t)6 * std::max(nl, 1));
    {
        std::vector<int> pos(seg.begin(), seg.end() - 1);
        for (int64_t i = 0; i < nlines; ++i) {
            if (!keep[i]) continue;
            const Rec &q = recs[i];
            const int o = pos[q.nu]++;
            l_nu[o] = q.nu;
            l_nucont[o] = q.nucont;
            l_nel[o] = q.nel;
            for (int k = 0; k < 5; ++k) l_const[5 * (size_t)o + k] = q.c[k];
            l_const[5 * (size_t)nl + o] = q.c[5];
        }
    }
    // seg[k] = first sorted line with nu >= k, k = 0..numnu+1
    A.nl = nl;
    A12_HIP(up(A.l_nu, l_nu.data(), nl, s));
    A12_HIP(up(A.l_nucont, l_nucont.data(), nl, s));
    A12_HIP(up(A.l_nel, l_nel.data(), nl, s));
    A12_HIP(up(A.l_const, l_const.data(), (size_t)6 * nl, s));
    A12_HIP(up(A.seg, seg.data(), seg.size(), s));
    A12_HIP(hipStreamSynchronize(s));
    A.lines_ok = true;
    ++A.grid_gen;
    return A12_OK;
}

// The per-iteration state LINOP1 and the RT pass read (T, XNE, HCKT, HKT,
// RHOX, TXNXN from XNF, XNFDOP, DOPPLE, TABCONT, RCOSET, TABVOIGT, TEFF),
// on the grid of the loaded store.  Invalidates XLINES, the KAPP tables and
// the RT-pass sums of the previous state.
int a12i_atlas_set_state(a12_ctx *ctx, const a12_atlas_atm *a)
{
    AtlasDev &A = ctx->atl;
    if (!A.lines_ok) {
        ctx->err = "a12_atlas_set_state: call a12_atlas_load_lines first";
        return A12_ERR_STATE;
    }
    if (!a) {
        ctx->err = "a12_atlas_set_state: null argument";
        return A12_ERR_ARG;
    }
    if (a->nrhox < 8 || a->nrhox > A12_NRMAX_LAYERS || a->ldion < a->nrhox ||
        a->ldtab < a->nrhox || !a->rcoset || !a->tabcont || !a->xnfdop || !a->dopple || !a->t ||
        !a->xne || !a->hckt || !a->hkt || !a->rhox || !a->xnf1 || !a->xnf3 || !a->xnf841 ||
        !a->h0tab || !a->h1tab || !a->h2tab) {
        ctx->err = "a12_atlas_set_state: bad dimensions or null array";
        return A12_ERR_ARG;
    }
    if (a->numnu != A.numnu || a->nwave1 != A.nwave1 || a->mion != A.mion ||
        (a->waveset && std::memcmp(a->waveset, A.waveset_h.data(), sizeof(double) * A.numnu)) ||
        (a->iwavetab && std::memcmp(a->iwavetab, A.iwavetab_h.data(), sizeof(int32_t) * A.nwave1))) {
        ctx->err = "a12_atlas_set_state: grid differs from the loaded line store's";
        return A12_ERR_STATE;
    }
    A.loaded = A.xl_ready = A.kapp_ready = A.rt_done = false;
    A.prof_lo = A.prof_rt = false;
    hipStream_t s = ctx->stream;
    const int n = a->nrhox;
    A.n = n;
    A.teff = a->teff;
    A.t_h.assign(a->t, a->t + n);
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
    // one pinned-free host staging buffer: depth3 | depth_lo | tabcont |
    // xnfdop | dopple, copied with one call per device array
    const size_t ntab = (size_t)A.nwave1 * n, nion = (size_t)A.mion * n;
    A.stage_h.resize(3 * (size_t)n + dset.size() + ntab + 2 * nion);
    double *h = A.stage_h.data();
    std::memcpy(h, dep.data(), sizeof(double) * 3 * n);
    std::memcpy(h + 3 * n, dset.data(), sizeof(double) * dset.size());
    double *ht = h + 3 * n + dset.size(), *hx = ht + ntab, *hd = hx + nion;
    for (int i = 0; i < A.nwave1; ++i)
        std::memcpy(ht + (size_t)i * n, a->tabcont + (size_t)i * a->ldtab, sizeof(double) * n);
    for (int i = 0; i < A.mion; ++i) {
        std::memcpy(hx + (size_t)i * n, a->xnfdop + (size_t)i * a->ldion, sizeof(double) * n);
        std::memcpy(hd + (size_t)i * n, a->dopple + (size_t)i * a->ldion, sizeof(double) * n);
    }
    A12_HIP(up(A.depth3, h, (size_t)3 * n, s));  // t hkt rhox for the RT pass
    A12_HIP(up(A.depth_lo, h + 3 * n, dset.size(), s));
    A12_HIP(up(A.tabcont, ht, ntab, s));
    A12_HIP(up(A.xnfdop, hx, nion, s));
    A12_HIP(up(A.dopple, hd, nion, s));
    A12_HIP(up(A.rcoset, a->rcoset, A.numnu, s));
    std::vector<double> tabs((size_t)3 * kTabN);
    std::memcpy(&tabs[0], a->h0tab, sizeof(double) * kTabN);
    std::memcpy(&tabs[kTabN], a->h1tab, sizeof(double) * kTabN);
    std::memcpy(&tabs[2 * kTabN], a->h2tab, sizeof(double) * kTabN);
    // FP32 rows (H0, H1, H2, 0) of the three tables: one 16-byte load per value
    std::vector<float4> tabs32(kTabN);
    for (int i = 0; i < kTabN; ++i)
        tabs32[i] = make_float4((float)tabs[i], (float)tabs[kTabN + i], (float)tabs[2 * kTabN + i], 0.0f);
    A12_HIP(up(A.tabs, tabs.data(), tabs.size(), s));
    A12_HIP(up(A.tabs32, tabs32.data(), tabs32.size(), s));
    A12_HIP(hipStreamSynchronize(s));  // the host staging is reused by the next state
    A.loaded = true;
    return A12_OK;
}

int a12i_atlas_load(a12_ctx *ctx, const a12_atlas_atm *a, const int32_t *linedata, int32_t nlines)
{
    const int rc = a12i_atlas_load_lines(ctx, a, linedata, nlines);
    return rc ? rc : a12i_atlas_set_state(ctx, a);
}

static int lo_gather_launch(a12_ctx *ctx, int prec, int shape)
{
    AtlasDev &A = ctx->atl;
    const int n = A.n, numnu = A.numnu;
    const double *tabs = A.tabs.as<double>();
    const float4 *tabs32 = A.tabs32.as<float4>();
    const LoShape &sh = kLoShapes[shape];
    if (sh.flat == 3) {
        auto g = prec == A12_PREC_FP64 ? sh.pfp64 : sh.pmixed;
        const int npb = (numnu + sh.pts - 1) / sh.pts, ng = (n + 7) / 8;
        const long long waves = (long long)npb * ng;
        hipLaunchKernelGGL(g, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, ctx->stream, n,
                           numnu, npb, ng, A.nl, A.seg.as<int>(), A.l_nu.as<int>(),
                           A.greach.as<unsigned short>(), A.state.as<int>(), A.par.as<double>(),
                           A.l_const.as<double>(), A.waveset.as<double>(), tabs, tabs + kTabN,
                           tabs + 2 * kTabN, tabs32, A.xlines.as<double>());
        A12_CHECK_LAUNCH();
        return A12_OK;
    }
    auto g = prec == A12_PREC_FP64 ? sh.fp64 : sh.mixed;
    const int tiles = (numnu + sh.pts - 1) / sh.pts, per = sh.block / n;
    if (per < 1) {
        ctx->err = "LINOP1 gather: block smaller than the layer count";
        return A12_ERR_ARG;
    }
    hipLaunchKernelGGL(g, dim3((tiles + per - 1) / per), dim3(sh.block), 0, ctx->stream, n, numnu,
                       tiles, A.seg.as<int>(), A.l_nu.as<int>(), A.reach.as<int>(), A.state.as<int>(),
                       A.par.as<double>(), A.l_const.as<double>(), A.waveset.as<double>(), tabs,
                       tabs + kTabN, tabs + 2 * kTabN, tabs32,
                       A.xlines.as<double>());
    A12_CHECK_LAUNCH();
    return A12_OK;
}

int a12i_atlas_linop1_run(a12_ctx *ctx, int prec)
{
    AtlasDev &A = ctx->atl;
    if (!A.loaded) {
        ctx->err = "a12_atlas_linop1_run: call a12_atlas_set_state (or a12_atlas_load) first";
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
    A12_HIP(A.reach.reserve(sizeof(int) * nl1));
    A12_HIP(A.greach.reserve(sizeof(unsigned short) * nl1 * ((n + 7) / 8)));
    A12_HIP(A.xlines.reserve(sizeof(double) * (size_t)numnu * n));
    const int nb = (int)((nl1 * n + 255) / 256);  // one lane per (line, depth)
    A12_HIP(A.blk.reserve(sizeof(long long) * (nb + kSumParts)));
    A12_HIP(A.evals.reserve(sizeof(long long)));
    const double *tabs = A.tabs.as<double>();
    if (ctx->prof) A12_HIP(hipEventRecord(ctx->ev[8], s));
    if (nl > 0) {
        hipLaunchKernelGGL(k_lo_setup, dim3(nb), dim3(256), 0, s, nl, n, numnu,
                           A.l_nu.as<int>(), A.l_nucont.as<int>(), A.l_nel.as<int>(),
                           A.l_const.as<double>(), A.waveset.as<double>(), A.tabcont.as<double>(),
                           A.xnfdop.as<double>(), A.dopple.as<double>(), A.depth_lo.as<double>(),
                           tabs, tabs + kTabN, tabs + 2 * kTabN, A.lg2inv,
                           prec == A12_PREC_MIXED ? 1 : 0, A.state.as<int>(),
                           A.par.as<double>(), A.blk.as<long long>());
        A12_CHECK_LAUNCH();
        const int np = std::min(nb, kSumParts);
        long long *part = A.blk.as<long long>() + nb;
        hipLaunchKernelGGL(k_lo_sum, dim3(np), dim3(256), 0, s, A.blk.as<long long>(), nb, part,
                           nullptr);
        A12_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_lo_sum, dim3(1), dim3(256), 0, s, part, np, nullptr,
                           A.evals.as<long long>());
        A12_CHECK_LAUNCH();
    } else {
        A12_HIP(hipMemsetAsync(A.evals.p, 0, sizeof(long long), s));
    }
    if (ctx->prof) A12_HIP(hipEventRecord(ctx->ev[9], s));
    if (nl > 0) {
        hipLaunchKernelGGL(k_lo_linereach, dim3((nl + 63) / 64), dim3(256), 0, s, nl, n,
                           A.state.as<int>(), A.reach.as<int>(), A.greach.as<unsigned short>());
        A12_CHECK_LAUNCH();
    }
    if (ctx->prof) A12_HIP(hipEventRecord(ctx->ev[10], s));
    {
        const int rc = lo_gather_launch(ctx, prec, A.shape);
        if (rc) return rc;
    }
    if (ctx->prof) A12_HIP(hipEventRecord(ctx->ev[11], s));
    A.xl_ready = true;
    A.prof_lo = ctx->prof;
    return A12_OK;
}

int a12i_atlas_shapes(int idx, int32_t *pbb)
{
    if (idx < 0 || idx >= kNumLoShapes) return kNumLoShapes;
    if (pbb) {
        pbb[0] = kLoShapes[idx].pts;
        pbb[1] = kLoShapes[idx].block;
        pbb[2] = kLoShapes[idx].batch;
        pbb[3] = kLoShapes[idx].flat;
    }
    return kNumLoShapes;
}

// Times every gather shape on the loaded store (after one full LINOP1 run
// for the per-(line, depth) state) and keeps the fastest.  All shapes add
// the lines in the same order at every point, so the choice changes the
// time only, never XLINES.
int a12i_atlas_autotune(a12_ctx *ctx, int prec, float *ms, int32_t *best)
{
    AtlasDev &A = ctx->atl;
    const int rc0 = a12i_atlas_linop1_run(ctx, prec);
    if (rc0) return rc0;
    hipEvent_t e0, e1;
    A12_HIP(hipEventCreate(&e0));
    A12_HIP(hipEventCreate(&e1));
    int bi = -1;
    float bt = 0.0f;
    int rc = A12_OK;
    for (int k = 0; k < kNumLoShapes && rc == A12_OK; ++k) {
        float t = -1.0f;
        if (A.n <= kLoShapes[k].block) {
            if ((rc = lo_gather_launch(ctx, prec, k)) != A12_OK) break;  // warm
            if (hipEventRecord(e0, ctx->stream) != hipSuccess) rc = A12_ERR_HIP;
            if (rc == A12_OK) rc = lo_gather_launch(ctx, prec, k);
            if (rc == A12_OK && hipEventRecord(e1, ctx->stream) != hipSuccess) rc = A12_ERR_HIP;
            if (rc == A12_OK && (hipEventSynchronize(e1) != hipSuccess ||
                                 hipEventElapsedTime(&t, e0, e1) != hipSuccess))
                rc = A12_ERR_HIP;
            if (rc == A12_OK && (bi < 0 || t < bt)) {
                bi = k;
                bt = t;
            }
        }
        if (ms) ms[k] = t;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != A12_OK) {
        if (ctx->err.empty()) ctx->err = "a12_atlas_autotune: timing failed";
        return rc;
    }
    A.shape = bi;
    if (best) *best = bi;
    return A12_OK;
}
