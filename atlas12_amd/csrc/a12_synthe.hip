// a12_synthe.hip -- SYNTHE line-opacity build on gfx950.
//
// Replaces the LTE part of SYNTHE's depth loop (src/synthe.f90:476-689).
// Pipeline per call, all on one stream:
//   k_depth_prep   per depth: xnfdop, txnxn (src/synthe.f90:518-551)
//   k_continuum    per (depth, point): FP32 continuum (src/synthe.f90:481-516)
//   k_line_setup   per (depth, line): cutoffs, damping, core/far-wing extent
//                  (src/synthe.f90:580-650) and the exact eval count
//   k_line_emax    per line: widest extent over depths -> extent class
//   k_cls_*        stable partition of the centre-sorted lines by class
//   k_tile_ranges  per (tile, class): the contiguous run of candidates
//   k_tile         per (tile, depth): gather every candidate's contribution
//                  to each point of the tile, accumulated in registers
// No atomics anywhere; every reduction has a fixed order.
//
// The reference scatters each line's profile into buffer(nbuff +- n); the
// gather here adds, at grid point i, profile(|i - nbuff|) for every line
// with |i - nbuff| <= nstep, which is the same set of contributions
// (src/synthe.f90:653-682 clip to [1, length]; the centre is added only for
// on-grid centres, :604-609).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "a12_common.h"
#include "a12_internal.h"

// Wavefront sum of the device library (DPP).
extern "C" __device__ long long __ockl_wfred_add_i64(long long);

namespace {

constexpr int kTile = 256;       // grid points per tile (one per thread)
// MIXED sums the lines of each run of 2^kGroupShift consecutive class
// positions in plain FP32 (at most 15 roundings, relative error <= 15 eps),
// and adds each run's sum into a Neumaier-compensated total.  Runs are fixed
// by position, so every tile shape forms the same partial sums.
constexpr int kGroupShift = 4;
constexpr int kChunk = 256;      // lines staged in LDS per pass
constexpr int kBaseExt = 64;     // extent of class 0
constexpr int kMaxProf = 1000000;  // MAXPROF, src/synthe_module.f90:219

__host__ __device__ inline int ext_class(int emax)
{
    if (emax < 0) return A12_NCLASS - 1;
    int c = 0;
    long r = kBaseExt;
    while (r < emax && c < A12_NCLASS - 2) {
        r <<= 1;
        ++c;
    }
    return c;
}

__host__ __device__ inline long class_radius(int c) { return (long)kBaseExt << c; }

// --------------------------------------------------------------------------
// Per-depth preparation, src/synthe.f90:518-551.  t, xne, xnfh, xnfhe and
// xnfh2 pass through REAL() (FP32) in the reference before use (:392-410).
// dscal[j*4 + 0] = txnxn (FP32 value), [1] = txnxn FP64, [2] = xne rounded
// through FP32, [3] = xne.
// --------------------------------------------------------------------------
__global__ void k_depth_prep(int nrhox, int mw6, const double *__restrict__ xnfpel,
                             const double *__restrict__ dopple, const double *__restrict__ depth,
                             double *__restrict__ xnfdop, double *__restrict__ dscal)
{
#pragma clang fp contract(off)
    const int j = blockIdx.x;
    const double *t = depth, *xne = depth + nrhox, *rho = depth + 2 * nrhox;
    const double *xnfh = depth + 3 * nrhox, *xnfhe1 = depth + 4 * nrhox,
                 *xnfh2 = depth + 5 * nrhox;
    for (int i = threadIdx.x; i < mw6; i += blockDim.x) {
        const double q = xnfpel[(size_t)j * mw6 + i];
        const double d = dopple[(size_t)j * mw6 + i];
        xnfdop[(size_t)j * mw6 + i] = d > 0.0 ? q / rho[j] / d : 0.0;
    }
    if (threadIdx.x == 0) {
        const double t32 = (double)(float)t[j];
        const double h32 = (double)(float)xnfh[j];
        const double he32 = (double)(float)xnfhe1[j];
        const double h232 = (double)(float)xnfh2[j];
        const float tx = (float)((h32 + 0.42 * he32 + 0.85 * h232) * pow(t32 / 10000.0, 0.3));
        const double tx64 = (xnfh[j] + 0.42 * xnfhe1[j] + 0.85 * xnfh2[j]) * pow(t[j] / 10000.0, 0.3);
        dscal[j * 4 + 0] = (double)tx;
        dscal[j * 4 + 1] = tx64;
        dscal[j * 4 + 2] = (double)(float)xne[j];
        dscal[j * 4 + 3] = xne[j];
    }
}

// --------------------------------------------------------------------------
// Continuum at every grid point, src/synthe.f90:481-516: quadratic Lagrange
// interpolation of log10 kappa over the edge interval that brackets the
// point, floored at -15, then 10**x in FP32.  The bracket is the interval
// the reference's monotone advance reaches (edges ascending).
// --------------------------------------------------------------------------
__global__ void k_continuum(int L, int nrhox, double wbegin, double ratio, int nedge,
                            const double *__restrict__ w, int nf,
                            const double *__restrict__ continall, float *__restrict__ cont32,
                            double *__restrict__ cont64)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= L) return;
    const double wave = wbegin * powi_f64(ratio, i);
    // ie = min(#{e >= 1 : w[e] <= wave}, nedge - 2)
    int lo = 1, hi = nedge - 1;  // search w[1..nedge-1]
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        if (w[mid] <= wave) lo = mid + 1;
        else hi = mid - 1;
    }
    int ie = lo - 1;
    if (ie > nedge - 2) ie = nedge - 2;
    if (ie < 0) ie = 0;
    const double w0 = w[ie], w1 = w[ie + 1];
    const double half = (w0 + w1) * 0.5;
    const double del = (w1 - w0) * (w1 - w0) * 0.5;
    const double *ab = continall + (size_t)j * nf + 3 * ie;
    const double x = ((wave - half) * (wave - w1) * ab[0] +
                      (w0 - wave) * (wave - w1) * 2.0 * ab[1] +
                      (wave - w0) * (wave - half) * ab[2]) / del;
    float xf = (float)x;
    if (xf < -15.0f) xf = -15.0f;
    cont32[(size_t)j * L + i] = (float)pow(10.0, (double)xf);
    if (cont64) cont64[(size_t)j * L + i] = pow(10.0, x < -15.0 ? -15.0 : x);
}

// In-grid contributions of one (line, depth) with extent nstep >= 0:
// centre (when on grid) plus the clipped red and blue wings.
__device__ inline long evals_of(int nb, int ns, int L)
{
    long e = (nb >= 1 && nb <= L) ? 1 : 0;
    // red: points nb+1 .. nb+ns within [1, L]
    long rlo = max(nb + 1, 1), rhi = min((long)nb + ns, (long)L);
    if (rhi >= rlo) e += rhi - rlo + 1;
    long blo = max((long)nb - ns, 1L), bhi = min(nb - 1, L);
    if (bhi >= blo) e += bhi - blo + 1;
    return e;
}

// --------------------------------------------------------------------------
// Line set-up per (depth, line), src/synthe.f90:580-650.
// MIXED writes the reference's own FP32/FP64 expression types, so kappa0,
// the damping a, the Doppler step and the extent nstep follow the
// reference's decisions; FP64 keeps every quantity in double.
// The core stops at the first n with kappa0 H(n dv, a) < kappa_min; H is
// decreasing in |v|, so that n is found by bisection after one test at
// n10dop (which also decides whether the n^-2 far wing follows).
// --------------------------------------------------------------------------
template <bool F64>
__global__ void k_line_setup(int nlines, int L, int mw6, double resolu, int nb_shift,
                             const int *__restrict__ nbuff, const int *__restrict__ nelion,
                             const float *__restrict__ cgf, const float *__restrict__ elo,
                             const float *__restrict__ gr, const float *__restrict__ gs,
                             const float *__restrict__ gw, const double *__restrict__ xnfdop,
                             const double *__restrict__ dopple, const double *__restrict__ dscal,
                             const double *__restrict__ hckt, const float *__restrict__ cont32,
                             const double *__restrict__ cont64, int2 *__restrict__ o_ns10,
                             void *__restrict__ o_par,
                             long long *__restrict__ blk_evals)
{
#pragma clang fp contract(off)
    using R = typename std::conditional<F64, double, float>::type;
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    long ev = 0, evc = 0;
    if (l < nlines) {
        const size_t o = (size_t)j * nlines + l;
        const int nb = nbuff[l] - nb_shift;
        const int nel = nelion[l] - 1;
        const int ci = nb < 1 ? 0 : (nb > L ? L - 1 : nb - 1);
        const double dop = dopple[(size_t)j * mw6 + nel];
        const double xd = xnfdop[(size_t)j * mw6 + nel];
        int nstep = -1, n10 = 0;
        R kap0 = 0, adamp = 0, dv = 0, xw = 0;
        if (!F64) {
            float k0 = cgf[l] * (float)xd;
            const float kmin = cont32[(size_t)j * L + ci] * 1.0e-3f;
            if (!(k0 < kmin)) {
                k0 = k0 * (float)exp(-(double)elo[l] * hckt[j]);
                if (!(k0 < kmin)) {
                    const float gwt = gw[l] * (float)dscal[j * 4 + 0];
                    const float a = (float)(((double)gr[l] + (double)gs[l] * dscal[j * 4 + 2] +
                                             (double)gwt) / dop);
                    n10 = (int)(10.0 * dop * resolu);
                    const float dnel = (float)dop;
                    const float d = 1.0f / dnel / (float)resolu;
                    float p10 = 0.0f;
                    bool wing = true;
                    int ne = n10 + 1;
                    if (n10 >= 1) {
                        p10 = k0 * voigt_ref((float)n10 * d, a);
                        if (p10 < kmin) {
                            wing = false;
                            // smallest n in [1, n10] with P(n) < kmin
                            int lo = 1, hi = n10;
                            while (lo < hi) {
                                const int mid = (lo + hi) >> 1;
                                if (k0 * voigt_ref((float)mid * d, a) < kmin) hi = mid;
                                else lo = mid + 1;
                            }
                            ne = lo;
                        }
                    }
                    if (wing) {
                        xw = p10 * ((float)n10 * (float)n10);
                        const float sq = sqrtf(xw / kmin);
                        nstep = sq >= 999999.0f ? kMaxProf : (int)sq + 1;
                    } else {
                        nstep = ne;
                    }
                    kap0 = k0;
                    adamp = a;
                    dv = d;
                }
            }
        } else {
            double k0 = (double)cgf[l] * xd;
            const double kmin = cont64[(size_t)j * L + ci] * 1.0e-3;
            if (!(k0 < kmin)) {
                k0 = k0 * exp(-(double)elo[l] * hckt[j]);
                if (!(k0 < kmin)) {
                    const double a = ((double)gr[l] + (double)gs[l] * dscal[j * 4 + 3] +
                                      (double)gw[l] * dscal[j * 4 + 1]) / dop;
                    n10 = (int)(10.0 * dop * resolu);
                    const double d = 1.0 / dop / resolu;
                    double p10 = 0.0;
                    bool wing = true;
                    int ne = n10 + 1;
                    if (n10 >= 1) {
                        p10 = k0 * voigt_f64((double)n10 * d, a);
                        if (p10 < kmin) {
                            wing = false;
                            int lo = 1, hi = n10;
                            while (lo < hi) {
                                const int mid = (lo + hi) >> 1;
                                if (k0 * voigt_f64((double)mid * d, a) < kmin) hi = mid;
                                else lo = mid + 1;
                            }
                            ne = lo;
                        }
                    }
                    if (wing) {
                        xw = p10 * ((double)n10 * (double)n10);
                        const double sq = sqrt(xw / kmin);
                        nstep = sq >= 999999.0 ? kMaxProf : (int)sq + 1;
                    } else {
                        nstep = ne;
                    }
                    kap0 = k0;
                    adamp = a;
                    dv = d;
                }
            }
        }
        o_ns10[o] = make_int2(nstep, n10);
        using R4 = typename std::conditional<F64, double4, float4>::type;
        R4 q;
        q.x = kap0;
        q.y = adamp;
        q.z = dv;
        q.w = xw;
        static_cast<R4 *>(o_par)[o] = q;
        // the reference skips lines whose whole window is off the grid (:653)
        if (nstep >= 0 && !(nb + nstep < 1 || nb - nstep > L)) {
            ev = evals_of(nb, nstep, L);
            evc = evals_of(nb, min(nstep, n10), L);  // centre + Voigt core
        }
    }
    // block totals of the counts: a DPP sum per wavefront, then the four wave
    // sums (integers: exact in any order)
    __shared__ long long s_w[4], s_wc[4];
    const long long ws = __ockl_wfred_add_i64(ev), wc = __ockl_wfred_add_i64(evc);
    if ((threadIdx.x & 63) == 0) {
        s_w[threadIdx.x >> 6] = ws;
        s_wc[threadIdx.x >> 6] = wc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const size_t b = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
        const size_t nb_all = (size_t)gridDim.x * gridDim.y;
        blk_evals[b] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
        blk_evals[nb_all + b] = s_wc[0] + s_wc[1] + s_wc[2] + s_wc[3];
    }
}

// out[0] = all contributions, out[1] = Voigt-core ones (the rest are n^-2
// far-wing terms).
__global__ void k_sum_evals(const long long *__restrict__ blk, int n, long long *out)
{
    __shared__ long long red[256], redc[256];
    long long s = 0, sc = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        s += blk[i];
        sc += blk[n + i];
    }
    red[threadIdx.x] = s;
    redc[threadIdx.x] = sc;
    __syncthreads();
    for (int k = blockDim.x >> 1; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) {
            red[threadIdx.x] += red[threadIdx.x + k];
            redc[threadIdx.x] += redc[threadIdx.x + k];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = red[0];
        out[1] = redc[0];
    }
}

// Zero fill (graph-capturable, unlike a memset node).
__global__ void k_zero_ll(long long *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Widest extent of each line over all depths, and its extent class.
__global__ void k_line_emax(int nlines, int nrhox, const int2 *__restrict__ ns10,
                            int *__restrict__ emax, int *__restrict__ lclass)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nlines) return;
    int e = -1;
    for (int j = 0; j < nrhox; ++j) e = max(e, ns10[(size_t)j * nlines + l].x);
    emax[l] = e;
    lclass[l] = ext_class(e);
}

// ---- stable partition of the lines by class (counting sort, no atomics) --
constexpr int kPartBlock = 256;

__global__ void k_cls_count(int nlines, const int *__restrict__ lclass, int *__restrict__ cnt)
{
    // cnt[c * nblk + b]
    const int l = blockIdx.x * kPartBlock + threadIdx.x;
    const int c = l < nlines ? lclass[l] : -1;
    for (int k = 0; k < A12_NCLASS; ++k) {
        const int n = __syncthreads_count(c == k);
        if (threadIdx.x == 0) cnt[(size_t)k * gridDim.x + blockIdx.x] = n;
    }
}

// exclusive scan of cnt (A12_NCLASS * nblk ints) by one workgroup
__global__ void k_cls_scan(int n, const int *__restrict__ cnt, int *__restrict__ off)
{
    __shared__ int part[1024];
    const int per = (n + blockDim.x - 1) / blockDim.x;
    const int b = threadIdx.x * per, e = min(b + per, n);
    int s = 0;
    for (int i = b; i < e; ++i) s += cnt[i];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int i = 0; i < (int)blockDim.x; ++i) {
            const int v = part[i];
            part[i] = run;
            run += v;
        }
    }
    __syncthreads();
    int run = part[threadIdx.x];
    for (int i = b; i < e; ++i) {
        off[i] = run;
        run += cnt[i];
    }
    if (threadIdx.x == blockDim.x - 1) off[n] = run;
}

__global__ void k_cls_scatter(int nlines, int nb_shift, const int *__restrict__ lclass,
                              const int *__restrict__ off, const int *__restrict__ nbuff,
                              int *__restrict__ perm, int *__restrict__ nbuffP)
{
    __shared__ int wcnt[kPartBlock / A12_WAVE][A12_NCLASS];
    const int l = blockIdx.x * kPartBlock + threadIdx.x;
    const int c = l < nlines ? lclass[l] : -1;
    const int lane = threadIdx.x & (A12_WAVE - 1), wv = threadIdx.x / A12_WAVE;
    int rank_in_wave = 0;
    for (int k = 0; k < A12_NCLASS; ++k) {
        const unsigned long long m = __ballot(c == k);
        if (lane == 0) wcnt[wv][k] = __popcll(m);
        if (c == k) rank_in_wave = __popcll(m & ((1ull << lane) - 1ull));
    }
    __syncthreads();
    if (c >= 0) {
        int pos = off[(size_t)c * gridDim.x + blockIdx.x] + rank_in_wave;
        for (int w = 0; w < wv; ++w) pos += wcnt[w][c];
        perm[pos] = l;
        nbuffP[pos] = nbuff[l] - nb_shift;
    }
}

// Per (tile, class): [lo, hi) of class positions whose centre lies within the
// class radius of the tile (1-based grid points t0+1 .. t1).
__global__ void k_tile_ranges(int ntiles, int tw, int L, const int *__restrict__ off, int nblk,
                              const int *__restrict__ nbuffP, int *__restrict__ rng)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles * (A12_NCLASS - 1)) return;
    const int tile = t / (A12_NCLASS - 1), c = t % (A12_NCLASS - 1);
    const int cs = off[(size_t)c * nblk], ce = off[(size_t)(c + 1) * nblk];
    const long R = class_radius(c);
    const long g0 = (long)tile * tw + 1, g1 = min((long)(tile + 1) * tw, (long)L);
    const long vlo = g0 - R, vhi = g1 + R;
    int a = cs, b = ce;  // first position with nbuffP >= vlo
    while (a < b) {
        const int m = (a + b) >> 1;
        if (nbuffP[m] < vlo) a = m + 1;
        else b = m;
    }
    const int lo = a;
    b = ce;  // first position with nbuffP > vhi
    while (a < b) {
        const int m = (a + b) >> 1;
        if (nbuffP[m] <= vhi) a = m + 1;
        else b = m;
    }
    rng[2 * t] = lo;
    rng[2 * t + 1] = a;
}

// --------------------------------------------------------------------------
// Tile gather: one workgroup of 256 threads per (tile, depth); a tile is
// 256 * PTS grid points, thread t owning points t, t + 256, ...  Candidate
// lines of each class are staged in LDS CHUNK at a time; lines whose window
// at this depth misses the tile are dropped while staging (wave ballot
// compaction, order preserved).  Each thread then adds the profile of every
// staged line that covers its points:
//   n = |i - nbuff| <= nstep:  n <= n10dop -> kappa0 H(n dv, a)
//                              else        -> x_wing / n^2
// MIXED evaluates H in FP32, the n^-2 wing with the hardware reciprocal
// (1 ulp), sums each run of 16 class positions in plain FP32 and adds the
// run sums with Neumaier compensation; FP64 does all of it in double with
// IEEE division.  Every point sees the lines in the same order
// (class, then centre position) whatever PTS and CHUNK are, so the result
// does not depend on the tile shape.
// --------------------------------------------------------------------------
struct TileArgs {
    int L, nlines;
    const int *rng, *perm, *nbuffP;
    const int2 *ns10;  // per (depth, line): (wing half extent nstep, core half width n10)
    const void *par;  // per (depth, line): (kappa0, a, dv, x_wing), float4 or double4
    void *out;
};

template <bool F64, int PTS, int CHUNK>
__device__ __forceinline__ void tile_body(const TileArgs &a, int tile, int j)
{
#pragma clang fp contract(off)
    using R = typename std::conditional<F64, double, float>::type;
    constexpr int kTw = kTile * PTS;
    __shared__ int s_nb[CHUNK], s_ns[CHUNK], s_n10[CHUNK], s_g[CHUNK];
    __shared__ R s_k0[CHUNK], s_a[CHUNK], s_dv[CHUNK], s_xw[CHUNK];
    __shared__ int s_wcnt[kTile / A12_WAVE];
    __shared__ int s_nact;

    const int tid = threadIdx.x, lane = tid & (A12_WAVE - 1), wv = tid / A12_WAVE;
    const int L = a.L;
    const int g0 = tile * kTw + 1;        // first 1-based point
    const int g1 = min(g0 + kTw - 1, L);  // last
    const size_t dj = (size_t)j * a.nlines;
    using R4 = typename std::conditional<F64, double4, float4>::type;
    const R4 *pa = static_cast<const R4 *>(a.par) + dj;

    Neumaier32 acc32[PTS];
    float part[PTS];  // plain FP32 sum of the current line group
    double acc64[PTS];
    for (int p = 0; p < PTS; ++p) acc64[p] = 0.0, part[p] = 0.0f;
    int grp = -1;

    const int ncls = a.nlines > 0 ? A12_NCLASS - 1 : 0;  // no lines: zeros
    for (int c = 0; c < ncls; ++c) {
        const int lo = a.rng[2 * (tile * (A12_NCLASS - 1) + c)];
        const int hi = a.rng[2 * (tile * (A12_NCLASS - 1) + c) + 1];
        for (int base = lo; base < hi; base += CHUNK) {
            int nact = 0;
            for (int sub = 0; sub < CHUNK; sub += kTile) {  // stage CHUNK candidates
                const int q = base + sub + tid;
                bool act = false;
                int l = 0, nb = 0, ns = -1, n10 = 0;
                if (sub + tid < CHUNK && q < hi) {
                    l = a.perm[q];
                    nb = a.nbuffP[q];
                    const int2 q10 = a.ns10[dj + l];
                    ns = q10.x;
                    n10 = q10.y;
                    act = ns >= 0 && nb - ns <= g1 && nb + ns >= g0;
                }
                const unsigned long long m = __ballot(act);
                if (lane == 0) s_wcnt[wv] = __popcll(m);
                __syncthreads();
                if (act) {
                    int pos = nact + __popcll(m & ((1ull << lane) - 1ull));
                    for (int w = 0; w < wv; ++w) pos += s_wcnt[w];
                    s_nb[pos] = nb;
                    s_ns[pos] = ns;
                    s_g[pos] = q >> kGroupShift;
                    s_n10[pos] = n10;
                    const R4 f = pa[l];
                    s_k0[pos] = f.x;
                    s_a[pos] = f.y;
                    s_dv[pos] = f.z;
                    s_xw[pos] = f.w;
                }
                if (tid == 0) {
                    int sm = nact;
                    for (int w = 0; w < kTile / A12_WAVE; ++w) sm += s_wcnt[w];
                    s_nact = sm;
                }
                __syncthreads();
                nact = s_nact;
                __syncthreads();  // s_wcnt / s_nact are rewritten by the next sub-chunk
            }
            for (int k = 0; k < nact; ++k) {
                const int nb = s_nb[k], ns = s_ns[k], n10 = s_n10[k];
                if (!F64) {
                    const int g = s_g[k];
                    if (g != grp) {
                        for (int p = 0; p < PTS; ++p) acc32[p].add(part[p]), part[p] = 0.0f;
                        grp = g;
                    }
                }
#pragma unroll
                for (int p = 0; p < PTS; ++p) {
                    const int gi = g0 + tid + p * kTile;
                    const int n = abs(gi - nb);
                    if (gi > g1 || n > ns) continue;
                    if (!F64) {
                        float v;
                        if (n <= n10) v = (float)s_k0[k] * voigt_mixed((float)n * (float)s_dv[k], (float)s_a[k]);
                        else v = (float)s_xw[k] * __builtin_amdgcn_rcpf((float)n * (float)n);
                        part[p] += v;
                    } else {
                        double v;
                        if (n <= n10) v = (double)s_k0[k] * voigt_f64((double)n * (double)s_dv[k], (double)s_a[k]);
                        else v = (double)s_xw[k] / ((double)n * (double)n);
                        acc64[p] += v;
                    }
                }
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int p = 0; p < PTS; ++p) {
        const int gi = g0 + tid + p * kTile;
        if (gi > g1) continue;
        acc32[p].add(part[p]);
        if (!F64) static_cast<float *>(a.out)[(size_t)j * L + gi - 1] = acc32[p].value();
        else static_cast<double *>(a.out)[(size_t)j * L + gi - 1] = acc64[p];
    }
}

template <bool F64, int PTS, int CHUNK>
__global__ __launch_bounds__(256) void k_tile(TileArgs a)
{
    tile_body<F64, PTS, CHUNK>(a, blockIdx.x, blockIdx.y);
}

// The whole batch in one persistent launch: work item w = (model, tile,
// depth), models in the order given (largest first), every workgroup
// striding through the items.  item_off (nmodel + 1): first item of each
// model; a model's items run tile-fastest.
struct BatchItemModel {
    TileArgs a;
    int ntiles;
};

template <bool F64, int PTS, int CHUNK>
__global__ __launch_bounds__(256) void k_tile_batch(const BatchItemModel *__restrict__ models,
                                                    const long long *__restrict__ item_off,
                                                    int nmodel)
{
    const long long nitems = item_off[nmodel];
    int m = 0;
    for (long long w = blockIdx.x; w < nitems; w += gridDim.x) {
        while (w >= item_off[m + 1]) ++m;  // items ascend, so the model only moves forward
        const long long r = w - item_off[m];
        const int nt = models[m].ntiles;
        tile_body<F64, PTS, CHUNK>(models[m].a, (int)(r % nt), (int)(r / nt));
    }
}

// Barrier-free form: wavefront w of the block owns the contiguous points
// g0 + w 64 PTS .. + 64 PTS - 1 (lane + 64 p for p < PTS) and works alone:
// it walks the tile's candidate ranges 64 at a time, keeps (ballot) the
// lines whose window at this depth meets its points and adds them in
// candidate order, lowest bit first -- the order of every other shape.  A
// kept line's NU and extent come from its lane by shuffle, its other
// parameters by wave-uniform loads fetched one line ahead.
template <bool F64, int PTS>
__device__ __forceinline__ void tile_body_w(const TileArgs &a, int tile, int j)
{
#pragma clang fp contract(off)
    using R = typename std::conditional<F64, double, float>::type;
    constexpr int kTw = kTile * PTS;
    const int tid = threadIdx.x, lane = tid & (A12_WAVE - 1), wv = tid / A12_WAVE;
    const int L = a.L;
    const int g0 = tile * kTw + 1;
    const int g1 = min(g0 + kTw - 1, L);
    const int w0 = g0 + wv * A12_WAVE * PTS;            // this wave's first point
    const int w1 = min(w0 + A12_WAVE * PTS - 1, g1);    // and last
    const size_t dj = (size_t)j * a.nlines;
    using R4 = typename std::conditional<F64, double4, float4>::type;
    const R4 *pa = static_cast<const R4 *>(a.par) + dj;
    Neumaier32 acc32[PTS];
    float part[PTS];  // plain FP32 sum of the current line group
    double acc64[PTS];
    for (int p = 0; p < PTS; ++p) acc64[p] = 0.0, part[p] = 0.0f;
    int grp = -1;
    const int ncls = (a.nlines > 0 && w0 <= g1) ? A12_NCLASS - 1 : 0;
    for (int c = 0; c < ncls; ++c) {
        const int lo = a.rng[2 * (tile * (A12_NCLASS - 1) + c)];
        const int hi = a.rng[2 * (tile * (A12_NCLASS - 1) + c) + 1];
        for (int c0 = lo; c0 < hi; c0 += A12_WAVE) {
            const int q = c0 + lane;
            int l = 0, nb = 0, ns = -1, n10 = 0;
            bool act = false;
            if (q < hi) {
                l = a.perm[q];
                nb = a.nbuffP[q];
                const int2 q10 = a.ns10[dj + l];
                ns = q10.x;
                n10 = q10.y;
                act = ns >= 0 && nb - ns <= w1 && nb + ns >= w0;
            }
            unsigned long long m = __ballot(act);
            if (!m) continue;
            int b = __builtin_ctzll(m);
            m &= m - 1;
            int cq = c0 + b;
            int cl = __shfl(l, b), cnb = __shfl(nb, b), cns = __shfl(ns, b);
            int cn10 = __shfl(n10, b);
            R4 cf = pa[cl];
            R ck0 = cf.x, ca = cf.y, cdv = cf.z, cxw = cf.w;
            for (;;) {
                const bool more = m != 0;
                int nl_ = 0, nnb = 0, nns = 0, nn10 = 0, nq = 0;
                R nk0 = 0, na = 0, ndv = 0, nxw = 0;
                if (more) {
                    const int bn = __builtin_ctzll(m);
                    m &= m - 1;
                    nq = c0 + bn;
                    nl_ = __shfl(l, bn);
                    nnb = __shfl(nb, bn);
                    nns = __shfl(ns, bn);
                    nn10 = __shfl(n10, bn);
                    const R4 nf = pa[nl_];
                    nk0 = nf.x;
                    na = nf.y;
                    ndv = nf.z;
                    nxw = nf.w;
                }
                if (!F64 && (cq >> kGroupShift) != grp) {
                    for (int p = 0; p < PTS; ++p) acc32[p].add(part[p]), part[p] = 0.0f;
                    grp = cq >> kGroupShift;
                }
                if (!F64 && (cnb + cn10 < w0 || cnb - cn10 > w1)) {
                    // no core point of this line in the wave's points: every item
                    // is an n^-2 wing term; points outside its extent add exactly 0
                    const float xw = (float)cxw;
#pragma unroll
                    for (int p = 0; p < PTS; ++p) {
                        const int gi = w0 + p * A12_WAVE + lane;
                        const int n = abs(gi - cnb);
                        const float w = xw * __builtin_amdgcn_rcpf((float)n * (float)n);
                        part[p] += (gi <= w1 && n <= cns) ? w : 0.0f;
                    }
                } else {
#pragma unroll
                    for (int p = 0; p < PTS; ++p) {
                        const int gi = w0 + p * A12_WAVE + lane;
                        const int n = abs(gi - cnb);
                        if (gi > w1 || n > cns) continue;
                        if (!F64) {
                            float v;
                            if (n <= cn10) v = (float)ck0 * voigt_mixed((float)n * (float)cdv, (float)ca);
                            else v = (float)cxw * __builtin_amdgcn_rcpf((float)n * (float)n);
                            part[p] += v;
                        } else {
                            double v;
                            if (n <= cn10) v = (double)ck0 * voigt_f64((double)n * (double)cdv, (double)ca);
                            else v = (double)cxw / ((double)n * (double)n);
                            acc64[p] += v;
                        }
                    }
                }
                if (!more) break;
                cq = nq;
                cl = nl_;
                cnb = nnb;
                cns = nns;
                cn10 = nn10;
                ck0 = nk0;
                ca = na;
                cdv = ndv;
                cxw = nxw;
            }
        }
    }
#pragma unroll
    for (int p = 0; p < PTS; ++p) {
        const int gi = w0 + p * A12_WAVE + lane;
        if (gi > w1) continue;
        acc32[p].add(part[p]);
        if (!F64) static_cast<float *>(a.out)[(size_t)j * L + gi - 1] = acc32[p].value();
        else static_cast<double *>(a.out)[(size_t)j * L + gi - 1] = acc64[p];
    }
}

template <bool F64, int PTS>
__global__ __launch_bounds__(256) void k_tile_w(TileArgs a)
{
    tile_body_w<F64, PTS>(a, blockIdx.x, blockIdx.y);
}

template <bool F64, int PTS>
__global__ __launch_bounds__(256) void k_tile_batch_w(const BatchItemModel *__restrict__ models,
                                                      const long long *__restrict__ item_off,
                                                      int nmodel)
{
    const long long nitems = item_off[nmodel];
    int m = 0;
    for (long long w = blockIdx.x; w < nitems; w += gridDim.x) {
        while (w >= item_off[m + 1]) ++m;
        const long long r = w - item_off[m];
        const int nt = models[m].ntiles;
        tile_body_w<F64, PTS>(models[m].a, (int)(r % nt), (int)(r / nt));
    }
}

// Gather shapes: points per thread and lines per LDS stage.
struct TileShape {
    int pts, chunk;
};
// chunk 0: the barrier-free per-wave form
constexpr TileShape kTileShapes[] = {{1, 256}, {1, 128}, {1, 512}, {2, 256}, {2, 512},
                                     {4, 256}, {1, 0},   {2, 0},   {4, 0},   {8, 0},
                                     {4, 512}, {8, 256}};
constexpr int kNumTileShapes = sizeof(kTileShapes) / sizeof(kTileShapes[0]);

template <bool F64>
using TileKernel = void (*)(TileArgs);
template <bool F64>
using TileBatchKernel = void (*)(const BatchItemModel *, const long long *, int);

template <bool F64>
TileKernel<F64> tile_kernel(int shape)
{
    switch (shape) {
    case 1: return k_tile<F64, 1, 128>;
    case 2: return k_tile<F64, 1, 512>;
    case 3: return k_tile<F64, 2, 256>;
    case 4: return k_tile<F64, 2, 512>;
    case 5: return k_tile<F64, 4, 256>;
    case 6: return k_tile_w<F64, 1>;
    case 7: return k_tile_w<F64, 2>;
    case 8: return k_tile_w<F64, 4>;
    case 9: return k_tile_w<F64, 8>;
    case 10: return k_tile<F64, 4, 512>;
    case 11: return k_tile<F64, 8, 256>;
    default: return k_tile<F64, 1, 256>;
    }
}

template <bool F64>
TileBatchKernel<F64> tile_batch_kernel(int shape)
{
    switch (shape) {
    case 1: return k_tile_batch<F64, 1, 128>;
    case 2: return k_tile_batch<F64, 1, 512>;
    case 3: return k_tile_batch<F64, 2, 256>;
    case 4: return k_tile_batch<F64, 2, 512>;
    case 5: return k_tile_batch<F64, 4, 256>;
    case 6: return k_tile_batch_w<F64, 1>;
    case 7: return k_tile_batch_w<F64, 2>;
    case 8: return k_tile_batch_w<F64, 4>;
    case 9: return k_tile_batch_w<F64, 8>;
    case 10: return k_tile_batch<F64, 4, 512>;
    case 11: return k_tile_batch<F64, 8, 256>;
    default: return k_tile_batch<F64, 1, 256>;
    }
}

// Batched H(v, a): prec 0 mixed FP32 path, 1 FP64 (the reference's
// arithmetic), 2 the reference's FP32-rounded result.
__global__ void k_voigt(int n, int prec, const float *__restrict__ v, const float *__restrict__ a,
                        double *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (prec == 0) out[i] = (double)voigt_mixed(v[i], a[i]);
    else if (prec == 1) out[i] = voigt_f64((double)v[i], (double)a[i]);
    else out[i] = (double)voigt_ref(v[i], a[i]);
}

template <class T>
hipError_t upload(DevBuf &b, const T *src, size_t n, hipStream_t s)
{
    hipError_t e = b.reserve(n * sizeof(T));
    if (e != hipSuccess) return e;
    return hipMemcpyAsync(b.p, src, n * sizeof(T), hipMemcpyHostToDevice, s);
}


}  // namespace

// --------------------------------------------------------------------------
// Host side
// --------------------------------------------------------------------------
int a12i_synthe_load(a12_ctx *ctx, const a12_synthe_atm *atm, const a12_lte_line *lines,
                     int32_t nlines)
{
    SyntheDev &S = ctx->syn;
    if (!atm || (!lines && nlines > 0) || nlines < 0) {
        ctx->err = "a12_synthe_load: null argument";
        return A12_ERR_ARG;
    }
    if (atm->nrhox < 2 || atm->nrhox > A12_NRMAX_LAYERS || atm->length < 1 || atm->nedge < 2 ||
        atm->nf < 3 * (atm->nedge - 1) || atm->mw6 < 6 || !atm->wledge || !atm->continall ||
        !atm->xnfpel || !atm->dopple || !atm->t || !atm->xne || !atm->rho || !atm->xnfh ||
        !atm->xnfhe1 || !atm->xnfh2 || !atm->hckt || !atm->hkt || !atm->hkt_atlas ||
        !atm->rhox || !(atm->ratio > 1.0) || !(atm->resolu > 0.0) || !(atm->wbegin > 0.0)) {
        ctx->err = "a12_synthe_load: bad atmosphere dimensions or null array";
        return A12_ERR_ARG;
    }
    for (int e = 1; e < atm->nedge; ++e)
        if (!(atm->wledge[e] > atm->wledge[e - 1])) {
            ctx->err = "a12_synthe_load: continuum edges must be strictly ascending";
            return A12_ERR_ARG;
        }
    for (int i = 0; i < nlines; ++i)
        if (lines[i].nelion < 1 || lines[i].nelion > atm->mw6) {
            ctx->err = "a12_synthe_load: line " + std::to_string(i) + " has nelion out of range";
            return A12_ERR_ARG;
        }
    hipStream_t st = ctx->stream;
    S.nrhox = atm->nrhox;
    S.L = atm->length;
    S.nedge = atm->nedge;
    S.nf = atm->nf;
    S.mw6 = atm->mw6;
    S.nlines = nlines;
    S.wbegin = atm->wbegin;
    S.ratio = atm->ratio;
    S.resolu = atm->resolu;
    const int n = S.nrhox;
    A12_HIP(upload(S.wledge, atm->wledge, (size_t)atm->nedge, st));
    A12_HIP(upload(S.continall, atm->continall, (size_t)atm->nf * n, st));
    if (atm->contabs) A12_HIP(upload(S.contabs, atm->contabs, (size_t)atm->nf * n, st));
    if (atm->contscat) A12_HIP(upload(S.contscat, atm->contscat, (size_t)atm->nf * n, st));
    A12_HIP(upload(S.xnfpel, atm->xnfpel, (size_t)atm->mw6 * n, st));
    A12_HIP(upload(S.dopple, atm->dopple, (size_t)atm->mw6 * n, st));
    std::vector<double> depth((size_t)10 * n);
    const double *src[10] = {atm->t, atm->xne, atm->rho, atm->xnfh, atm->xnfhe1,
                             atm->xnfh2, atm->hckt, atm->hkt, atm->hkt_atlas, atm->rhox};
    for (int k = 0; k < 10; ++k) std::memcpy(&depth[(size_t)k * n], src[k], sizeof(double) * n);
    A12_HIP(upload(S.depth, depth.data(), depth.size(), st));
    // lines sorted by centre (stable: equal centres keep input order)
    std::vector<int> order(nlines);
    std::iota(order.begin(), order.end(), 0);
    bool sorted = true;
    for (int i = 1; i < nlines; ++i)
        if (lines[i].nbuff < lines[i - 1].nbuff) { sorted = false; break; }
    if (!sorted)
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return lines[a].nbuff < lines[b].nbuff; });
    std::vector<int> nb(nlines), ne(nlines);
    std::vector<float> cg(nlines), el(nlines), g1(nlines), g2(nlines), g3(nlines);
    for (int i = 0; i < nlines; ++i) {
        const a12_lte_line &r = lines[order[i]];
        nb[i] = r.nbuff; ne[i] = r.nelion; cg[i] = r.cgf; el[i] = r.elo;
        g1[i] = r.gamrf; g2[i] = r.gamsf; g3[i] = r.gamwf;
    }
    const size_t nl = (size_t)std::max(nlines, 1);
    A12_HIP(upload(S.l_nbuff, nb.data(), nl * (nlines > 0), st));
    A12_HIP(upload(S.l_nelion, ne.data(), nl * (nlines > 0), st));
    A12_HIP(upload(S.l_cgf, cg.data(), nl * (nlines > 0), st));
    A12_HIP(upload(S.l_elo, el.data(), nl * (nlines > 0), st));
    A12_HIP(upload(S.l_gr, g1.data(), nl * (nlines > 0), st));
    A12_HIP(upload(S.l_gs, g2.data(), nl * (nlines > 0), st));
    A12_HIP(upload(S.l_gw, g3.data(), nl * (nlines > 0), st));
    A12_HIP(hipStreamSynchronize(st));
    S.loaded = true;
    ++S.gen;
    a12_cplx_clear(ctx);  // complex lines belong to the atmosphere they were loaded with
    S.opac_prec = -1;
    S.transferred = false;
    return A12_OK;
}

int a12i_synthe_prep(a12_ctx *ctx, double *xnfdop, double *dscal, float *cont32)
{
    SyntheDev &S = ctx->syn;
    if (!S.loaded) {
        ctx->err = "call a12_synthe_load first";
        return A12_ERR_STATE;
    }
    hipStream_t st = ctx->stream;
    const int n = S.nrhox, L = S.L;
    hipLaunchKernelGGL(k_depth_prep, dim3(n), dim3(256), 0, st, n, S.mw6, S.xnfpel.as<double>(),
                       S.dopple.as<double>(), S.depth.as<double>(), xnfdop, dscal);
    A12_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_continuum, dim3((L + 255) / 256, n), dim3(256), 0, st, L, n, S.wbegin, S.ratio,
                       S.nedge, S.wledge.as<double>(), S.nf, S.continall.as<double>(), cont32,
                       (double *)nullptr);
    A12_CHECK_LAUNCH();
    return A12_OK;
}

namespace {

// One model's device arrays for the opacity pipeline: inputs, work arrays
// and the output.  nb_shift is subtracted from every nbuff of the shared line
// arrays (0 for a single model; the model's window offset in a batch).
struct SynModel {
    int n = 0, L = 0, N = 0, mw6 = 0, nedge = 0, nf = 0, nb_shift = 0;
    double wbegin = 0, ratio = 0, resolu = 0;
    const double *wledge = nullptr, *continall = nullptr, *xnfpel = nullptr,
                 *dopple = nullptr, *depth = nullptr;
    const int *l_nbuff = nullptr, *l_nelion = nullptr;
    const float *l_cgf = nullptr, *l_elo = nullptr, *l_gr = nullptr, *l_gs = nullptr,
                *l_gw = nullptr;
    double *xnfdop = nullptr, *dscal = nullptr, *cont64 = nullptr;
    float *cont32 = nullptr;
    int2 *p_ns10 = nullptr;
    void *p_par = nullptr;
    int *emax = nullptr, *lclass = nullptr, *cls_cnt = nullptr, *cls_off = nullptr,
        *perm = nullptr, *nbuffP = nullptr, *cls_range = nullptr;
    long long *blk_evals = nullptr, *evals_out = nullptr;
    void *opacity = nullptr;
};

int tile_width(int shape) { return kTile * kTileShapes[shape].pts; }
int ntiles_of(int L, int shape) { return (L + tile_width(shape) - 1) / tile_width(shape); }
int nblk_lines(int N) { return (std::max(N, 1) + 255) / 256; }
int nblk_part(int N) { return (std::max(N, 1) + kPartBlock - 1) / kPartBlock; }

TileArgs tile_args(const SynModel &M)
{
    TileArgs a;
    a.L = M.L;
    a.nlines = M.N;
    a.rng = M.cls_range;
    a.perm = M.perm;
    a.nbuffP = M.nbuffP;
    a.ns10 = M.p_ns10;
    a.par = M.p_par;
    a.out = M.opacity;
    return a;
}

int launch_ranges(a12_ctx *ctx, const SynModel &M, int shape, hipStream_t st)
{
    if (M.N == 0) return A12_OK;
    const int ntiles = ntiles_of(M.L, shape);
    const int nr = ntiles * (A12_NCLASS - 1);
    hipLaunchKernelGGL(k_tile_ranges, dim3((nr + 255) / 256), dim3(256), 0, st, ntiles,
                       tile_width(shape), M.L, M.cls_off, nblk_part(M.N), M.nbuffP, M.cls_range);
    A12_CHECK_LAUNCH();
    return A12_OK;
}

// Everything before the gather, on stream st: per-depth prep, continuum,
// line set-up with the evaluation count, extent classes, the class
// partition and the per-tile candidate ranges for tile shape `shape`.
// ev (may be null): events recorded after the continuum and the set-up.
int launch_prep(a12_ctx *ctx, const SynModel &M, bool f64, int shape, hipStream_t st,
                hipEvent_t ev_cont, hipEvent_t ev_setup)
{
    const int n = M.n, L = M.L, N = M.N;
    hipLaunchKernelGGL(k_depth_prep, dim3(n), dim3(256), 0, st, n, M.mw6, M.xnfpel, M.dopple,
                       M.depth, M.xnfdop, M.dscal);
    A12_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_continuum, dim3((L + 255) / 256, n), dim3(256), 0, st, L, n, M.wbegin,
                       M.ratio, M.nedge, M.wledge, M.nf, M.continall, M.cont32,
                       f64 ? M.cont64 : nullptr);
    A12_CHECK_LAUNCH();
    if (ev_cont) A12_HIP(hipEventRecord(ev_cont, st));
    if (N == 0) {
        hipLaunchKernelGGL(k_zero_ll, dim3(1), dim3(64), 0, st, M.evals_out, 2);
        A12_CHECK_LAUNCH();
        return A12_OK;
    }
    const int nblkl = nblk_lines(N), nblk = nblk_part(N);
    auto setup = f64 ? k_line_setup<true> : k_line_setup<false>;
    hipLaunchKernelGGL(setup, dim3(nblkl, n), dim3(256), 0, st, N, L, M.mw6, M.resolu, M.nb_shift,
                       M.l_nbuff, M.l_nelion, M.l_cgf, M.l_elo, M.l_gr, M.l_gs, M.l_gw, M.xnfdop,
                       M.dopple, M.dscal, M.depth + 6 * n, M.cont32, f64 ? M.cont64 : nullptr,
                       M.p_ns10, M.p_par, M.blk_evals);
    A12_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_sum_evals, dim3(1), dim3(256), 0, st, M.blk_evals, nblkl * n,
                       M.evals_out);
    A12_CHECK_LAUNCH();
    if (ev_setup) A12_HIP(hipEventRecord(ev_setup, st));
    hipLaunchKernelGGL(k_line_emax, dim3(nblkl), dim3(256), 0, st, N, n, M.p_ns10, M.emax,
                       M.lclass);
    A12_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_cls_count, dim3(nblk), dim3(kPartBlock), 0, st, N, M.lclass, M.cls_cnt);
    A12_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_cls_scan, dim3(1), dim3(1024), 0, st, A12_NCLASS * nblk, M.cls_cnt,
                       M.cls_off);
    A12_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_cls_scatter, dim3(nblk), dim3(kPartBlock), 0, st, N, M.nb_shift,
                       M.lclass, M.cls_off, M.l_nbuff, M.perm, M.nbuffP);
    A12_CHECK_LAUNCH();
    return launch_ranges(ctx, M, shape, st);
}

int launch_gather(a12_ctx *ctx, const SynModel &M, bool f64, int shape, hipStream_t st)
{
    auto tk = f64 ? tile_kernel<true>(shape) : tile_kernel<false>(shape);
    hipLaunchKernelGGL(tk, dim3(ntiles_of(M.L, shape), M.n), dim3(256), 0, st, tile_args(M));
    A12_CHECK_LAUNCH();
    return A12_OK;
}

// Work-array bytes of one model (everything but inputs and the opacity).
size_t work_bytes(int n, int L, int N, int mw6, bool f64, int max_tiles)
{
    const size_t rb = f64 ? sizeof(double) : sizeof(float);
    const size_t nl = (size_t)std::max(N, 1);
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    size_t b = al(sizeof(double) * mw6 * n) + al(sizeof(double) * 4 * n) +
               al(sizeof(float) * (size_t)L * n) + (f64 ? al(sizeof(double) * (size_t)L * n) : 0);
    b += al(sizeof(int2) * nl * n) + al(4 * rb * nl * n);  // (nstep, n10), profile records
    b += 2 * al(sizeof(int) * nl);
    b += al(2 * sizeof(long long) * (size_t)nblk_lines(N) * n) + al(2 * sizeof(long long));
    b += al(sizeof(int) * (size_t)A12_NCLASS * nblk_part(N)) +
         al(sizeof(int) * ((size_t)A12_NCLASS * nblk_part(N) + 1));
    b += 2 * al(sizeof(int) * nl) + al(sizeof(int) * 2 * (size_t)max_tiles * (A12_NCLASS - 1));
    return b;
}

// Carve a model's work arrays out of one allocation (same order as
// work_bytes).
void carve_work(SynModel &M, char *base, bool f64, int max_tiles)
{
    const size_t rb = f64 ? sizeof(double) : sizeof(float);
    const int n = M.n, L = M.L, N = M.N;
    const size_t nl = (size_t)std::max(N, 1);
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    char *p = base;
    auto take = [&](size_t bytes) {
        char *q = p;
        p += al(bytes);
        return q;
    };
    M.xnfdop = reinterpret_cast<double *>(take(sizeof(double) * M.mw6 * n));
    M.dscal = reinterpret_cast<double *>(take(sizeof(double) * 4 * n));
    M.cont32 = reinterpret_cast<float *>(take(sizeof(float) * (size_t)L * n));
    M.cont64 = f64 ? reinterpret_cast<double *>(take(sizeof(double) * (size_t)L * n)) : nullptr;
    M.p_ns10 = reinterpret_cast<int2 *>(take(sizeof(int2) * nl * n));
    M.p_par = take(4 * rb * nl * n);
    M.emax = reinterpret_cast<int *>(take(sizeof(int) * nl));
    M.lclass = reinterpret_cast<int *>(take(sizeof(int) * nl));
    M.blk_evals = reinterpret_cast<long long *>(take(2 * sizeof(long long) * (size_t)nblk_lines(N) * n));
    M.evals_out = reinterpret_cast<long long *>(take(2 * sizeof(long long)));
    M.cls_cnt = reinterpret_cast<int *>(take(sizeof(int) * (size_t)A12_NCLASS * nblk_part(N)));
    M.cls_off = reinterpret_cast<int *>(take(sizeof(int) * ((size_t)A12_NCLASS * nblk_part(N) + 1)));
    M.perm = reinterpret_cast<int *>(take(sizeof(int) * nl));
    M.nbuffP = reinterpret_cast<int *>(take(sizeof(int) * nl));
    M.cls_range = reinterpret_cast<int *>(take(sizeof(int) * 2 * (size_t)max_tiles * (A12_NCLASS - 1)));
}

}  // namespace

int a12i_synthe_run(a12_ctx *ctx, int prec)
{
    SyntheDev &S = ctx->syn;
    if (!S.loaded) {
        ctx->err = "a12_synthe_opacity_run: call a12_synthe_load first";
        return A12_ERR_STATE;
    }
    if (prec != A12_PREC_MIXED && prec != A12_PREC_FP64) {
        ctx->err = "a12_synthe_opacity_run: prec must be A12_PREC_MIXED or A12_PREC_FP64";
        return A12_ERR_ARG;
    }
    const bool f64 = prec == A12_PREC_FP64;
    hipStream_t st = ctx->stream;
    const int n = S.nrhox, L = S.L, N = S.nlines;
    const size_t rb = f64 ? sizeof(double) : sizeof(float);
    const int max_tiles = (L + kTile - 1) / kTile;  // the narrowest shape has the most tiles
    A12_HIP(S.work.reserve(work_bytes(n, L, N, S.mw6, f64, max_tiles)));
    A12_HIP(S.opacity.reserve(rb * (size_t)L * n));
    SynModel M;
    M.n = n;
    M.L = L;
    M.N = N;
    M.mw6 = S.mw6;
    M.nedge = S.nedge;
    M.nf = S.nf;
    M.wbegin = S.wbegin;
    M.ratio = S.ratio;
    M.resolu = S.resolu;
    M.wledge = S.wledge.as<double>();
    M.continall = S.continall.as<double>();
    M.xnfpel = S.xnfpel.as<double>();
    M.dopple = S.dopple.as<double>();
    M.depth = S.depth.as<double>();
    M.l_nbuff = S.l_nbuff.as<int>();
    M.l_nelion = S.l_nelion.as<int>();
    M.l_cgf = S.l_cgf.as<float>();
    M.l_elo = S.l_elo.as<float>();
    M.l_gr = S.l_gr.as<float>();
    M.l_gs = S.l_gs.as<float>();
    M.l_gw = S.l_gw.as<float>();
    M.opacity = S.opacity.p;
    carve_work(M, S.work.as<char>(), f64, max_tiles);
    const int shape = S.shape;
    if (ctx->prof) A12_HIP(hipEventRecord(ctx->ev[0], st));
    int rc = launch_prep(ctx, M, f64, shape, st, ctx->prof ? ctx->ev[1] : nullptr,
                         ctx->prof ? ctx->ev[2] : nullptr);
    if (rc != A12_OK) return rc;
    if (ctx->prof) A12_HIP(hipEventRecord(ctx->ev[3], st));
    rc = launch_gather(ctx, M, f64, shape, st);
    if (rc != A12_OK) return rc;
    if (ctx->prof) A12_HIP(hipEventRecord(ctx->ev[4], st));
    // complex-profile lines, when loaded: compute_line_opacity's buffer is
    // the depth's starting BUFFER in the reference (src/synthe.f90:552)
    if (a12i_cplx_active(ctx)) {
        rc = a12i_cplx_add(ctx, M.xnfdop, M.cont32, M.opacity, f64);
        if (rc != A12_OK) return rc;
    }
    S.prof_valid = ctx->prof && N > 0;
    S.evals_dev = M.evals_out;
    S.opac_prec = prec;
    S.transferred = false;
    return A12_OK;
}

int a12i_synthe_shapes(int idx, int32_t *pts_chunk)
{
    if (idx < 0 || idx >= kNumTileShapes) return -1;
    if (pts_chunk) {
        pts_chunk[0] = kTileShapes[idx].pts;
        pts_chunk[1] = kTileShapes[idx].chunk;
    }
    return kNumTileShapes;
}

int a12i_synthe_set_shape(a12_ctx *ctx, int shape)
{
    if (shape < 0 || shape >= kNumTileShapes) {
        ctx->err = "a12_synthe_set_shape: no such shape";
        return A12_ERR_ARG;
    }
    ctx->syn.shape = shape;
    return A12_OK;
}

int a12i_voigt(a12_ctx *ctx, int prec, int n, const float *v, const float *a, double *h)
{
    if (n < 0 || prec < 0 || prec > 2 || (n > 0 && (!v || !a || !h))) {
        ctx->err = "a12_voigt_profile: bad arguments";
        return A12_ERR_ARG;
    }
    if (n == 0) return A12_OK;
    DevBuf &b = ctx->jb_in;
    // layout: v (n floats), a (n floats), pad to 8 bytes, h (n doubles)
    const size_t off_h = ((sizeof(float) * 2 * (size_t)n + 7) / 8) * 8;
    A12_HIP(b.reserve(off_h + sizeof(double) * (size_t)n));
    float *dv = b.as<float>(), *da = dv + n;
    double *dh = reinterpret_cast<double *>(static_cast<char *>(b.p) + off_h);
    A12_HIP(hipMemcpyAsync(dv, v, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    A12_HIP(hipMemcpyAsync(da, a, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_voigt, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, prec, dv, da, dh);
    A12_CHECK_LAUNCH();
    A12_HIP(hipMemcpyAsync(h, dh, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    A12_HIP(hipStreamSynchronize(ctx->stream));
    return A12_OK;
}

// ==========================================================================
// Batched SYNTHE: many models, each with its own atmosphere, depth count and
// wavelength window, over one shared line list sorted by centre.  A model's
// lines are a contiguous range of that list and its nbuff is the list's
// nbuff minus the model's window offset (all windows on one log-lambda
// grid).  Per group of models whose work arrays fit the work budget: the
// set-up kernels per model, then ONE persistent gather launch over every
// (model, tile, depth) item of the group; then, optionally, the FP64
// transfer per model.  The whole sequence can be captured once into a
// hipGraph and replayed.  Results per model are those of the single-model
// pipeline on the same inputs, bit for bit.
// ==========================================================================
struct SynBatch {
    int nmodel = 0, mw6 = 0;
    int64_t nlines = 0;
    std::vector<SynModel> m;           // inputs and output set; work carved per run
    std::vector<int> order;            // models, largest first (gather item order)
    std::vector<size_t> in_off, out_off, h_off;
    std::vector<int> has_tables;       // contabs/contscat given
    std::vector<const double *> cab, csc;  // device contabs / contscat per model
    int grid = 0;                      // persistent gather workgroups
    DevBuf lines, inputs, opacity, hout, work, evals, desc, item_off;
    std::vector<std::vector<int>> groups;
    std::vector<size_t> wk_off;        // work offset of each model within its group
    size_t budget = 0, max_model_work = 0;
    int shape = 0, prec = -1;
    bool transfer = false, ran = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_prec = -1, graph_shape = -1;
    bool graph_transfer = false;
    void drop_graph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        graph_prec = graph_shape = -1;
    }
};

void a12i_synthe_batch_free(a12_ctx *ctx)
{
    SynBatch *B = ctx->sbatch;
    if (!B) return;
    B->drop_graph();
    for (DevBuf *b : {&B->lines, &B->inputs, &B->opacity, &B->hout, &B->work, &B->evals, &B->desc,
                      &B->item_off})
        b->release();
    delete B;
    ctx->sbatch = nullptr;
}

int a12i_synthe_batch_load(a12_ctx *ctx, int32_t nmodel, const a12_synthe_atm *atms,
                           const int32_t *nb_shift, const int64_t *line_lo,
                           const int64_t *line_hi, const a12_lte_line *lines, int64_t nlines)
{
    a12i_synthe_batch_free(ctx);
    if (nmodel < 1 || !atms || !nb_shift || !line_lo || !line_hi || nlines < 0 ||
        (nlines > 0 && !lines) || nlines > INT32_MAX) {
        ctx->err = "a12_synthe_batch_load: bad arguments";
        return A12_ERR_ARG;
    }
    for (int64_t i = 1; i < nlines; ++i)
        if (lines[i].nbuff < lines[i - 1].nbuff) {
            ctx->err = "a12_synthe_batch_load: the shared line list must be sorted by nbuff";
            return A12_ERR_ARG;
        }
    const int mw6 = atms[0].mw6;
    for (int k = 0; k < nmodel; ++k) {
        const a12_synthe_atm *a = &atms[k];
        const std::string id = "a12_synthe_batch_load: model " + std::to_string(k) + ": ";
        if (a->nrhox < 2 || a->nrhox > A12_NRMAX_LAYERS || a->length < 1 || a->nedge < 2 ||
            a->nf < 3 * (a->nedge - 1) || a->mw6 != mw6 || mw6 < 6 || !a->wledge ||
            !a->continall || !a->xnfpel || !a->dopple || !a->t || !a->xne || !a->rho ||
            !a->xnfh || !a->xnfhe1 || !a->xnfh2 || !a->hckt || !a->hkt || !a->hkt_atlas ||
            !a->rhox || !(a->ratio > 1.0) || !(a->resolu > 0.0) || !(a->wbegin > 0.0) ||
            (!a->contabs) != (!a->contscat)) {
            ctx->err = id + "bad dimensions or null array (mw6 must be the same for all)";
            return A12_ERR_ARG;
        }
        for (int e = 1; e < a->nedge; ++e)
            if (!(a->wledge[e] > a->wledge[e - 1])) {
                ctx->err = id + "continuum edges must be strictly ascending";
                return A12_ERR_ARG;
            }
        if (line_lo[k] < 0 || line_hi[k] < line_lo[k] || line_hi[k] > nlines ||
            line_hi[k] - line_lo[k] > INT32_MAX) {
            ctx->err = id + "line range outside the list";
            return A12_ERR_ARG;
        }
    }
    for (int64_t i = 0; i < nlines; ++i)
        if (lines[i].nelion < 1 || lines[i].nelion > mw6) {
            ctx->err = "a12_synthe_batch_load: line " + std::to_string(i) + " has nelion out of range";
            return A12_ERR_ARG;
        }
    SynBatch *B = new SynBatch;
    ctx->sbatch = B;
    B->nmodel = nmodel;
    B->mw6 = mw6;
    B->nlines = nlines;
    hipStream_t st = ctx->stream;
    // shared lines (SoA): nbuff, nelion, cgf, elo, gr, gs, gw
    const size_t nl = (size_t)std::max<int64_t>(nlines, 1);
    A12_HIP(B->lines.reserve(7 * 4 * nl));
    {
        std::vector<int32_t> buf(7 * nl, 0);
        int32_t *nb = buf.data(), *ne = nb + nl;
        float *f = reinterpret_cast<float *>(buf.data() + 2 * nl);
        for (int64_t i = 0; i < nlines; ++i) {
            nb[i] = lines[i].nbuff;
            ne[i] = lines[i].nelion;
            f[i] = lines[i].cgf;
            f[nl + i] = lines[i].elo;
            f[2 * nl + i] = lines[i].gamrf;
            f[3 * nl + i] = lines[i].gamsf;
            f[4 * nl + i] = lines[i].gamwf;
        }
        A12_HIP(hipMemcpyAsync(B->lines.p, buf.data(), 7 * 4 * nl, hipMemcpyHostToDevice, st));
        A12_HIP(hipStreamSynchronize(st));
    }
    const int32_t *d_nb = B->lines.as<int32_t>(), *d_ne = d_nb + nl;
    const float *d_f = reinterpret_cast<const float *>(d_nb + 2 * nl);
    // per-model inputs in one allocation
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    B->in_off.resize(nmodel);
    B->out_off.resize(nmodel);
    B->h_off.resize(nmodel);
    B->has_tables.resize(nmodel);
    B->cab.resize(nmodel);
    B->csc.resize(nmodel);
    size_t in_bytes = 0, out_elems = 0, h_elems = 0;
    for (int k = 0; k < nmodel; ++k) {
        const a12_synthe_atm *a = &atms[k];
        const size_t n = a->nrhox;
        B->in_off[k] = in_bytes;
        in_bytes += al(8 * a->nedge) + 3 * al(8 * (size_t)a->nf * n) + 2 * al(8 * (size_t)mw6 * n) +
                    al(8 * 10 * n);
        B->out_off[k] = out_elems;
        out_elems += (size_t)a->length * n;
        B->h_off[k] = h_elems;
        h_elems += 2 * (size_t)a->length;
    }
    A12_HIP(B->inputs.reserve(in_bytes));
    A12_HIP(B->hout.reserve(sizeof(double) * h_elems));
    A12_HIP(B->evals.reserve(2 * sizeof(long long) * nmodel));
    B->m.resize(nmodel);
    std::vector<double> depth;
    for (int k = 0; k < nmodel; ++k) {
        const a12_synthe_atm *a = &atms[k];
        const int n = a->nrhox;
        SynModel &M = B->m[k];
        M.n = n;
        M.L = a->length;
        M.N = (int)(line_hi[k] - line_lo[k]);
        M.mw6 = mw6;
        M.nedge = a->nedge;
        M.nf = a->nf;
        M.nb_shift = nb_shift[k];
        M.wbegin = a->wbegin;
        M.ratio = a->ratio;
        M.resolu = a->resolu;
        char *p = B->inputs.as<char>() + B->in_off[k];
        auto put = [&](const double *src, size_t cnt) -> const double * {
            char *q = p;
            p += al(8 * cnt);
            if (src && hipMemcpyAsync(q, src, 8 * cnt, hipMemcpyHostToDevice, st) != hipSuccess)
                return nullptr;
            return reinterpret_cast<const double *>(q);
        };
        M.wledge = put(a->wledge, a->nedge);
        M.continall = put(a->continall, (size_t)a->nf * n);
        B->cab[k] = put(a->contabs, (size_t)a->nf * n);
        B->csc[k] = put(a->contscat, (size_t)a->nf * n);
        B->has_tables[k] = a->contabs != nullptr;
        M.xnfpel = put(a->xnfpel, (size_t)mw6 * n);
        M.dopple = put(a->dopple, (size_t)mw6 * n);
        depth.assign((size_t)10 * n, 0.0);
        const double *src[10] = {a->t, a->xne, a->rho, a->xnfh, a->xnfhe1,
                                 a->xnfh2, a->hckt, a->hkt, a->hkt_atlas, a->rhox};
        for (int q = 0; q < 10; ++q) std::memcpy(&depth[(size_t)q * n], src[q], sizeof(double) * n);
        M.depth = put(depth.data(), depth.size());
        if (!M.wledge || !M.continall || !M.xnfpel || !M.dopple || !M.depth) {
            ctx->err = "a12_synthe_batch_load: upload failed";
            return A12_ERR_HIP;
        }
        A12_HIP(hipStreamSynchronize(st));  // depth is reused by the next model
        const int64_t lo = line_lo[k];
        M.l_nbuff = d_nb + lo;
        M.l_nelion = d_ne + lo;
        M.l_cgf = d_f + lo;
        M.l_elo = d_f + nl + lo;
        M.l_gr = d_f + 2 * nl + lo;
        M.l_gs = d_f + 3 * nl + lo;
        M.l_gw = d_f + 4 * nl + lo;
    }
    B->order.resize(nmodel);
    std::iota(B->order.begin(), B->order.end(), 0);
    std::stable_sort(B->order.begin(), B->order.end(), [&](int x, int y) {
        return (size_t)B->m[x].L * B->m[x].n > (size_t)B->m[y].L * B->m[y].n;
    });
    size_t fr = 0, tot = 0;
    (void)hipMemGetInfo(&fr, &tot);
    B->budget = std::min<size_t>(fr / 2, (size_t)64 << 30);
    (void)out_elems;
    return A12_OK;
}

int a12i_synthe_batch_set_budget(a12_ctx *ctx, int64_t bytes)
{
    if (!ctx->sbatch) {
        ctx->err = "a12_synthe_batch_set_budget: no batch loaded";
        return A12_ERR_STATE;
    }
    if (bytes < (int64_t)1 << 20) {
        ctx->err = "a12_synthe_batch_set_budget: budget below 1 MiB";
        return A12_ERR_ARG;
    }
    ctx->sbatch->budget = (size_t)bytes;
    ctx->sbatch->drop_graph();
    return A12_OK;
}

namespace {

// Work layout for precision f64: groups in model order, each model's
// arrays carved within its group's allocation; per-group gather
// descriptors.  Returns the allocation it needs.
int batch_plan(a12_ctx *ctx, SynBatch *B, bool f64, size_t *work_need)
{
    B->groups.clear();
    B->wk_off.assign(B->nmodel, 0);
    size_t cur = 0, need = 0;
    std::vector<int> g;
    B->max_model_work = 0;
    for (int k : B->order) {
        const SynModel &M = B->m[k];
        const size_t w = (work_bytes(M.n, M.L, M.N, M.mw6, f64, (M.L + kTile - 1) / kTile) + 255) &
                         ~(size_t)255;
        B->max_model_work = std::max(B->max_model_work, w);
        if (w > B->budget) {
            ctx->err = "a12_synthe_batch: model " + std::to_string(k) +
                       " alone needs more work memory than the budget";
            return A12_ERR_OOM;
        }
        if (!g.empty() && cur + w > B->budget) {
            B->groups.push_back(g);
            need = std::max(need, cur);
            g.clear();
            cur = 0;
        }
        B->wk_off[k] = cur;
        g.push_back(k);
        cur += w;
    }
    if (!g.empty()) {
        B->groups.push_back(g);
        need = std::max(need, cur);
    }
    *work_need = need;
    return A12_OK;
}

}  // namespace

// Launch (or capture) the prepared batch on stream st: no allocation, no
// host copies, no memset nodes.
static int batch_launch(a12_ctx *ctx, SynBatch *B, bool f64, int shape, bool transfer,
                        hipStream_t st)
{
    const size_t dstride = (sizeof(BatchItemModel) * B->nmodel + 255) & ~(size_t)255;
    int rc;
    for (size_t gi = 0; gi < B->groups.size(); ++gi) {
        const std::vector<int> &g = B->groups[gi];
        for (int k : g) {
            rc = launch_prep(ctx, B->m[k], f64, shape, st, nullptr, nullptr);
            if (rc != A12_OK) return rc;
            A12_HIP(hipMemcpyAsync(B->evals.as<long long>() + 2 * k, B->m[k].evals_out,
                                   2 * sizeof(long long), hipMemcpyDeviceToDevice, st));
        }
        auto kb = f64 ? tile_batch_kernel<true>(shape) : tile_batch_kernel<false>(shape);
        const BatchItemModel *dm =
            reinterpret_cast<const BatchItemModel *>(B->desc.as<char>() + gi * dstride);
        const long long *dio = B->item_off.as<long long>() + gi * (size_t)(B->nmodel + 1);
        hipLaunchKernelGGL(kb, dim3(B->grid), dim3(256), 0, st, dm, dio, (int)g.size());
        A12_CHECK_LAUNCH();
    }
    if (transfer)
        for (int k = 0; k < B->nmodel; ++k) {
            const SynModel &M = B->m[k];
            double *h = B->hout.as<double>() + B->h_off[k];
            rc = a12i_synthe_transfer_launch(ctx, st, f64, M.L, M.n, M.wbegin, M.ratio, M.nedge,
                                             M.wledge, M.nf, B->cab[k], B->csc[k], M.depth,
                                             M.opacity, h, h + M.L);
            if (rc != A12_OK) return rc;
        }
    return A12_OK;
}

// Prepare (carve, descriptors) for precision/shape; allocations only here.
static int batch_prepare(a12_ctx *ctx, SynBatch *B, int prec, int shape, bool transfer)
{
    const bool f64 = prec == A12_PREC_FP64;
    if (transfer) {
        if (!ctx->have_ops) {
            ctx->err = "a12_synthe_batch: the transfer needs a12_set_operators first";
            return A12_ERR_STATE;
        }
        for (int k = 0; k < B->nmodel; ++k)
            if (!B->has_tables[k]) {
                ctx->err = "a12_synthe_batch: model " + std::to_string(k) +
                           " was loaded without contabs/contscat";
                return A12_ERR_ARG;
            }
    }
    size_t need = 0;
    int rc = batch_plan(ctx, B, f64, &need);
    if (rc != A12_OK) return rc;
    const size_t rb = f64 ? sizeof(double) : sizeof(float);
    size_t out = 0;
    for (int k = 0; k < B->nmodel; ++k) out += (size_t)B->m[k].L * B->m[k].n;
    A12_HIP(B->opacity.reserve(rb * out));
    A12_HIP(B->work.reserve(need));
    const size_t dstride = (sizeof(BatchItemModel) * B->nmodel + 255) & ~(size_t)255;
    A12_HIP(B->desc.reserve(dstride * B->groups.size()));
    A12_HIP(B->item_off.reserve(sizeof(long long) * (size_t)(B->nmodel + 1) * B->groups.size()));
    for (int k = 0; k < B->nmodel; ++k) {
        SynModel &M = B->m[k];
        M.opacity = B->opacity.as<char>() + rb * B->out_off[k];
        carve_work(M, B->work.as<char>() + B->wk_off[k], f64, (M.L + kTile - 1) / kTile);
    }
    std::vector<char> dbuf(dstride * B->groups.size(), 0);
    std::vector<long long> ibuf((size_t)(B->nmodel + 1) * B->groups.size(), 0);
    for (size_t gi = 0; gi < B->groups.size(); ++gi) {
        BatchItemModel *d = reinterpret_cast<BatchItemModel *>(dbuf.data() + gi * dstride);
        long long *io = ibuf.data() + gi * (size_t)(B->nmodel + 1);
        long long run = 0;
        for (size_t q = 0; q < B->groups[gi].size(); ++q) {
            const SynModel &M = B->m[B->groups[gi][q]];
            d[q].a = tile_args(M);
            d[q].ntiles = ntiles_of(M.L, shape);
            io[q] = run;
            run += (long long)d[q].ntiles * M.n;
        }
        io[B->groups[gi].size()] = run;
    }
    A12_HIP(hipMemcpyAsync(B->desc.p, dbuf.data(), dbuf.size(), hipMemcpyHostToDevice, ctx->stream));
    A12_HIP(hipMemcpyAsync(B->item_off.p, ibuf.data(), sizeof(long long) * ibuf.size(),
                           hipMemcpyHostToDevice, ctx->stream));
    A12_HIP(hipStreamSynchronize(ctx->stream));
    int ncu = 256, per = 1;
    A12_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device));
    A12_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per, reinterpret_cast<const void *>(f64 ? tile_batch_kernel<true>(shape)
                                                 : tile_batch_kernel<false>(shape)),
        256, 0));
    B->grid = ncu * std::max(per, 1);
    B->prec = prec;
    B->shape = shape;
    B->transfer = transfer;
    return A12_OK;
}

int a12i_synthe_batch_run(a12_ctx *ctx, int prec, int transfer, int use_graph)
{
    SynBatch *B = ctx->sbatch;
    if (!B) {
        ctx->err = "a12_synthe_batch_run: call a12_synthe_batch_load first";
        return A12_ERR_STATE;
    }
    if (prec != A12_PREC_MIXED && prec != A12_PREC_FP64) {
        ctx->err = "a12_synthe_batch_run: prec must be A12_PREC_MIXED or A12_PREC_FP64";
        return A12_ERR_ARG;
    }
    const int shape = ctx->syn.shape;
    hipStream_t st = ctx->stream;
    const bool same = B->exec && B->graph_prec == prec && B->graph_shape == shape &&
                      B->graph_transfer == (transfer != 0);
    if (!(use_graph && same)) {
        B->drop_graph();
        int rc = batch_prepare(ctx, B, prec, shape, transfer != 0);
        if (rc != A12_OK) return rc;
    }
    if (!use_graph) {
        int rc = batch_launch(ctx, B, prec == A12_PREC_FP64, shape, transfer != 0, st);
        if (rc != A12_OK) return rc;
    } else {
        if (!same) {
            A12_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeGlobal));
            int rc = batch_launch(ctx, B, prec == A12_PREC_FP64, shape, transfer != 0, st);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(st, &g);
            if (rc != A12_OK) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            A12_HIP(e);
            B->graph = g;
            A12_HIP(hipGraphInstantiate(&B->exec, g, nullptr, nullptr, 0));
            B->graph_prec = prec;
            B->graph_shape = shape;
            B->graph_transfer = transfer != 0;
        }
        A12_HIP(hipGraphLaunch(B->exec, st));
    }
    A12_HIP(hipStreamSynchronize(st));
    B->ran = true;
    return A12_OK;
}

int a12i_synthe_batch_fetch(a12_ctx *ctx, int32_t model, void *opacity_out, double *hline,
                            double *hcont, int64_t *evals)
{
    SynBatch *B = ctx->sbatch;
    if (!B || !B->ran) {
        ctx->err = "a12_synthe_batch_fetch: nothing computed";
        return A12_ERR_STATE;
    }
    if (model < 0 || model >= B->nmodel) {
        ctx->err = "a12_synthe_batch_fetch: no such model";
        return A12_ERR_ARG;
    }
    if ((hline || hcont) && !B->transfer) {
        ctx->err = "a12_synthe_batch_fetch: the last run had no transfer";
        return A12_ERR_STATE;
    }
    const SynModel &M = B->m[model];
    hipStream_t st = ctx->stream;
    const size_t rb = B->prec == A12_PREC_FP64 ? sizeof(double) : sizeof(float);
    if (opacity_out)
        A12_HIP(hipMemcpyAsync(opacity_out, M.opacity, rb * (size_t)M.L * M.n,
                               hipMemcpyDeviceToHost, st));
    const double *h = B->hout.as<double>() + B->h_off[model];
    if (hline) A12_HIP(hipMemcpyAsync(hline, h, sizeof(double) * M.L, hipMemcpyDeviceToHost, st));
    if (hcont)
        A12_HIP(hipMemcpyAsync(hcont, h + M.L, sizeof(double) * M.L, hipMemcpyDeviceToHost, st));
    long long ev[2] = {0, 0};
    A12_HIP(hipMemcpyAsync(ev, B->evals.as<long long>() + 2 * model, sizeof(ev),
                           hipMemcpyDeviceToHost, st));
    A12_HIP(hipStreamSynchronize(st));
    if (evals) {
        evals[0] = ev[0];
        evals[1] = ev[1];
    }
    return A12_OK;
}

int a12i_synthe_batch_info(a12_ctx *ctx, int64_t *info)
{
    SynBatch *B = ctx->sbatch;
    if (!B) {
        ctx->err = "a12_synthe_batch_info: no batch loaded";
        return A12_ERR_STATE;
    }
    info[0] = B->nmodel;
    info[1] = (int64_t)B->groups.size();
    info[2] = (int64_t)B->work.bytes;
    info[3] = (int64_t)B->budget;
    info[4] = (int64_t)B->max_model_work;
    return A12_OK;
}
