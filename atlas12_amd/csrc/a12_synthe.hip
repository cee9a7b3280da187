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

namespace {

constexpr int kTile = 256;       // grid points per tile (one per thread)
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
__global__ void k_line_setup(int nlines, int L, int mw6, double resolu,
                             const int *__restrict__ nbuff, const int *__restrict__ nelion,
                             const float *__restrict__ cgf, const float *__restrict__ elo,
                             const float *__restrict__ gr, const float *__restrict__ gs,
                             const float *__restrict__ gw, const double *__restrict__ xnfdop,
                             const double *__restrict__ dopple, const double *__restrict__ dscal,
                             const double *__restrict__ hckt, const float *__restrict__ cont32,
                             const double *__restrict__ cont64, int *__restrict__ o_nstep,
                             int *__restrict__ o_n10, void *__restrict__ o_kap0,
                             void *__restrict__ o_adamp, void *__restrict__ o_dv,
                             void *__restrict__ o_xwing, long long *__restrict__ blk_evals)
{
#pragma clang fp contract(off)
    using R = typename std::conditional<F64, double, float>::type;
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    long ev = 0, evc = 0;
    if (l < nlines) {
        const size_t o = (size_t)j * nlines + l;
        const int nb = nbuff[l];
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
        o_nstep[o] = nstep;
        o_n10[o] = n10;
        static_cast<R *>(o_kap0)[o] = kap0;
        static_cast<R *>(o_adamp)[o] = adamp;
        static_cast<R *>(o_dv)[o] = dv;
        static_cast<R *>(o_xwing)[o] = xw;
        // the reference skips lines whose whole window is off the grid (:653)
        if (nstep >= 0 && !(nb + nstep < 1 || nb - nstep > L)) {
            ev = evals_of(nb, nstep, L);
            evc = evals_of(nb, min(nstep, n10), L);  // centre + Voigt core
        }
    }
    // block reduction of the counts in a fixed order
    __shared__ long long red[256], redc[256];
    red[threadIdx.x] = ev;
    redc[threadIdx.x] = evc;
    __syncthreads();
    for (int s = blockDim.x >> 1; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            red[threadIdx.x] += red[threadIdx.x + s];
            redc[threadIdx.x] += redc[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const size_t b = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
        const size_t nb_all = (size_t)gridDim.x * gridDim.y;
        blk_evals[b] = red[0];
        blk_evals[nb_all + b] = redc[0];
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

// Widest extent of each line over all depths, and its extent class.
__global__ void k_line_emax(int nlines, int nrhox, const int *__restrict__ nstep,
                            int *__restrict__ emax, int *__restrict__ lclass)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nlines) return;
    int e = -1;
    for (int j = 0; j < nrhox; ++j) e = max(e, nstep[(size_t)j * nlines + l]);
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

__global__ void k_cls_scatter(int nlines, const int *__restrict__ lclass,
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
        nbuffP[pos] = nbuff[l];
    }
}

// Per (tile, class): [lo, hi) of class positions whose centre lies within the
// class radius of the tile (1-based grid points t0+1 .. t1).
__global__ void k_tile_ranges(int ntiles, int L, const int *__restrict__ off, int nblk,
                              const int *__restrict__ nbuffP, int *__restrict__ rng)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles * (A12_NCLASS - 1)) return;
    const int tile = t / (A12_NCLASS - 1), c = t % (A12_NCLASS - 1);
    const int cs = off[(size_t)c * nblk], ce = off[(size_t)(c + 1) * nblk];
    const long R = class_radius(c);
    const long g0 = (long)tile * kTile + 1, g1 = min((long)(tile + 1) * kTile, (long)L);
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
// Tile gather: one workgroup per (tile of 256 points, depth), one point per
// thread.  Candidate lines of each class are staged in LDS in chunks of 256;
// lines whose window at this depth misses the tile are dropped while staging
// (wave ballot compaction, order preserved).  Each thread then adds the
// profile of every staged line that covers its point:
//   n = |i - nbuff| <= nstep:  n <= n10dop -> kappa0 H(n dv, a)
//                              else        -> x_wing / n^2
// MIXED evaluates H in FP32 and accumulates with Neumaier compensation;
// FP64 does both in double.
// --------------------------------------------------------------------------
template <bool F64>
__global__ __launch_bounds__(256) void k_tile(int L, int nlines, const int *__restrict__ rng,
                                              const int *__restrict__ perm,
                                              const int *__restrict__ nbuffP,
                                              const int *__restrict__ p_nstep,
                                              const int *__restrict__ p_n10,
                                              const void *__restrict__ p_kap0,
                                              const void *__restrict__ p_adamp,
                                              const void *__restrict__ p_dv,
                                              const void *__restrict__ p_xwing,
                                              void *__restrict__ out)
{
    using R = typename std::conditional<F64, double, float>::type;
    __shared__ int s_nb[kChunk], s_ns[kChunk], s_n10[kChunk];
    __shared__ R s_k0[kChunk], s_a[kChunk], s_dv[kChunk], s_xw[kChunk];
    __shared__ int s_wcnt[kChunk / A12_WAVE];
    __shared__ int s_nact;

    const int tile = blockIdx.x, j = blockIdx.y;
    const int tid = threadIdx.x, lane = tid & (A12_WAVE - 1), wv = tid / A12_WAVE;
    const int g0 = tile * kTile + 1;                  // first 1-based point
    const int g1 = min((tile + 1) * kTile, L);        // last
    const int gi = g0 + tid;                          // this thread's point
    const size_t dj = (size_t)j * nlines;
    const R *k0a = static_cast<const R *>(p_kap0) + dj;
    const R *aa = static_cast<const R *>(p_adamp) + dj;
    const R *dva = static_cast<const R *>(p_dv) + dj;
    const R *xwa = static_cast<const R *>(p_xwing) + dj;

    Neumaier32 acc32;
    double acc64 = 0.0;

    for (int c = 0; c < A12_NCLASS - 1; ++c) {
        const int lo = rng[2 * (tile * (A12_NCLASS - 1) + c)];
        const int hi = rng[2 * (tile * (A12_NCLASS - 1) + c) + 1];
        for (int base = lo; base < hi; base += kChunk) {
            const int q = base + tid;
            bool act = false;
            int l = 0, nb = 0, ns = -1;
            if (q < hi) {
                l = perm[q];
                nb = nbuffP[q];
                ns = p_nstep[dj + l];
                act = ns >= 0 && nb - ns <= g1 && nb + ns >= g0;
            }
            const unsigned long long m = __ballot(act);
            if (lane == 0) s_wcnt[wv] = __popcll(m);
            __syncthreads();
            if (act) {
                int pos = __popcll(m & ((1ull << lane) - 1ull));
                for (int w = 0; w < wv; ++w) pos += s_wcnt[w];
                s_nb[pos] = nb;
                s_ns[pos] = ns;
                s_n10[pos] = p_n10[dj + l];
                s_k0[pos] = k0a[l];
                s_a[pos] = aa[l];
                s_dv[pos] = dva[l];
                s_xw[pos] = xwa[l];
            }
            if (tid == 0) {
                int s = 0;
                for (int w = 0; w < kChunk / A12_WAVE; ++w) s += s_wcnt[w];
                s_nact = s;
            }
            __syncthreads();
            const int nact = s_nact;
            if (gi <= g1) {
                for (int k = 0; k < nact; ++k) {
                    const int n = abs(gi - s_nb[k]);
                    if (n > s_ns[k]) continue;
                    if (!F64) {
                        float v;
                        if (n <= s_n10[k]) v = (float)s_k0[k] * voigt_mixed((float)n * (float)s_dv[k], (float)s_a[k]);
                        else v = (float)s_xw[k] / ((float)n * (float)n);
                        acc32.add(v);
                    } else {
                        double v;
                        if (n <= s_n10[k]) v = (double)s_k0[k] * voigt_f64((double)n * (double)s_dv[k], (double)s_a[k]);
                        else v = (double)s_xw[k] / ((double)n * (double)n);
                        acc64 += v;
                    }
                }
            }
            __syncthreads();
        }
    }
    if (gi <= g1) {
        if (!F64) static_cast<float *>(out)[(size_t)j * L + gi - 1] = acc32.value();
        else static_cast<double *>(out)[(size_t)j * L + gi - 1] = acc64;
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
    S.opac_prec = -1;
    S.transferred = false;
    return A12_OK;
}

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
    const size_t nl = (size_t)std::max(N, 1);
    A12_HIP(S.xnfdop.reserve(sizeof(double) * S.mw6 * n));
    A12_HIP(S.dscal.reserve(sizeof(double) * 4 * n));
    A12_HIP(S.cont32.reserve(sizeof(float) * (size_t)L * n));
    if (f64) A12_HIP(S.cont64.reserve(sizeof(double) * (size_t)L * n));
    A12_HIP(S.opacity.reserve(rb * (size_t)L * n));
    A12_HIP(S.p_nstep.reserve(sizeof(int) * nl * n));
    A12_HIP(S.p_n10.reserve(sizeof(int) * nl * n));
    A12_HIP(S.p_kap0.reserve(rb * nl * n));
    A12_HIP(S.p_adamp.reserve(rb * nl * n));
    A12_HIP(S.p_dv.reserve(rb * nl * n));
    A12_HIP(S.p_xwing.reserve(rb * nl * n));
    A12_HIP(S.emax.reserve(sizeof(int) * nl));
    A12_HIP(S.lclass.reserve(sizeof(int) * nl));
    const int nblkl = (int)((nl + 255) / 256);
    A12_HIP(S.blk_evals.reserve(2 * sizeof(long long) * (size_t)nblkl * n));
    A12_HIP(S.evals_out.reserve(2 * sizeof(long long)));
    const int nblk = (int)((nl + kPartBlock - 1) / kPartBlock);
    A12_HIP(S.cls_cnt.reserve(sizeof(int) * (size_t)A12_NCLASS * nblk));
    A12_HIP(S.cls_off.reserve(sizeof(int) * ((size_t)A12_NCLASS * nblk + 1)));
    A12_HIP(S.perm.reserve(sizeof(int) * nl));
    A12_HIP(S.nbuffP.reserve(sizeof(int) * nl));
    const int ntiles = (L + kTile - 1) / kTile;
    A12_HIP(S.cls_range.reserve(sizeof(int) * 2 * (size_t)ntiles * (A12_NCLASS - 1)));

    const double *dep = S.depth.as<double>();
    if (ctx->prof) A12_HIP(hipEventRecord(ctx->ev[0], st));
    hipLaunchKernelGGL(k_depth_prep, dim3(n), dim3(256), 0, st, n, S.mw6, S.xnfpel.as<double>(),
                       S.dopple.as<double>(), dep, S.xnfdop.as<double>(), S.dscal.as<double>());
    A12_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_continuum, dim3((L + 255) / 256, n), dim3(256), 0, st, L, n, S.wbegin,
                       S.ratio, S.nedge, S.wledge.as<double>(), S.nf, S.continall.as<double>(),
                       S.cont32.as<float>(), f64 ? S.cont64.as<double>() : nullptr);
    A12_CHECK_LAUNCH();
    if (ctx->prof) A12_HIP(hipEventRecord(ctx->ev[1], st));
    if (N > 0) {
        auto setup = f64 ? k_line_setup<true> : k_line_setup<false>;
        hipLaunchKernelGGL(setup, dim3(nblkl, n), dim3(256), 0, st, N, L, S.mw6, S.resolu,
                           S.l_nbuff.as<int>(), S.l_nelion.as<int>(), S.l_cgf.as<float>(),
                           S.l_elo.as<float>(), S.l_gr.as<float>(), S.l_gs.as<float>(),
                           S.l_gw.as<float>(), S.xnfdop.as<double>(), S.dopple.as<double>(),
                           S.dscal.as<double>(), dep + 6 * n, S.cont32.as<float>(),
                           f64 ? S.cont64.as<double>() : nullptr, S.p_nstep.as<int>(),
                           S.p_n10.as<int>(), S.p_kap0.p, S.p_adamp.p, S.p_dv.p, S.p_xwing.p,
                           S.blk_evals.as<long long>());
        A12_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_sum_evals, dim3(1), dim3(256), 0, st, S.blk_evals.as<long long>(),
                           nblkl * n, S.evals_out.as<long long>());
        A12_CHECK_LAUNCH();
        if (ctx->prof) A12_HIP(hipEventRecord(ctx->ev[2], st));
        hipLaunchKernelGGL(k_line_emax, dim3(nblkl), dim3(256), 0, st, N, n, S.p_nstep.as<int>(),
                           S.emax.as<int>(), S.lclass.as<int>());
        A12_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_cls_count, dim3(nblk), dim3(kPartBlock), 0, st, N,
                           S.lclass.as<int>(), S.cls_cnt.as<int>());
        A12_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_cls_scan, dim3(1), dim3(1024), 0, st, A12_NCLASS * nblk,
                           S.cls_cnt.as<int>(), S.cls_off.as<int>());
        A12_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_cls_scatter, dim3(nblk), dim3(kPartBlock), 0, st, N,
                           S.lclass.as<int>(), S.cls_off.as<int>(), S.l_nbuff.as<int>(),
                           S.perm.as<int>(), S.nbuffP.as<int>());
        A12_CHECK_LAUNCH();
        const int nr = ntiles * (A12_NCLASS - 1);
        hipLaunchKernelGGL(k_tile_ranges, dim3((nr + 255) / 256), dim3(256), 0, st, ntiles, L,
                           S.cls_off.as<int>(), nblk, S.nbuffP.as<int>(), S.cls_range.as<int>());
        A12_CHECK_LAUNCH();
        if (ctx->prof) A12_HIP(hipEventRecord(ctx->ev[3], st));
        auto tk = f64 ? k_tile<true> : k_tile<false>;
        hipLaunchKernelGGL(tk, dim3(ntiles, n), dim3(256), 0, st, L, N, S.cls_range.as<int>(),
                           S.perm.as<int>(), S.nbuffP.as<int>(), S.p_nstep.as<int>(),
                           S.p_n10.as<int>(), S.p_kap0.p, S.p_adamp.p, S.p_dv.p, S.p_xwing.p,
                           S.opacity.p);
        A12_CHECK_LAUNCH();
        if (ctx->prof) A12_HIP(hipEventRecord(ctx->ev[4], st));
        S.prof_valid = ctx->prof;
    } else {
        S.prof_valid = false;
        A12_HIP(hipMemsetAsync(S.opacity.p, 0, rb * (size_t)L * n, st));
        A12_HIP(hipMemsetAsync(S.evals_out.p, 0, 2 * sizeof(long long), st));
    }
    S.opac_prec = prec;
    S.transferred = false;
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
