// a12_select.hip -- SELECTLINES' line filter on gfx950.
//
// SELECTLINES (src/atlas12_modules.f90:19171-19781) streams the raw line
// lists and keeps a record only if its centre opacity can exceed the
// continuum at some depth:
//   XNFDOPMAX(NELION, NU) = max_J XNFDOP(J, NELION) / TABCONT(J, NU)
//   CENRATIO = CEN_PREFAC * 10**(IGFLOG) * XNFDOPMAX / FREQ >= 1
//   CENRATIO * exp(-E_lo * HCKT(NRHOX)) >= 1
// with NU the continuum bin that the running wavelength walk has reached
// and FREQ the frequency of the bin before it.  Two of its paths are here:
// the predicted 16-byte records (path 1, kept as read) and the TiO MOL2
// 32-byte records (path 5, repacked into the 16-byte layout).
//
//   k_sel_xmax    XNFDOPMAX per (species, bin);
//   k_sel_pred    one thread per predicted record: bin by binary search
//                 (equal to the walk for a wavelength-sorted list), keep
//                 flag;
//   k_sel_tio     one thread per MOL2 record: TiO isotope correction,
//                 keep flag, packed record;
//   k_sel_count / k_sel_scan / k_sel_scatter: order-preserving compaction.
// TABLOG comes from a host table built with the same libm pow the
// reference's TABLOG uses.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "a12_internal.h"

namespace {

constexpr double kCLightNms = 2.99792458e17;
constexpr double kCenPrefac = 0.026538 / 1.7724538509055159;  // :19179
constexpr int kSelItems = 8;                                   // items per thread
constexpr int kSelBlock = 256;
constexpr int kSelTile = kSelItems * kSelBlock;

__global__ void k_sel_xmax(int n, int mion, int nw1, const double *__restrict__ xnfdop,
                           const double *__restrict__ tabcont, double *__restrict__ xm)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= mion * nw1) return;
    const int nel = i / nw1, nu = i - nel * nw1;
    double v = 0.0;
    for (int j = 0; j < n; ++j) v = fmax(v, xnfdop[(size_t)nel * n + j] / tabcont[(size_t)nu * n + j]);
    xm[i] = v;
}

// Number of IWAVETAB entries <= iwl: the walk's final NU (0-based).  The
// host admits only lines at or above the first bin and below the sentinel,
// so this lies in 1..nw1-1; the clamp keeps every index in bounds anyway.
__device__ __forceinline__ int sel_bin(int iwl, const int *__restrict__ iwt, int nw1)
{
    int lo = 0, hi = nw1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (iwt[mid] <= iwl) lo = mid + 1;
        else hi = mid;
    }
    return min(max(lo, 1), nw1 - 1);
}

__global__ void k_sel_pred(long long np, int mion, int nw1, const int4 *__restrict__ rec,
                           const int *__restrict__ iwt, const double *__restrict__ wavetab,
                           const double *__restrict__ xm, const double *__restrict__ tablog,
                           double hcktn, int *__restrict__ flag)
{
#pragma clang fp contract(off)
    const long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= np) return;
    const int4 r = rec[l];
    const int nu = sel_bin(r.x, iwt, nw1);
    const double freq = kCLightNms / wavetab[nu - 1];
    const int ielion = (short)(r.y & 0xffff), ielo = (short)((unsigned)r.y >> 16);
    const int igflog = (short)(r.z & 0xffff);
    const int nelion = abs(ielion / 10);
    int keep = 0;
    if (nelion >= 1 && nelion <= mion) {
        const double x = xm[(size_t)(nelion - 1) * nw1 + nu];
        if (x > 1.0e-37) {
            const double cen = kCenPrefac * tablog[igflog + 32768] * x / freq;
            if (cen >= 1.0) keep = !(cen * exp(-tablog[ielo + 32768] * hcktn) < 1.0);
        }
    }
    flag[l] = keep;
}

__global__ void k_sel_tio(long long nm, int nw1, const a12_mol2_rec *__restrict__ m,
                          const int *__restrict__ iwt, const double *__restrict__ wavetab,
                          const double *__restrict__ xm, const double *__restrict__ tablog,
                          double hcktn, int4 *__restrict__ out, int *__restrict__ flag)
{
#pragma clang fp contract(off)
    const long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nm) return;
    const a12_mol2_rec r = m[l];
    const int iso = r.iso - 45;
    const double wlv = (double)r.wl;
    int keep = 0;
    int4 o = make_int4(0, 0, 0, 0);
    if (r.code == 822 && iso >= 1 && iso <= 5 && wlv > 0.0) {
        const double ratiolg = log(1.0 + 1.0 / 2000000.0);
        const int iwl = (int)(log(wlv) / ratiolg + 0.5);
        const int nu = sel_bin(iwl, iwt, nw1);
        const double freq = kCLightNms / wavetab[nu - 1];
        const int corr = iso == 1 ? -1101 : iso == 2 ? -1138 : iso == 3 ? -131 : iso == 4 ? -1259 : -1272;
        const int kgflog = (int)round((double)r.loggf * 1000.0) + 16384;
        const int igflog = (short)max(kgflog + corr, 1);
        const double x = xm[(size_t)(895 - 1) * nw1 + nu];
        if (x != 0.0) {
            const double cen = kCenPrefac * tablog[igflog + 32768] * x / freq;
            const double elotio = fmin(fabs((double)r.e1), fabs((double)r.e2));
            if (cen >= 1.0 && !(cen * exp(-elotio * hcktn) < 1.0)) {
                keep = 1;
                const int ielion = 8949 + iso;
                const int ielo = (short)(int)(log10(fmax(elotio, 1.0e-10)) * 1000.0 + 16384.5);
                const double gr = r.gr100 == 0 ? 2.223e13 / (wlv * wlv) : pow(10.0, (double)r.gr100 * 0.01);
                const int igr = (short)(int)(log10(gr) * 1000.0 + 16384.5);
                o.x = iwl;
                o.y = (int)((unsigned)(unsigned short)ielion | ((unsigned)(unsigned short)ielo << 16));
                o.z = (int)((unsigned)(unsigned short)igflog | ((unsigned)(unsigned short)igr << 16));
                o.w = (int)(1u | (9384u << 16));
            }
        }
    }
    out[l] = o;
    flag[l] = keep;
}

// Kept records per tile of kSelTile items.
__global__ __launch_bounds__(kSelBlock) void k_sel_count(long long n, const int *__restrict__ flag,
                                                         int *__restrict__ cnt)
{
    __shared__ int red[kSelBlock / 64];
    const long long base = (long long)blockIdx.x * kSelTile + (long long)threadIdx.x * kSelItems;
    int c = 0;
    for (int k = 0; k < kSelItems; ++k)
        if (base + k < n) c += flag[base + k];
    for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < kSelBlock / 64; ++w) t += red[w];
        cnt[blockIdx.x] = t;
    }
}

// Exclusive scan of the tile counts in one block; off[nt] = total.
__global__ __launch_bounds__(1024) void k_sel_scan(int nt, const int *__restrict__ cnt,
                                                   long long *__restrict__ off)
{
    __shared__ long long part[1024];
    const int per = (nt + 1023) / 1024;
    const int t = threadIdx.x, b = t * per, e = min(nt, b + per);
    long long s = 0;
    for (int i = b; i < e; ++i) s += cnt[i];
    part[t] = s;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        const long long v = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    long long run = part[t] - s;
    for (int i = b; i < e; ++i) {
        off[i] = run;
        run += cnt[i];
    }
    if (t == 1023) off[nt] = part[1023];
}

__global__ __launch_bounds__(kSelBlock) void k_sel_scatter(long long n, const int *__restrict__ flag,
                                                           const int4 *__restrict__ cand,
                                                           const long long *__restrict__ off,
                                                           int4 *__restrict__ out)
{
    __shared__ int pre[kSelBlock];
    const long long base = (long long)blockIdx.x * kSelTile + (long long)threadIdx.x * kSelItems;
    int c = 0;
    for (int k = 0; k < kSelItems; ++k)
        if (base + k < n) c += flag[base + k];
    pre[threadIdx.x] = c;
    __syncthreads();
    for (int d = 1; d < kSelBlock; d <<= 1) {
        const int v = threadIdx.x >= d ? pre[threadIdx.x - d] : 0;
        __syncthreads();
        pre[threadIdx.x] += v;
        __syncthreads();
    }
    long long o = off[blockIdx.x] + pre[threadIdx.x] - c;
    for (int k = 0; k < kSelItems; ++k)
        if (base + k < n && flag[base + k]) out[o++] = cand[base + k];
}

template <class T>
hipError_t upload(DevBuf &b, const T *src, size_t n, hipStream_t s)
{
    hipError_t e = b.reserve(sizeof(T) * (n ? n : 1));
    if (e != hipSuccess || n == 0) return e;
    return hipMemcpyAsync(b.p, src, sizeof(T) * n, hipMemcpyHostToDevice, s);
}

}  // namespace

int a12i_atlas_select(a12_ctx *ctx, const a12_atlas_atm *a, const double *wavetab,
                      const int32_t *pred, int64_t npred, const a12_mol2_rec *mol2, int64_t nmol2,
                      int32_t *out, int64_t cap, int64_t *nkept)
{
    if (!a || !wavetab || !nkept || npred < 0 || nmol2 < 0 || (npred && !pred) ||
        (nmol2 && !mol2) || a->nrhox < 1 || a->nrhox > A12_NRMAX_LAYERS || a->mion < 1 ||
        a->ldion < a->nrhox || a->ldtab < a->nrhox || a->nwave1 < 2 || !a->tabcont ||
        !a->iwavetab || !a->xnfdop || !a->hckt || cap < 0 || (cap && !out)) {
        ctx->err = "a12_atlas_select_lines: bad dimensions or null array";
        return A12_ERR_ARG;
    }
    const int n = a->nrhox, nw1 = a->nwave1, mion = a->mion;
    const int *iwt = a->iwavetab;
    // the walk's invariants: ascending lists that start at or above the first bin
    for (int64_t l = 1; l < npred; ++l)
        if (pred[4 * l] < pred[4 * (l - 1)]) {
            ctx->err = "a12_atlas_select_lines: predicted records must ascend in IWL";
            return A12_ERR_ARG;
        }
    if (npred && (pred[0] < iwt[0] || pred[4 * (npred - 1)] >= iwt[nw1 - 1])) {
        ctx->err = "a12_atlas_select_lines: a predicted line lies outside the continuum bins";
        return A12_ERR_ARG;
    }
    const bool tio = nmol2 > 0 && a->teff <= 8000.0;  // :19557
    if (tio) {
        if (mion < 895) {
            ctx->err = "a12_atlas_select_lines: TiO needs mion >= 895";
            return A12_ERR_ARG;
        }
        for (int64_t l = 1; l < nmol2; ++l)
            if (mol2[l].wl < mol2[l - 1].wl) {
                ctx->err = "a12_atlas_select_lines: MOL2 records must ascend in wavelength";
                return A12_ERR_ARG;
            }
        const double ratiolg = std::log(1.0 + 1.0 / 2000000.0);
        for (int64_t l = 0; l < nmol2; ++l) {
            const a12_mol2_rec &r = mol2[l];
            if (r.code != 822 || r.iso < 46 || r.iso > 50 || !(r.wl > 0.0f)) continue;
            if ((int)(std::log((double)r.wl) / ratiolg + 0.5) < iwt[0]) {
                ctx->err = "a12_atlas_select_lines: a TiO line lies below the first continuum bin";
                return A12_ERR_ARG;
            }
            break;
        }
    }
    hipStream_t s = ctx->stream;
    SelectDev &S = ctx->sel;
    std::vector<double> tmp((size_t)n * std::max(mion, nw1));
    for (int i = 0; i < nw1; ++i)
        std::memcpy(&tmp[(size_t)i * n], a->tabcont + (size_t)i * a->ldtab, sizeof(double) * n);
    A12_HIP(upload(S.tabcont, tmp.data(), (size_t)nw1 * n, s));
    A12_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < mion; ++i)
        std::memcpy(&tmp[(size_t)i * n], a->xnfdop + (size_t)i * a->ldion, sizeof(double) * n);
    A12_HIP(upload(S.xnfdop, tmp.data(), (size_t)mion * n, s));
    A12_HIP(upload(S.iwt, iwt, nw1, s));
    A12_HIP(upload(S.wavetab, wavetab, nw1 - 1, s));
    if (!S.tablog_ready) {
        // TABLOG(I) = 10**((I-16384)*0.001), src/atlas12_modules.f90:2083-2086, all int16 codes
        std::vector<double> tl(65536);
        for (int i = 0; i < 65536; ++i) tl[i] = std::pow(10.0, (double)(i - 32768 - 16384) * .001);
        A12_HIP(upload(S.tablog, tl.data(), tl.size(), s));
        A12_HIP(hipStreamSynchronize(s));
        S.tablog_ready = true;
    }
    const long long ntot = npred + (tio ? nmol2 : 0);
    A12_HIP(S.xm.reserve(sizeof(double) * (size_t)mion * nw1));
    A12_HIP(S.cand.reserve(sizeof(int4) * (size_t)std::max(ntot, 1LL)));
    A12_HIP(S.flag.reserve(sizeof(int) * (size_t)std::max(ntot, 1LL)));
    const long long ntile = (ntot + kSelTile - 1) / kSelTile;
    A12_HIP(S.cnt.reserve(sizeof(int) * (size_t)std::max(ntile, 1LL)));
    A12_HIP(S.off.reserve(sizeof(long long) * (size_t)(ntile + 1)));
    if (npred) A12_HIP(hipMemcpyAsync(S.cand.p, pred, sizeof(int4) * npred, hipMemcpyHostToDevice, s));
    if (tio) A12_HIP(upload(S.mol2, mol2, nmol2, s));
    const double hcktn = a->hckt[n - 1];
    const double *tablog = S.tablog.as<double>();
    hipLaunchKernelGGL(k_sel_xmax, dim3((mion * nw1 + 255) / 256), dim3(256), 0, s, n, mion, nw1,
                       S.xnfdop.as<double>(), S.tabcont.as<double>(), S.xm.as<double>());
    A12_CHECK_LAUNCH();
    if (npred) {
        hipLaunchKernelGGL(k_sel_pred, dim3((unsigned)((npred + 255) / 256)), dim3(256), 0, s,
                           (long long)npred, mion, nw1, S.cand.as<int4>(), S.iwt.as<int>(),
                           S.wavetab.as<double>(), S.xm.as<double>(), tablog, hcktn, S.flag.as<int>());
        A12_CHECK_LAUNCH();
    }
    if (tio) {
        hipLaunchKernelGGL(k_sel_tio, dim3((unsigned)((nmol2 + 255) / 256)), dim3(256), 0, s,
                           (long long)nmol2, nw1, S.mol2.as<a12_mol2_rec>(), S.iwt.as<int>(),
                           S.wavetab.as<double>(), S.xm.as<double>(), tablog, hcktn,
                           S.cand.as<int4>() + npred, S.flag.as<int>() + npred);
        A12_CHECK_LAUNCH();
    }
    long long total = 0;
    if (ntot) {
        hipLaunchKernelGGL(k_sel_count, dim3((unsigned)ntile), dim3(kSelBlock), 0, s, ntot,
                           S.flag.as<int>(), S.cnt.as<int>());
        A12_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_sel_scan, dim3(1), dim3(1024), 0, s, (int)ntile, S.cnt.as<int>(),
                           S.off.as<long long>());
        A12_CHECK_LAUNCH();
        A12_HIP(hipMemcpyAsync(&total, S.off.as<long long>() + ntile, sizeof(long long),
                               hipMemcpyDeviceToHost, s));
        A12_HIP(hipStreamSynchronize(s));
        A12_HIP(S.kept.reserve(sizeof(int4) * (size_t)std::max(total, 1LL)));
        hipLaunchKernelGGL(k_sel_scatter, dim3((unsigned)ntile), dim3(kSelBlock), 0, s, ntot,
                           S.flag.as<int>(), S.cand.as<int4>(), S.off.as<long long>(),
                           S.kept.as<int4>());
        A12_CHECK_LAUNCH();
    }
    *nkept = total;
    if (total > cap) {
        A12_HIP(hipStreamSynchronize(s));
        ctx->err = "a12_atlas_select_lines: output capacity too small (count in *nkept)";
        return A12_ERR_ARG;
    }
    if (total)
        A12_HIP(hipMemcpyAsync(out, S.kept.p, sizeof(int4) * total, hipMemcpyDeviceToHost, s));
    A12_HIP(hipStreamSynchronize(s));
    return A12_OK;
}
