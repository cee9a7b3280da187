// a12_select.hip -- SELECTLINES' line filter on gfx950.
//
// SELECTLINES (src/atlas12_modules.f90:19171-19781) streams the raw line
// lists and keeps a record only if its centre opacity can exceed the
// continuum at some depth:
//   XNFDOPMAX(NELION, NU) = max_J XNFDOP(J, NELION) / TABCONT(J, NU)
//   CENRATIO = CEN_PREFAC * 10**(IGFLOG) * XNFDOPMAX / FREQ >= 1
//   CENRATIO * exp(-E_lo * HCKT(NRHOX)) >= 1
// with NU the continuum bin that the running wavelength walk has reached
// and FREQ the frequency of the bin before it.  Its seven lists, in store
// order: (1) predicted and (2) observed low lines, (3) hilines and (7) H3+
// (16-byte records; 1-3 kept as read, 7 repacked), (4) the ASCII diatomics
// (diatomic_record_t, sorted by wavelength key), (5) TiO MOL2 (32-byte) and
// (6) H2O (8-byte), all repacked into the 16-byte layout.
//
//   k_sel_xmax    XNFDOPMAX per (species, bin);
//   k_sel_rec16   one thread per 16-byte record: bin by binary search
//                 (equal to the walk for a wavelength-sorted list; path 2
//                 starts where path 1's walk ended), keep flag;
//   k_sel_dia     one thread per diatomic record in key order: MOLCODE ->
//                 ISOX, integer encodings, keep flag, packed record;
//   k_sel_tio     one thread per MOL2 record: TiO isotope correction, keep
//                 flag, packed record;
//   k_sel_h2o     one thread per H2O record: isotope from the signs,
//                 radiative width of the walk's bin, packed record;
//   k_sel_count / k_sel_scan / k_sel_scatter: order-preserving compaction.
// TABLOG and the H2O widths per bin come from host tables built with the
// same libm the reference uses; the diatomic wavelength keys and their
// heapsort (whose order among equal keys the store inherits) run on the
// host as well.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <string>
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

__device__ __forceinline__ int pack2(int lo, int hi)
{
    return (int)((unsigned)(unsigned short)lo | ((unsigned)(unsigned short)hi << 16));
}

// Float to default INTEGER as x86-64's truncating conversion gives it (the
// reference's INT(LOG10(0.0) ...) for an H2O line with E_lo = 0).
__device__ __forceinline__ int f2i_trunc(double x)
{
    if (!(x > -2147483649.0 && x < 2147483648.0)) return (int)0x80000000u;
    return (int)x;
}

enum SelPath { kPathPred = 1, kPathLowObs = 2, kPathHiLines = 3, kPathH3p = 7 };

// Paths 1, 2, 3 and 7 (:19262-19395, :19723-19767).  nu_floor: where the
// walk starts (path 2 goes on from path 1's last bin).  A NELION outside
// 1..mion is skipped by path 1 and counted in *bad by paths 2 and 3, which
// index without a bounds check.
__global__ void k_sel_rec16(long long np, int path, int nu_floor, int mion, int nw1,
                            int4 *__restrict__ rec, const int *__restrict__ iwt,
                            const double *__restrict__ wavetab, const double *__restrict__ xm,
                            const double *__restrict__ tablog, double hcktn,
                            int *__restrict__ flag, int *__restrict__ bad)
{
#pragma clang fp contract(off)
    const long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= np) return;
    int4 r = rec[l];
    const int nu = max(nu_floor, sel_bin(r.x, iwt, nw1));
    const double freq = kCLightNms / wavetab[nu - 1];
    const int ielion = (short)(r.y & 0xffff), ielo = (short)((unsigned)r.y >> 16);
    int igflog = (short)(r.z & 0xffff);
    int nelion = abs(ielion / 10);
    if (path == kPathH3p) {
        igflog = (short)max(igflog - 1272, 1);
        nelion = 895;
    }
    int keep = 0;
    if (nelion < 1 || nelion > mion) {
        if (path != kPathPred) atomicAdd(bad, 1);
    } else {
        const double x = xm[(size_t)(nelion - 1) * nw1 + nu];
        const bool skip = path <= kPathLowObs ? x <= 1.0e-37 : x == 0.0;
        if (!skip) {
            const double cen = kCenPrefac * tablog[igflog + 32768] * x / freq;
            if (!(cen < 1.0)) keep = !(cen * exp(-tablog[ielo + 32768] * hcktn) < 1.0);
        }
    }
    if (keep && path == kPathH3p) {
        r.z = pack2(igflog, (short)((unsigned)r.z >> 16));
        r.w = pack2(1, 9384);
        rec[l] = r;
    }
    flag[l] = keep;
}

// SELECTLINES' MOLCODES and ISOX (:19181-19200)
__constant__ int kMolcodes[46] = {
    8410, 8411, 8460, 8461, 8470, 8471, 8480, 8481, 8482, 8510, 8511, 8512, 8530, 8531, 8532,
    8580, 8581, 8582, 8583, 8584, 8620, 8621, 8622, 8623, 8640, 8641, 8642, 8643, 8680, 8681,
    8682, 8690, 8691, 8692, 8693, 8700, 8701, 8702, 8703, 8704, 8705, 8890, 8891, 8892, 8896,
    8960};
__constant__ int kIsox[46] = {
    0, -4469, -5, -1955, -2, -2444, -1, -3398, -2690, -105, -996, -947, -35, -1331, -1516,
    -13, -2189, -2870, -1681, -4398, -1362, -77, -1022, -1626, -1237, -38, -1658, -2553, -10,
    -1960, -3910, -7, -1957, -2449, -4399, -6, -1956, -3403, -5353, -2695, -4645, -36, -1332,
    -1517, -2725, -2};

// Path 4 (:19417-19536) on the records in key order: position l holds
// record order[l] with wavelength key keys[l].  The host has checked every
// MOLCODE against the table and mion >= MOLCODE / 10.
__global__ void k_sel_dia(long long n, int nw1, const a12_diatomic_rec *__restrict__ dia,
                          const int *__restrict__ keys, const int *__restrict__ order,
                          const int *__restrict__ iwt, const double *__restrict__ wavetab,
                          const double *__restrict__ xm, const double *__restrict__ tablog,
                          double hcktn, int4 *__restrict__ out, int *__restrict__ flag)
{
#pragma clang fp contract(off)
    const long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    const a12_diatomic_rec d = dia[order[l]];
    const int iwl = keys[l];
    const int nu = sel_bin(iwl, iwt, nw1);
    const double freq = kCLightNms / wavetab[nu - 1];
    int imol = 0;
    for (int k = 0; k < 46; ++k)
        if (kMolcodes[k] == d.molcode) imol = k;
    const int isox = kIsox[imol];
    const double gflog = d.gflog_dex - (double)isox * 1.0e-3;
    const int ielion = (short)d.molcode;
    const int ielo = (short)(int)llround(log10(fmax(d.elo_cm, 1.0e-10)) * 1000.0 + 16384.0);
    const int igf0 = (short)(int)llround(gflog * 1000.0 + 16384.0);
    const int igr = (short)(int)llround(d.gammar_log * 1000.0 + 16384.0);
    const int igw = (short)(int)llround(d.gammaw_log * 1000.0 + 16384.0);
    const int igflog = (short)max(igf0 + isox, 1);
    const int nelion = abs(ielion / 10);
    const double x = xm[(size_t)(nelion - 1) * nw1 + nu];
    int keep = 0;
    if (!(x == 0.0)) {
        const double cen = kCenPrefac * tablog[igflog + 32768] * x / freq;
        if (!(cen < 1.0)) keep = !(cen * exp(-tablog[ielo + 32768] * hcktn) < 1.0);
    }
    out[l] = make_int4(iwl, pack2(ielion, ielo), pack2(igflog, igr), pack2(1, igw));
    flag[l] = keep;
}

// Path 6 (:19650-19720).  igr[i]: the radiative-width code the walk sets
// on stepping past bin i (a host table).  The host has checked mion >= 940.
__global__ void k_sel_h2o(long long n, int nw1, const a12_h2o_rec *__restrict__ h,
                          const int *__restrict__ iwt, const double *__restrict__ wavetab,
                          const short *__restrict__ igr, const double *__restrict__ xm,
                          const double *__restrict__ tablog, double hcktn,
                          int4 *__restrict__ out, int *__restrict__ flag)
{
#pragma clang fp contract(off)
    const long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    const a12_h2o_rec r = h[l];
    const int nu = sel_bin(r.iwl, iwt, nw1);
    const double freq = kCLightNms / wavetab[nu - 1];
    const int iso = r.ielo > 0 ? (r.igflog > 0 ? 1 : 2) : (r.igflog > 0 ? 3 : 4);
    const int corr = iso == 1 ? -1 : iso == 2 ? -3398 : iso == 3 ? -2690 : -5000;
    const double elo = (double)(short)abs((int)r.ielo);           // ABS of an INTEGER(2)
    const int kgflog = (short)abs((int)r.igflog);
    const int igflog = (short)max(kgflog + corr, 1);
    const double x = xm[(size_t)(940 - 1) * nw1 + nu];
    int keep = 0;
    if (!(x == 0.0)) {
        const double cen = kCenPrefac * tablog[igflog + 32768] * x / freq;
        if (!(cen < 1.0)) keep = !(cen * exp(-elo * hcktn) < 1.0);
    }
    const int ielo = (short)f2i_trunc(log10(elo) * 1000.0 + 16384.5);
    out[l] = make_int4(r.iwl, pack2(-(9399 + iso), ielo), pack2(igflog, igr[nu - 1]),
                       pack2(1, 9384));
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

// sort_indirect_by_int (src/atlas12_modules.f90:1985-2039): the heapsort
// the diatomic store's order among equal wavelength keys comes from.  A
// stable or parallel sort orders ties differently, so the library runs this
// one as written (1-based heap indices).
void heapsort_keys(std::vector<int> &keys, std::vector<int> &order)
{
    const long long n = (long long)keys.size();
    if (n <= 1) return;
    auto K = [&](long long i) -> int & { return keys[(size_t)(i - 1)]; };
    auto swap_at = [&](long long a, long long b) {
        std::swap(keys[(size_t)(a - 1)], keys[(size_t)(b - 1)]);
        std::swap(order[(size_t)(a - 1)], order[(size_t)(b - 1)]);
    };
    auto sift = [&](long long parent, long long last) {
        for (;;) {
            long long child = 2 * parent;
            if (child > last) break;
            if (child + 1 <= last && K(child + 1) > K(child)) child = child + 1;
            if (K(parent) >= K(child)) break;
            swap_at(parent, child);
            parent = child;
        }
    };
    for (long long i = n / 2; i >= 1; --i) sift(i, n);
    for (long long i = n; i >= 2; --i) {
        swap_at(1, i);
        sift(1, i - 1);
    }
}

const int kHostMolcodes[46] = {
    8410, 8411, 8460, 8461, 8470, 8471, 8480, 8481, 8482, 8510, 8511, 8512, 8530, 8531, 8532,
    8580, 8581, 8582, 8583, 8584, 8620, 8621, 8622, 8623, 8640, 8641, 8642, 8643, 8680, 8681,
    8682, 8690, 8691, 8692, 8693, 8700, 8701, 8702, 8703, 8704, 8705, 8890, 8891, 8892, 8896,
    8960};

// The walk's final NU (0-based) for iwl, as sel_bin computes it.
int host_bin(const int *iwt, int nw1, int iwl)
{
    const int lo = (int)(std::upper_bound(iwt, iwt + nw1, iwl) - iwt);
    return std::min(std::max(lo, 1), nw1 - 1);
}

// A 16-byte list (or the H2O list, stride 2 int32): ascending IWL within
// the continuum bins, which makes the binary search equal to the walk.
const char *check_iwl(const int32_t *rec, int64_t n, int stride, const int *iwt, int nw1)
{
    for (int64_t l = 1; l < n; ++l)
        if (rec[stride * l] < rec[stride * (l - 1)]) return "records must ascend in IWL";
    if (n && (rec[0] < iwt[0] || rec[stride * (n - 1)] >= iwt[nw1 - 1]))
        return "a line lies outside the continuum bins";
    return nullptr;
}

}  // namespace

int a12i_atlas_select_all(a12_ctx *ctx, const a12_atlas_atm *a, const double *wavetab,
                          const a12_sel_lists *L, int32_t *out, int64_t cap, int64_t *nkept)
{
    auto fail = [&](const char *msg) {
        ctx->err = msg;
        return A12_ERR_ARG;
    };
    if (!a || !wavetab || !nkept || !L || a->nrhox < 1 || a->nrhox > A12_NRMAX_LAYERS ||
        a->mion < 1 || a->ldion < a->nrhox || a->ldtab < a->nrhox || a->nwave1 < 2 ||
        !a->tabcont || !a->iwavetab || !a->xnfdop || !a->hckt || cap < 0 || (cap && !out))
        return fail("a12_atlas_select: bad dimensions or null array");
    const int64_t counts[7] = {L->npred, L->nlowobs, L->nhilines, L->ndia, L->nmol2, L->nh2o,
                               L->nh3p};
    const void *ptrs[7] = {L->pred, L->lowobs, L->hilines, L->dia, L->mol2, L->h2o, L->h3p};
    for (int k = 0; k < 7; ++k)
        if (counts[k] < 0 || (counts[k] && !ptrs[k]))
            return fail("a12_atlas_select: negative count or null list");
    const int n = a->nrhox, nw1 = a->nwave1, mion = a->mion;
    const int *iwt = a->iwavetab;
    const bool cool = a->teff <= 8000.0;   // TiO and H2O only at or below 8000 K (:19557, :19650)
    const int64_t nmol2 = cool ? L->nmol2 : 0, nh2o = cool ? L->nh2o : 0;
    // the walk's invariants, per list
    const struct { const int32_t *r; int64_t n; int stride; const char *msg; } lists16[5] = {
        {L->pred, L->npred, 4, "a12_atlas_select: predicted list: "},
        {L->lowobs, L->nlowobs, 4, "a12_atlas_select: observed low-line list: "},
        {L->hilines, L->nhilines, 4, "a12_atlas_select: hilines list: "},
        {L->h3p, L->nh3p, 4, "a12_atlas_select: H3+ list: "},
        {reinterpret_cast<const int32_t *>(L->h2o), nh2o, 2, "a12_atlas_select: H2O list: "}};
    for (const auto &li : lists16)
        if (const char *e = check_iwl(li.r, li.n, li.stride, iwt, nw1)) {
            ctx->err = std::string(li.msg) + e;
            return A12_ERR_ARG;
        }
    if ((L->nh3p && mion < 895) || (nmol2 && mion < 895) || (nh2o && mion < 940))
        return fail("a12_atlas_select: TiO and H3+ need mion >= 895, H2O mion >= 940");
    if (nmol2) {
        const a12_mol2_rec *mol2 = L->mol2;
        for (int64_t l = 1; l < nmol2; ++l)
            if (mol2[l].wl < mol2[l - 1].wl)
                return fail("a12_atlas_select: MOL2 records must ascend in wavelength");
        const double ratiolg = std::log(1.0 + 1.0 / 2000000.0);
        for (int64_t l = 0; l < nmol2; ++l) {
            const a12_mol2_rec &r = mol2[l];
            if (r.code != 822 || r.iso < 46 || r.iso > 50 || !(r.wl > 0.0f)) continue;
            if ((int)(std::log((double)r.wl) / ratiolg + 0.5) < iwt[0])
                return fail("a12_atlas_select: a TiO line lies below the first continuum bin");
            break;
        }
    }
    // path 4: wavelength keys NINT(LOG(wl)/RATIOLG) and their heapsort (:19440-19446)
    const int64_t ndia = L->ndia;
    std::vector<int> dkeys((size_t)ndia), dorder((size_t)ndia);
    if (ndia) {
        const double ratiolg = std::log(1.0 + 1.0 / 2000000.0);
        int maxcode = 0;
        for (int64_t i = 0; i < ndia; ++i) {
            const a12_diatomic_rec &d = L->dia[i];
            if (!(d.wlvac_nm > 0.0))
                return fail("a12_atlas_select: a diatomic wavelength is not positive");
            const int mc = d.molcode;
            if (std::find(kHostMolcodes, kHostMolcodes + 46, mc) == kHostMolcodes + 46)
                return fail("a12_atlas_select: unknown diatomic MOLCODE");
            maxcode = std::max(maxcode, mc);
            dkeys[(size_t)i] = (int)std::lround(std::log(d.wlvac_nm) / ratiolg);
            dorder[(size_t)i] = (int)i;
        }
        if (mion < maxcode / 10) return fail("a12_atlas_select: mion below a diatomic NELION");
        heapsort_keys(dkeys, dorder);
        if (dkeys.front() < iwt[0] || dkeys.back() >= iwt[nw1 - 1])
            return fail("a12_atlas_select: a diatomic line lies outside the continuum bins");
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
    // store segments: pred | lowobs | hilines | dia | TiO | H2O | H3+
    int64_t seg[8] = {0};
    const int64_t used[7] = {L->npred, L->nlowobs, L->nhilines, ndia, nmol2, nh2o, L->nh3p};
    for (int k = 0; k < 7; ++k) seg[k + 1] = seg[k] + used[k];
    const long long ntot = seg[7];
    A12_HIP(S.xm.reserve(sizeof(double) * (size_t)mion * nw1));
    A12_HIP(S.cand.reserve(sizeof(int4) * (size_t)std::max(ntot, 1LL)));
    A12_HIP(S.flag.reserve(sizeof(int) * (size_t)std::max(ntot, 1LL)));
    const long long ntile = (ntot + kSelTile - 1) / kSelTile;
    A12_HIP(S.cnt.reserve(sizeof(int) * (size_t)std::max(ntile, 1LL)));
    A12_HIP(S.off.reserve(sizeof(long long) * (size_t)(ntile + 2)));
    int4 *cand = S.cand.as<int4>();
    int *flag = S.flag.as<int>();
    for (int k : {0, 1, 2, 6})
        if (used[k])
            A12_HIP(hipMemcpyAsync(cand + seg[k], ptrs[k], sizeof(int4) * used[k],
                                   hipMemcpyHostToDevice, s));
    if (nmol2) A12_HIP(upload(S.mol2, L->mol2, nmol2, s));
    if (ndia) {
        A12_HIP(upload(S.dia, L->dia, ndia, s));
        A12_HIP(upload(S.dkey, dkeys.data(), ndia, s));
        A12_HIP(upload(S.dord, dorder.data(), ndia, s));
    }
    if (nh2o) {
        A12_HIP(upload(S.h2o, L->h2o, nh2o, s));
        // the walk's GAMMAR = 2.474D-22 * FREQ**2 * 0.001 (a REAL(4) literal)
        // per bin stepped past, as INT(LOG10 * 1000 + 16384.5) (:19666-19671)
        std::vector<short> igr((size_t)nw1);
        for (int i = 0; i + 1 < nw1; ++i) {
            const double freq = 2.99792458e17 / wavetab[i];
            const double g = std::log10(2.474e-22 * (freq * freq) * (double)0.001f) * 1000.0 + 16384.5;
            igr[(size_t)i] = (short)(g > -2147483649.0 && g < 2147483648.0 ? (int)g : INT32_MIN);
        }
        A12_HIP(upload(S.igr, igr.data(), igr.size(), s));
    }
    // the path-2 walk starts where path 1's ended (:19316-19326)
    const int nu_after_pred = L->npred ? host_bin(iwt, nw1, L->pred[4 * (L->npred - 1)]) : 0;
    const double hcktn = a->hckt[n - 1];
    const double *tablog = S.tablog.as<double>();
    int *bad = reinterpret_cast<int *>(S.off.as<long long>() + ntile + 1);
    A12_HIP(hipMemsetAsync(bad, 0, sizeof(long long), s));
    hipLaunchKernelGGL(k_sel_xmax, dim3((mion * nw1 + 255) / 256), dim3(256), 0, s, n, mion, nw1,
                       S.xnfdop.as<double>(), S.tabcont.as<double>(), S.xm.as<double>());
    A12_CHECK_LAUNCH();
    auto grid = [](int64_t m) { return dim3((unsigned)((m + 255) / 256)); };
    const struct { int k, path, floor; } rec16[4] = {
        {0, kPathPred, 0}, {1, kPathLowObs, nu_after_pred}, {2, kPathHiLines, 0}, {6, kPathH3p, 0}};
    for (const auto &p : rec16) {
        if (!used[p.k]) continue;
        hipLaunchKernelGGL(k_sel_rec16, grid(used[p.k]), dim3(256), 0, s, (long long)used[p.k],
                           p.path, p.floor, mion, nw1, cand + seg[p.k], S.iwt.as<int>(),
                           S.wavetab.as<double>(), S.xm.as<double>(), tablog, hcktn,
                           flag + seg[p.k], bad);
        A12_CHECK_LAUNCH();
    }
    if (ndia) {
        hipLaunchKernelGGL(k_sel_dia, grid(ndia), dim3(256), 0, s, (long long)ndia, nw1,
                           S.dia.as<a12_diatomic_rec>(), S.dkey.as<int>(), S.dord.as<int>(),
                           S.iwt.as<int>(), S.wavetab.as<double>(), S.xm.as<double>(), tablog,
                           hcktn, cand + seg[3], flag + seg[3]);
        A12_CHECK_LAUNCH();
    }
    if (nmol2) {
        hipLaunchKernelGGL(k_sel_tio, grid(nmol2), dim3(256), 0, s, (long long)nmol2, nw1,
                           S.mol2.as<a12_mol2_rec>(), S.iwt.as<int>(), S.wavetab.as<double>(),
                           S.xm.as<double>(), tablog, hcktn, cand + seg[4], flag + seg[4]);
        A12_CHECK_LAUNCH();
    }
    if (nh2o) {
        hipLaunchKernelGGL(k_sel_h2o, grid(nh2o), dim3(256), 0, s, (long long)nh2o, nw1,
                           S.h2o.as<a12_h2o_rec>(), S.iwt.as<int>(), S.wavetab.as<double>(),
                           S.igr.as<short>(), S.xm.as<double>(), tablog, hcktn, cand + seg[5],
                           flag + seg[5]);
        A12_CHECK_LAUNCH();
    }
    long long head[2] = {0, 0};   // total kept, bad-NELION count
    if (ntot) {
        hipLaunchKernelGGL(k_sel_count, dim3((unsigned)ntile), dim3(kSelBlock), 0, s, ntot, flag,
                           S.cnt.as<int>());
        A12_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_sel_scan, dim3(1), dim3(1024), 0, s, (int)ntile, S.cnt.as<int>(),
                           S.off.as<long long>());
        A12_CHECK_LAUNCH();
    }
    A12_HIP(hipMemcpyAsync(head, S.off.as<long long>() + ntile, 2 * sizeof(long long),
                           hipMemcpyDeviceToHost, s));
    A12_HIP(hipStreamSynchronize(s));
    const long long total = ntot ? head[0] : 0;
    if ((int)head[1] != 0)
        return fail("a12_atlas_select: a NELION outside 1..mion in the observed-low-line or "
                    "hilines list");
    *nkept = total;
    if (total > cap) return fail("a12_atlas_select: output capacity too small (count in *nkept)");
    if (total) {
        A12_HIP(S.kept.reserve(sizeof(int4) * (size_t)total));
        hipLaunchKernelGGL(k_sel_scatter, dim3((unsigned)ntile), dim3(kSelBlock), 0, s, ntot, flag,
                           cand, S.off.as<long long>(), S.kept.as<int4>());
        A12_CHECK_LAUNCH();
        A12_HIP(hipMemcpyAsync(out, S.kept.p, sizeof(int4) * total, hipMemcpyDeviceToHost, s));
    }
    A12_HIP(hipStreamSynchronize(s));
    return A12_OK;
}

int a12i_atlas_select(a12_ctx *ctx, const a12_atlas_atm *a, const double *wavetab,
                      const int32_t *pred, int64_t npred, const a12_mol2_rec *mol2, int64_t nmol2,
                      int32_t *out, int64_t cap, int64_t *nkept)
{
    a12_sel_lists L;
    std::memset(&L, 0, sizeof L);
    L.pred = pred;
    L.npred = npred;
    L.mol2 = mol2;
    L.nmol2 = nmol2;
    return a12i_atlas_select_all(ctx, a, wavetab, &L, out, cap, nkept);
}
