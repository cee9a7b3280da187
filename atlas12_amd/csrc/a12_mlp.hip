// a12_mlp.hip -- the starting-structure network on gfx950: 4 -> 256 -> 256
// -> 256 -> 256 -> 144 (SiLU on the hidden layers), BF16 operands on MFMA
// with FP32 accumulation, FP32 master weights and Adam.  No reference
// counterpart: it proposes the T(tau), log Pgas(tau) start that SCALE_MODEL
// otherwise provides.
//
//   k_gemm<EPI>  C = op(A) op(B) on v_mfma_f32_16x16x16_bf16: a 64 x 64
//                block tile, four waves of 2 x 2 MFMA tiles, K in steps of
//                16 staged through LDS (any strides, so transposes are free);
//                epilogues: store; + bias, keep Z, SiLU, BF16 copy (forward);
//                x SiLU'(Z) and BF16 copy (backward)
//   k_colsum     bias gradients: column sums by row slices, slices added in order
//   k_loss       dY = 2 (Y - T) / (n 144) and per-block squared error
//   k_adam       Adam on the FP32 master weights, BF16 copies refreshed
// A training step (forward, loss, backward, Adam) is captured once in a
// hipGraph and replayed; the step counter Adam needs lives on the device.
// Every reduction has a fixed order: results are deterministic.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <algorithm>
#include <cstring>
#include <vector>

#include "a12_internal.h"

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short bf16;

constexpr int kIn = 4, kInPad = 16, kHid = 256, kOut = 144, kLayers = 5;
constexpr int lay_k(int l) { return l == 0 ? kInPad : kHid; }
constexpr int lay_n(int l) { return l == kLayers - 1 ? kOut : kHid; }

// offsets of W_l (lay_k(l) x lay_n(l), row-major) and b_l in the padded flat
// parameter array
constexpr size_t w_off(int l)
{
    size_t o = 0;
    for (int i = 0; i < l; ++i) o += (size_t)lay_k(i) * lay_n(i) + lay_n(i);
    return o;
}
constexpr size_t b_off(int l) { return w_off(l) + (size_t)lay_k(l) * lay_n(l); }
constexpr size_t kParamsPadded = w_off(kLayers);

__device__ __forceinline__ bf16 f2bf(float f)
{
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (bf16)(u >> 16);
}
__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ __forceinline__ float dsilu(float z)
{
    const float s = 1.0f / (1.0f + expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

enum { EPI_STORE = 0, EPI_FWD = 1, EPI_BWD = 2 };

// C (M x N, ldc) = A (M x K; element (m, k) at A[m sam + k sak]) times
// B (K x N; element (k, n) at B[k sbk + n sbn]).  Operand tiles are staged
// in LDS (zero outside the matrices); lane l of an MFMA holds A row l % 16,
// k = 4 (l / 16) .. +3; B column l % 16, same k; result column l % 16,
// rows 4 (l / 16) .. +3.
// Split K (gridDim.z > 1, EPI_STORE only): block z sums k in
// [z kchunk, (z+1) kchunk) into its own slice C + z M ldc.
template <int EPI>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const bf16 *__restrict__ A,
                                              long sam, long sak, const bf16 *__restrict__ B,
                                              long sbk, long sbn, float *__restrict__ C, int ldc,
                                              bf16 *__restrict__ Cb, const float *__restrict__ bias,
                                              float *__restrict__ Z, const float *__restrict__ Zp,
                                              int act, int kchunk)
{
    __shared__ bf16 sA[64 * 17];  // [m][k], padded row
    __shared__ bf16 sB[16 * 65];  // [k][n], padded row
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    const int kbeg = blockIdx.z * kchunk, kend = min(K, kbeg + kchunk);
    C += (long)blockIdx.z * M * ldc;
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    const int li = lane & 15, kb = 4 * (lane >> 4);
    f32x4 acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = kbeg; k0 < kend; k0 += 16) {
        for (int e = t; e < 1024; e += 256) {
            const int r = e >> 4, c = e & 15, gm = m0 + r, gk = k0 + c;
            sA[r * 17 + c] = (gm < M && gk < kend) ? A[(long)gm * sam + (long)gk * sak] : (bf16)0;
        }
        for (int e = t; e < 1024; e += 256) {
            const int r = e >> 6, c = e & 63, gk = k0 + r, gn = n0 + c;
            sB[r * 65 + c] = (gk < kend && gn < N) ? B[(long)gk * sbk + (long)gn * sbn] : (bf16)0;
        }
        __syncthreads();
        for (int i = 0; i < 2; ++i) {
            bf16x4 a;
            for (int q = 0; q < 4; ++q) a[q] = (short)sA[(wm + i * 16 + li) * 17 + kb + q];
            for (int j = 0; j < 2; ++j) {
                bf16x4 b;
                for (int q = 0; q < 4; ++q) b[q] = (short)sB[(kb + q) * 65 + wn + j * 16 + li];
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[i][j], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int q = 0; q < 4; ++q) {
                const int gm = m0 + wm + i * 16 + kb + q, gn = n0 + wn + j * 16 + li;
                if (gm >= M || gn >= N) continue;
                const long o = (long)gm * ldc + gn;
                float v = acc[i][j][q];
                if (EPI == EPI_FWD) {
                    v += bias[gn];
                    Z[o] = v;
                    if (act) v = silu(v);
                }
                if (EPI == EPI_BWD) v *= dsilu(Zp[o]);
                C[o] = v;
                if (Cb) Cb[o] = f2bf(v);
            }
}

// out[i] = sum over the split slices in slice order
__global__ void k_sum_split(long n, int split, const float *__restrict__ part,
                            float *__restrict__ out)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int z = 0; z < split; ++z) s += part[(long)z * n + i];
    out[i] = s;
}

// inputs (n x 4, FP32) -> padded BF16 (rows x 16), zero beyond
__global__ void k_prep_x(int n, int rows, const float *__restrict__ x, bf16 *__restrict__ xb)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows * kInPad) return;
    const int r = i / kInPad, c = i % kInPad;
    xb[i] = (r < n && c < kIn) ? f2bf(x[r * kIn + c]) : (bf16)0;
}

// dY = 2 (Y - T) / (n 144) on the n valid rows (0 on padding rows), its
// BF16 copy, and per-block partial sums of (Y - T)^2 in a fixed tree order
__global__ __launch_bounds__(256) void k_loss(int n, int rows, const float *__restrict__ y,
                                              const float *__restrict__ tgt, float *__restrict__ dy,
                                              bf16 *__restrict__ dyb, float *__restrict__ part)
{
    __shared__ float red[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    float e2 = 0.0f;
    if (i < rows * kOut) {
        float g = 0.0f;
        if (i / kOut < n) {
            const float d = y[i] - tgt[i];
            e2 = d * d;
            g = 2.0f * d / ((float)n * (float)kOut);
        }
        dy[i] = g;
        dyb[i] = f2bf(g);
    }
    red[threadIdx.x] = e2;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// loss = (sum of the nb block partials, in a fixed order) / (n 144): each
// thread sums a strided share, then a tree in LDS
__global__ __launch_bounds__(256) void k_loss_sum(int nb, int n, const float *__restrict__ part,
                                                  float *__restrict__ loss)
{
    __shared__ float red[256];
    float s = 0.0f;
    for (int b = threadIdx.x; b < nb; b += 256) s += part[b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss = red[0] / ((float)n * (float)kOut);
}

// Column sums of g (M x N, row stride ldg) over row slice blockIdx.y of
// `chunk` rows: block = 64 columns x 4 row groups, each group summing every
// 4th row of the slice, the groups combined in order; part[y N + c].
// k_sum_split then adds the slices in slice order: a fixed order throughout.
__global__ __launch_bounds__(256) void k_colsum(int M, int N, const float *__restrict__ g,
                                                int ldg, int chunk, float *__restrict__ part)
{
    __shared__ float red[4][64];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int c = blockIdx.x * 64 + tx;
    const int r0 = blockIdx.y * chunk, r1 = min(M, r0 + chunk);
    float s = 0.0f;
    if (c < N)
        for (int r = r0 + ty; r < r1; r += 4) s += g[(long)r * ldg + c];
    red[ty][tx] = s;
    __syncthreads();
    if (ty == 0 && c < N)
        part[(long)blockIdx.y * N + c] = ((red[0][tx] + red[1][tx]) + red[2][tx]) + red[3][tx];
}

__global__ void k_step(int *__restrict__ t)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *t += 1;
}

// Adam (beta1 0.9, beta2 0.999, eps 1e-8), bias-corrected by the device step
__global__ void k_adam(size_t np, const int *__restrict__ step, const float *__restrict__ lrp,
                       float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, bf16 *__restrict__ wb)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const float b1 = 0.9f, b2 = 0.999f;
    const int t = *step;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.0f - powf(b1, (float)t)), vh = vi / (1.0f - powf(b2, (float)t));
    const float wi = w[i] - *lrp * mh / (sqrtf(vh) + 1.0e-8f);
    w[i] = wi;
    wb[i] = f2bf(wi);
}

__global__ void k_to_bf16(size_t n, const float *__restrict__ a, bf16 *__restrict__ b)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = f2bf(a[i]);
}

// deterministic host RNG for the initial weights
uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct MlpDev {
    int rows = 0;  // padded batch capacity (multiple of 64)
    int n = 0;     // samples of the loaded data
    DevBuf w, wb, g, m, v;                    // parameters (padded layout)
    DevBuf xb, tgt, y, z[kLayers], h, hb[kLayers - 1];
    DevBuf dz, dzb, dz2, dz2b, part, loss, step, lr, xin, xbi, split;
    hipGraphExec_t graph = nullptr;
    bool ready = false;
    ~MlpDev()
    {
        if (graph) (void)hipGraphExecDestroy(graph);
        DevBuf *b[] = {&w, &wb, &g, &m, &v, &xb, &tgt, &y, &h, &dz, &dzb, &dz2, &dz2b,
                       &part, &loss, &step, &lr, &xin, &xbi, &split};
        for (DevBuf *p : b) p->release();
        for (DevBuf &p : z) p.release();
        for (DevBuf &p : hb) p.release();
    }
};

static MlpDev &mlp(a12_ctx *ctx)
{
    if (!ctx->mlp) ctx->mlp = new MlpDev();
    return *ctx->mlp;
}

void a12i_mlp_free(a12_ctx *ctx)
{
    delete ctx->mlp;
    ctx->mlp = nullptr;
}

static void gemm(hipStream_t s, int epi, int M, int N, int K, const bf16 *A, long sam, long sak,
                 const bf16 *B, long sbk, long sbn, float *C, int ldc, bf16 *Cb,
                 const float *bias, float *Z, const float *Zp, int act)
{
    const dim3 grid((N + 63) / 64, (M + 63) / 64);
    if (epi == EPI_FWD)
        hipLaunchKernelGGL(k_gemm<EPI_FWD>, grid, dim3(256), 0, s, M, N, K, A, sam, sak, B, sbk, sbn,
                           C, ldc, Cb, bias, Z, Zp, act, K);
    else if (epi == EPI_BWD)
        hipLaunchKernelGGL(k_gemm<EPI_BWD>, grid, dim3(256), 0, s, M, N, K, A, sam, sak, B, sbk, sbn,
                           C, ldc, Cb, bias, Z, Zp, act, K);
    else
        hipLaunchKernelGGL(k_gemm<EPI_STORE>, grid, dim3(256), 0, s, M, N, K, A, sam, sak, B, sbk,
                           sbn, C, ldc, Cb, bias, Z, Zp, act, K);
}

// C (M x N, dense) = A B with K split `split` ways into slices of `part`,
// summed in slice order (deterministic).
static void gemm_splitk(hipStream_t s, int M, int N, int K, const bf16 *A, long sam, long sak,
                        const bf16 *B, long sbk, long sbn, float *C, float *part, int split)
{
    const int kchunk = ((K + split - 1) / split + 15) / 16 * 16;
    const dim3 grid((N + 63) / 64, (M + 63) / 64, split);
    hipLaunchKernelGGL(k_gemm<EPI_STORE>, grid, dim3(256), 0, s, M, N, K, A, sam, sak, B, sbk, sbn,
                       part, N, nullptr, nullptr, nullptr, nullptr, 0, kchunk);
    const long n = (long)M * N;
    hipLaunchKernelGGL(k_sum_split, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, split,
                       part, C);
}

// forward over `rows` rows of the BF16 inputs xb into y (and z, hb)
static void mlp_forward(a12_ctx *ctx, MlpDev &D, int rows, const bf16 *xb)
{
    hipStream_t s = ctx->stream;
    const bf16 *wb = D.wb.as<bf16>();
    const float *w = D.w.as<float>();
    const bf16 *hin = xb;
    for (int l = 0; l < kLayers; ++l) {
        const int K = lay_k(l), N = lay_n(l);
        const bool last = l == kLayers - 1;
        float *out = last ? D.y.as<float>() : D.h.as<float>();
        bf16 *outb = last ? nullptr : D.hb[l].as<bf16>();
        gemm(s, EPI_FWD, rows, N, K, hin, K, 1, wb + w_off(l), N, 1, out, N, outb, w + b_off(l),
             D.z[l].as<float>(), nullptr, last ? 0 : 1);
        hin = outb;
    }
}

// row-splits of the weight-gradient GEMMs: about 128 rows per slice, <= 32
static int split_of(int rows) { return std::max(1, std::min(32, rows / 128)); }

// loss, backward and (if adam) the update, over D.n valid of D.rows rows
static void mlp_backward(a12_ctx *ctx, MlpDev &D, bool adam)
{
    hipStream_t s = ctx->stream;
    const int rows = D.rows;
    const int nb = (rows * kOut + 255) / 256;
    hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, s, D.n, rows, D.y.as<float>(),
                       D.tgt.as<float>(), D.dz.as<float>(), D.dzb.as<bf16>(), D.part.as<float>());
    hipLaunchKernelGGL(k_loss_sum, dim3(1), dim3(256), 0, s, nb, D.n, D.part.as<float>(),
                       D.loss.as<float>());
    const bf16 *wb = D.wb.as<bf16>();
    float *g = D.g.as<float>();
    float *dz = D.dz.as<float>(), *dz2 = D.dz2.as<float>();
    bf16 *dzb = D.dzb.as<bf16>(), *dz2b = D.dz2b.as<bf16>();
    for (int l = kLayers - 1; l >= 0; --l) {
        const int K = lay_k(l), N = lay_n(l);
        const bf16 *hin = l == 0 ? D.xb.as<bf16>() : D.hb[l - 1].as<bf16>();
        // dW_l = Hin^T dZ_l  (K x N, summed over the rows, K split over the rows)
        gemm_splitk(s, K, N, rows, hin, 1, K, dzb, N, 1, g + w_off(l), D.split.as<float>(),
                    split_of(rows));
        // bias gradient: column sums of dZ_l by row slices (reusing the split buffer)
        const int sp = split_of(rows), chunk = (rows + sp - 1) / sp;
        hipLaunchKernelGGL(k_colsum, dim3((N + 63) / 64, sp), dim3(256), 0, s, rows, N, dz, N,
                           chunk, D.split.as<float>());
        hipLaunchKernelGGL(k_sum_split, dim3((N + 255) / 256), dim3(256), 0, s, (long)N, sp,
                           D.split.as<float>(), g + b_off(l));
        if (l > 0) {
            // dZ_{l-1} = (dZ_l W_l^T) * SiLU'(Z_{l-1})
            gemm(s, EPI_BWD, rows, K, N, dzb, N, 1, wb + w_off(l), 1, N, dz2, K, dz2b, nullptr,
                 nullptr, D.z[l - 1].as<float>(), 0);
            std::swap(dz, dz2);
            std::swap(dzb, dz2b);
        }
    }
    if (adam) {
        hipLaunchKernelGGL(k_step, dim3(1), dim3(64), 0, s, D.step.as<int>());
        hipLaunchKernelGGL(k_adam, dim3((unsigned)((kParamsPadded + 255) / 256)), dim3(256), 0, s,
                           kParamsPadded, D.step.as<int>(), D.lr.as<float>(), D.w.as<float>(), g,
                           D.m.as<float>(), D.v.as<float>(), D.wb.as<bf16>());
    }
}

int a12i_mlp_init(a12_ctx *ctx, int batch, uint64_t seed)
{
    if (batch < 1 || batch > (1 << 20)) {
        ctx->err = "a12_mlp_init: batch out of range";
        return A12_ERR_ARG;
    }
    MlpDev &D = mlp(ctx);
    if (D.graph) {
        (void)hipGraphExecDestroy(D.graph);
        D.graph = nullptr;
    }
    hipStream_t s = ctx->stream;
    D.rows = (batch + 63) / 64 * 64;
    D.n = 0;
    const size_t P = kParamsPadded, R = D.rows;
    A12_HIP(D.w.reserve(4 * P));
    A12_HIP(D.g.reserve(4 * P));
    A12_HIP(D.m.reserve(4 * P));
    A12_HIP(D.v.reserve(4 * P));
    A12_HIP(D.wb.reserve(2 * P));
    A12_HIP(D.xb.reserve(2 * R * kInPad));
    A12_HIP(D.xbi.reserve(2 * R * kInPad));
    A12_HIP(D.split.reserve(4 * (size_t)split_of(D.rows) * kHid * kHid));
    A12_HIP(D.xin.reserve(4 * R * kIn));
    A12_HIP(D.tgt.reserve(4 * R * kOut));
    A12_HIP(D.y.reserve(4 * R * kOut));
    A12_HIP(D.h.reserve(4 * R * kHid));
    for (int l = 0; l < kLayers; ++l) A12_HIP(D.z[l].reserve(4 * R * lay_n(l)));
    for (int l = 0; l < kLayers - 1; ++l) A12_HIP(D.hb[l].reserve(2 * R * kHid));
    A12_HIP(D.dz.reserve(4 * R * kHid));
    A12_HIP(D.dz2.reserve(4 * R * kHid));
    A12_HIP(D.dzb.reserve(2 * R * kHid));
    A12_HIP(D.dz2b.reserve(2 * R * kHid));
    A12_HIP(D.part.reserve(4 * ((R * kOut + 255) / 256)));
    A12_HIP(D.loss.reserve(4));
    A12_HIP(D.step.reserve(4));
    A12_HIP(D.lr.reserve(4));
    // Glorot-uniform weights from a seeded generator, zero biases; the
    // padding rows of W_1 (inputs 5..16) stay zero
    std::vector<float> w(P, 0.0f);
    uint64_t st = seed;
    for (int l = 0; l < kLayers; ++l) {
        const int fin = l == 0 ? kIn : lay_k(l), N = lay_n(l);
        const double lim = std::sqrt(6.0 / (fin + N));
        for (int k = 0; k < fin; ++k)
            for (int c = 0; c < N; ++c) {
                const double u = (double)(splitmix64(st) >> 11) * (1.0 / 9007199254740992.0);
                w[w_off(l) + (size_t)k * N + c] = (float)((2.0 * u - 1.0) * lim);
            }
    }
    A12_HIP(hipMemcpyAsync(D.w.p, w.data(), 4 * P, hipMemcpyHostToDevice, s));
    A12_HIP(hipMemsetAsync(D.m.p, 0, 4 * P, s));
    A12_HIP(hipMemsetAsync(D.v.p, 0, 4 * P, s));
    A12_HIP(hipMemsetAsync(D.step.p, 0, 4, s));
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, s, P,
                       D.w.as<float>(), D.wb.as<bf16>());
    A12_CHECK_LAUNCH();
    A12_HIP(hipStreamSynchronize(s));
    D.ready = true;
    return A12_OK;
}

static int need_mlp(a12_ctx *ctx, const char *who)
{
    if (!ctx->mlp || !ctx->mlp->ready) {
        ctx->err = std::string(who) + ": call a12_mlp_init first";
        return A12_ERR_STATE;
    }
    return A12_OK;
}

// logical order: W1 (4 x 256), b1, W2 (256 x 256), b2, ..., W5 (256 x 144), b5
static void logical_to_padded(const float *src, std::vector<float> &dst)
{
    dst.assign(kParamsPadded, 0.0f);
    size_t o = 0;
    for (int l = 0; l < kLayers; ++l) {
        const int fin = l == 0 ? kIn : lay_k(l), N = lay_n(l);
        std::memcpy(&dst[w_off(l)], src + o, sizeof(float) * fin * N);
        o += (size_t)fin * N;
        std::memcpy(&dst[b_off(l)], src + o, sizeof(float) * N);
        o += N;
    }
}

static void padded_to_logical(const std::vector<float> &src, float *dst)
{
    size_t o = 0;
    for (int l = 0; l < kLayers; ++l) {
        const int fin = l == 0 ? kIn : lay_k(l), N = lay_n(l);
        std::memcpy(dst + o, &src[w_off(l)], sizeof(float) * fin * N);
        o += (size_t)fin * N;
        std::memcpy(dst + o, &src[b_off(l)], sizeof(float) * N);
        o += N;
    }
}

int a12i_mlp_params(a12_ctx *ctx, float *get, const float *set)
{
    if (int rc = need_mlp(ctx, "a12_mlp_params")) return rc;
    MlpDev &D = *ctx->mlp;
    hipStream_t s = ctx->stream;
    std::vector<float> buf;
    if (set) {
        logical_to_padded(set, buf);
        A12_HIP(hipMemcpyAsync(D.w.p, buf.data(), 4 * kParamsPadded, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((kParamsPadded + 255) / 256)), dim3(256), 0,
                           s, kParamsPadded, D.w.as<float>(), D.wb.as<bf16>());
        A12_CHECK_LAUNCH();
        A12_HIP(hipMemsetAsync(D.m.p, 0, 4 * kParamsPadded, s));
        A12_HIP(hipMemsetAsync(D.v.p, 0, 4 * kParamsPadded, s));
        A12_HIP(hipMemsetAsync(D.step.p, 0, 4, s));
    }
    if (get) {
        buf.resize(kParamsPadded);
        A12_HIP(hipMemcpyAsync(buf.data(), D.w.p, 4 * kParamsPadded, hipMemcpyDeviceToHost, s));
        A12_HIP(hipStreamSynchronize(s));
        padded_to_logical(buf, get);
    }
    A12_HIP(hipStreamSynchronize(s));
    return A12_OK;
}

int a12i_mlp_set_data(a12_ctx *ctx, const float *x, const float *t, int n)
{
    if (int rc = need_mlp(ctx, "a12_mlp_set_data")) return rc;
    MlpDev &D = *ctx->mlp;
    if (n < 1 || n > D.rows || !x || !t) {
        ctx->err = "a12_mlp_set_data: need 1 <= n <= the batch of a12_mlp_init";
        return A12_ERR_ARG;
    }
    hipStream_t s = ctx->stream;
    A12_HIP(hipMemcpyAsync(D.xin.p, x, sizeof(float) * kIn * n, hipMemcpyHostToDevice, s));
    A12_HIP(hipMemsetAsync(D.tgt.p, 0, sizeof(float) * kOut * D.rows, s));
    A12_HIP(hipMemcpyAsync(D.tgt.p, t, sizeof(float) * kOut * n, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_prep_x, dim3((D.rows * kInPad + 255) / 256), dim3(256), 0, s, n, D.rows,
                       D.xin.as<float>(), D.xb.as<bf16>());
    A12_CHECK_LAUNCH();
    A12_HIP(hipStreamSynchronize(s));
    D.n = n;
    if (D.graph) {
        (void)hipGraphExecDestroy(D.graph);
        D.graph = nullptr;
    }
    return A12_OK;
}

int a12i_mlp_grads(a12_ctx *ctx, float *loss, float *grads)
{
    if (int rc = need_mlp(ctx, "a12_mlp_grads")) return rc;
    MlpDev &D = *ctx->mlp;
    if (D.n < 1) {
        ctx->err = "a12_mlp_grads: call a12_mlp_set_data first";
        return A12_ERR_STATE;
    }
    mlp_forward(ctx, D, D.rows, D.xb.as<bf16>());
    mlp_backward(ctx, D, false);
    A12_CHECK_LAUNCH();
    std::vector<float> g(kParamsPadded);
    A12_HIP(hipMemcpyAsync(g.data(), D.g.p, 4 * kParamsPadded, hipMemcpyDeviceToHost, ctx->stream));
    float l = 0.0f;
    A12_HIP(hipMemcpyAsync(&l, D.loss.p, 4, hipMemcpyDeviceToHost, ctx->stream));
    A12_HIP(hipStreamSynchronize(ctx->stream));
    if (grads) padded_to_logical(g, grads);
    if (loss) *loss = l;
    return A12_OK;
}

int a12i_mlp_train(a12_ctx *ctx, int steps, float lr, float *loss)
{
    if (int rc = need_mlp(ctx, "a12_mlp_train")) return rc;
    MlpDev &D = *ctx->mlp;
    if (D.n < 1 || steps < 0 || !(lr > 0.0f)) {
        ctx->err = "a12_mlp_train: set data first; steps >= 0, lr > 0";
        return A12_ERR_ARG;
    }
    hipStream_t s = ctx->stream;
    A12_HIP(hipMemcpyAsync(D.lr.p, &lr, 4, hipMemcpyHostToDevice, s));
    if (!D.graph) {
        // capture one step: forward, loss, backward, Adam (kernels only)
        hipGraph_t g = nullptr;
        A12_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        mlp_forward(ctx, D, D.rows, D.xb.as<bf16>());
        mlp_backward(ctx, D, true);
        hipError_t e = hipStreamEndCapture(s, &g);
        if (e != hipSuccess) {
            ctx->err = std::string("a12_mlp_train: capture: ") + hipGetErrorString(e);
            return A12_ERR_HIP;
        }
        e = hipGraphInstantiate(&D.graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            D.graph = nullptr;
            ctx->err = std::string("a12_mlp_train: instantiate: ") + hipGetErrorString(e);
            return A12_ERR_HIP;
        }
    }
    for (int i = 0; i < steps; ++i) A12_HIP(hipGraphLaunch(D.graph, s));
    float l = 0.0f;
    A12_HIP(hipMemcpyAsync(&l, D.loss.p, 4, hipMemcpyDeviceToHost, s));
    A12_HIP(hipStreamSynchronize(s));
    if (loss) *loss = l;
    return A12_OK;
}

int a12i_mlp_infer(a12_ctx *ctx, const float *x, int n, float *y)
{
    if (int rc = need_mlp(ctx, "a12_mlp_infer")) return rc;
    MlpDev &D = *ctx->mlp;
    if (n < 0 || (n && (!x || !y))) {
        ctx->err = "a12_mlp_infer: bad arguments";
        return A12_ERR_ARG;
    }
    hipStream_t s = ctx->stream;
    // inference has its own input buffer: the training set stays loaded
    for (int i0 = 0; i0 < n; i0 += D.rows) {
        const int c = std::min(D.rows, n - i0);
        A12_HIP(hipMemcpyAsync(D.xin.p, x + (size_t)kIn * i0, sizeof(float) * kIn * c,
                               hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_prep_x, dim3((D.rows * kInPad + 255) / 256), dim3(256), 0, s, c,
                           D.rows, D.xin.as<float>(), D.xbi.as<bf16>());
        mlp_forward(ctx, D, (c + 63) / 64 * 64, D.xbi.as<bf16>());
        A12_CHECK_LAUNCH();
        A12_HIP(hipMemcpyAsync(y + (size_t)kOut * i0, D.y.p, sizeof(float) * kOut * c,
                               hipMemcpyDeviceToHost, s));
    }
    A12_HIP(hipStreamSynchronize(s));
    return A12_OK;
}

int a12i_mlp_gemm_check(a12_ctx *ctx, int M, int N, int K, const uint16_t *a, const uint16_t *b,
                        float *c, int split)
{
    if (M < 1 || N < 1 || K < 1 || M > 4096 || N > 4096 || K > 4096 || !a || !b || !c ||
        split < 1 || split > 64) {
        ctx->err = "a12_mlp_gemm: bad arguments";
        return A12_ERR_ARG;
    }
    DevBuf da, db, dc;
    hipStream_t s = ctx->stream;
    A12_HIP(da.reserve(2 * (size_t)M * K));
    A12_HIP(db.reserve(2 * (size_t)K * N));
    A12_HIP(dc.reserve(4 * (size_t)M * N));
    A12_HIP(hipMemcpyAsync(da.p, a, 2 * (size_t)M * K, hipMemcpyHostToDevice, s));
    A12_HIP(hipMemcpyAsync(db.p, b, 2 * (size_t)K * N, hipMemcpyHostToDevice, s));
    DevBuf dp;
    if (split == 1) {
        gemm(s, EPI_STORE, M, N, K, da.as<bf16>(), K, 1, db.as<bf16>(), N, 1, dc.as<float>(), N,
             nullptr, nullptr, nullptr, nullptr, 0);
    } else {
        A12_HIP(dp.reserve(4 * (size_t)split * M * N));
        gemm_splitk(s, M, N, K, da.as<bf16>(), K, 1, db.as<bf16>(), N, 1, dc.as<float>(),
                    dp.as<float>(), split);
    }
    A12_CHECK_LAUNCH();
    A12_HIP(hipMemcpyAsync(c, dc.p, 4 * (size_t)M * N, hipMemcpyDeviceToHost, s));
    A12_HIP(hipStreamSynchronize(s));
    da.release();
    db.release();
    dc.release();
    dp.release();
    return A12_OK;
}
