// a12_common.h -- shared device helpers for the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define A12_WAVE 64
#define A12_NXTAU 51
#define A12_NRMAX 128

// Physical constants, mod_constants (src/atlas12_modules.f90:315-355).
#define A12_SQRTPI 1.7724538509055159
#define A12_CLIGHT_NMS 2.99792458e17
#define A12_HPLANCK 6.62607015e-27
#define A12_CLIGHT 2.99792458e10

// --------------------------------------------------------------------------
// Voigt function H(v, a) = Re w(v + i a): Weideman (1994) N = 32 rational
// approximation, Humlicek (1982) asymptote for |v| + a > 15 -- the method
// of src/synthe_module.f90:463-537.  The 32 Weideman coefficients (the
// FFT-derived constants of that method, L = sqrt(N / sqrt 2)) are the
// published ones and must be identical.
// --------------------------------------------------------------------------
__constant__ static const double kWeiA64[32] = {
    -1.30317978630500875e-12, 3.74088129316536249e-12, 8.03036789996388945e-12,
    -2.15436320778387687e-11, -5.54423594816646238e-11, 1.16582510935237737e-10,
    4.15374309183345315e-10, -5.23102048119632885e-10, -3.20801509172336887e-09,
    8.12488945684665156e-10, 2.37975567798974168e-08, 2.29304390650999664e-08,
    -1.48130789151209774e-07, -4.18407637021697758e-07, 4.25583313757500854e-07,
    4.40153173157854990e-06, 6.82103194400198485e-06, -2.14096192017107501e-05,
    -1.30754492546153456e-04, -2.45329802700214317e-04, 3.92591360700703107e-04,
    4.51954110534921738e-03, 1.90061557848454077e-02, 5.73044035298372195e-02,
    1.40607162268937685e-01, 2.95444510715087316e-01, 5.46013972063934094e-01,
    9.01925489364799882e-01, 1.34554416923454490e+00, 1.82566962963248147e+00,
    2.26353729990026764e+00, 2.57225340812456960e+00};
#define A12_WEI_L 4.75682846001088411

// FP64 evaluation, operation for operation as the reference routine (no
// contraction), so that FP32-rounded profile values and the cutoff decisions
// taken on them agree with the reference.
__device__ __forceinline__ double voigt_f64(double v, double a)
{
#pragma clang fp contract(off)
    const double x = fabs(v), y = a, s = x + y;
    if (s > 15.0) {
        const double dr = x * x - y * y - 0.5;
        const double di = 2.0 * x * y;
        const double d2 = dr * dr + di * di;
        return (-y * dr + x * di) / (A12_SQRTPI * d2);
    }
    const double dr = A12_WEI_L + y, di = -x;
    const double nr = A12_WEI_L - y, ni = x;
    const double d2 = dr * dr + di * di;
    const double zr = (nr * dr + ni * di) / d2;
    const double zi = (ni * dr - nr * di) / d2;
    double pr = kWeiA64[0], pim = 0.0;
#pragma unroll
    for (int k = 1; k < 32; ++k) {
        const double t = pr * zr - pim * zi + kWeiA64[k];
        pim = pr * zi + pim * zr;
        pr = t;
    }
    const double sqr = dr * dr - di * di;
    const double sqi = 2.0 * dr * di;
    const double sqd = sqr * sqr + sqi * sqi;
    return 2.0 * (pr * sqr + pim * sqi) / sqd + dr / (A12_SQRTPI * d2);
}

// Reference-typed evaluation: FP32 argument and result, FP64 inside.
__device__ __forceinline__ float voigt_ref(float v, float a)
{
    return (float)voigt_f64((double)fabsf(v), (double)a);
}

// FP32 evaluation for the mixed-precision path: the same rational function
// with FP32 coefficients; the two divisions become one reciprocal each.
__constant__ static const float kWeiA32[32] = {
    -1.30317978630500875e-12f, 3.74088129316536249e-12f, 8.03036789996388945e-12f,
    -2.15436320778387687e-11f, -5.54423594816646238e-11f, 1.16582510935237737e-10f,
    4.15374309183345315e-10f, -5.23102048119632885e-10f, -3.20801509172336887e-09f,
    8.12488945684665156e-10f, 2.37975567798974168e-08f, 2.29304390650999664e-08f,
    -1.48130789151209774e-07f, -4.18407637021697758e-07f, 4.25583313757500854e-07f,
    4.40153173157854990e-06f, 6.82103194400198485e-06f, -2.14096192017107501e-05f,
    -1.30754492546153456e-04f, -2.45329802700214317e-04f, 3.92591360700703107e-04f,
    4.51954110534921738e-03f, 1.90061557848454077e-02f, 5.73044035298372195e-02f,
    1.40607162268937685e-01f, 2.95444510715087316e-01f, 5.46013972063934094e-01f,
    9.01925489364799882e-01f, 1.34554416923454490e+00f, 1.82566962963248147e+00f,
    2.26353729990026764e+00f, 2.57225340812456960e+00f};

__device__ __forceinline__ float voigt_f32(float v, float a)
{
#pragma clang fp contract(off)
    const float x = fabsf(v), y = a, s = x + y;
    if (s > 15.0f) {
        const float dr = x * x - y * y - 0.5f;
        const float di = 2.0f * x * y;
        const float d2 = dr * dr + di * di;
        return (-y * dr + x * di) / ((float)A12_SQRTPI * d2);
    }
    const float L = (float)A12_WEI_L;
    const float dr = L + y, di = -x;
    const float nr = L - y, ni = x;
    const float d2 = dr * dr + di * di;
    const float inv = 1.0f / d2;
    const float zr = (nr * dr + ni * di) * inv;
    const float zi = (ni * dr - nr * di) * inv;
    float pr = kWeiA32[0], pim = 0.0f;
#pragma unroll
    for (int k = 1; k < 32; ++k) {
        const float t = fmaf(pr, zr, fmaf(-pim, zi, kWeiA32[k]));
        pim = fmaf(pr, zi, pim * zr);
        pr = t;
    }
    const float sqr = dr * dr - di * di;
    const float sqi = 2.0f * dr * di;
    const float sqd = sqr * sqr + sqi * sqi;
    return 2.0f * (pr * sqr + pim * sqi) / sqd + dr * inv * (float)(1.0 / A12_SQRTPI);
}

// Mixed-precision H(v, a) for the FP32 path.  The Weideman form computes
// w(z) and takes its real part; when a is small and |v| large, Re w is tiny
// against |w| and FP32 loses ~eps |w| / Re w.  So:
//   |v| + a > 15           : the reference's Humlicek asymptote (FP32; its
//                            real part has no cancellation written as below)
//   |v| >= 5               : Laplace continued fraction of w, 12 levels,
//                            whose imaginary parts only ever add (Re w is
//                            formed without cancellation), plus the
//                            Re e^{-z^2} term the truncated fraction lacks
//                            near the real axis (a < 1)
//   2.5 <= |v| < 5, a < 0.1: FP64 Weideman
//   otherwise              : FP32 Weideman
// The asymptote and the continued fraction divide with the hardware
// reciprocal (1 ulp).  Worst relative error against Re w(z) is below 2e-5
// (see tests).
__device__ __forceinline__ float voigt_mixed(float v, float a)
{
    // fused steps are written as fmaf; nothing else is contracted, so the
    // value is the same in every kernel this is inlined into
#pragma clang fp contract(off)
    const float x = fabsf(v), y = a;
    if (x + y > 15.0f) {
        const float x2 = x * x, y2 = y * y;
        const float dr = x2 - y2 - 0.5f;
        const float di = 2.0f * x * y;
        return y * (x2 + y2 + 0.5f) *
               __builtin_amdgcn_rcpf((float)A12_SQRTPI * (dr * dr + di * di));
    }
    if (x >= 5.0f) {
        float tr = x, ti = y;
#pragma unroll
        for (int k = 12; k >= 1; --k) {
            const float q = (0.5f * (float)k) * __builtin_amdgcn_rcpf(fmaf(tr, tr, ti * ti));
            tr = fmaf(-q, tr, x);
            ti = fmaf(q, ti, y);
        }
        float h = ti * __builtin_amdgcn_rcpf((float)A12_SQRTPI * fmaf(tr, tr, ti * ti));
        if (y < 1.0f) h += __expf(y * y - x * x) * __cosf(2.0f * x * y);
        return h;
    }
    if (x >= 2.5f && y < 0.1f) return (float)voigt_f64((double)x, (double)y);
    return voigt_f32(x, y);
}

// Integer power by repeated squaring -- how the compiled reference evaluates
// REAL(8)**INTEGER with a run-time exponent (a constant exponent is expanded
// into left-to-right products; the two agree up to the cube).
// (wave = wbegin * ratio**(i-1), src/synthe.f90:495).
__device__ __forceinline__ double powi_f64(double x, int m)
{
    unsigned n = m < 0 ? -m : m;
    double y = (n & 1u) ? x : 1.0;
    while (n >>= 1) {
        x = x * x;
        if (n & 1u) y = y * x;
    }
    return m < 0 ? 1.0 / y : y;
}

// Neumaier (improved Kahan-Babuska) compensated accumulation.
struct Neumaier32 {
    float s = 0.0f, c = 0.0f;
    __device__ __forceinline__ void add(float x)
    {
        const float t = s + x;
        if (fabsf(s) >= fabsf(x)) c += (s - t) + x;
        else c += (x - t) + s;
        s = t;
    }
    __device__ __forceinline__ float value() const { return s + c; }
};
