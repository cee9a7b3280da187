// a12_wave.h -- one-wavefront FP64 building blocks shared by the solvers
// that keep a whole system in LDS (JOSH, the molecular-equilibrium Newton).
#pragma once

#include <hip/hip_runtime.h>

namespace {

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

__device__ __forceinline__ void wsync() { __syncthreads(); }

// SOLVIT(A, N, B, IPIVOT), src/atlas12_modules.f90:5739-5847, on one
// wavefront: A (n x n, column-major, leading dimension n) and B (n) in LDS,
// ipiv (n) LDS scratch, n <= 64.  Lane K - 1 owns row K.  The pivot is the
// first row with the largest |A(K, I)| (the reference's strict '>' scan);
// rows swap in columns I+1..N only; the reciprocal pivot is stored on the
// diagonal and a pivot below 1e-30 in magnitude is floored to +-1e-30.
// Every element update is the reference's single expression, so results
// agree with the serial routine bit for bit.
__device__ void solvit_lds(double *A, double *B, int *ipiv, int n)
{
#pragma clang fp contract(off)
    const int ln = lane_id();
    const double TOL = 1.0e-30;
#define AM(i, j) A[((j) - 1) * n + ((i) - 1)]
    for (int I = 1; I <= n - 1; ++I) {
        // (a) pivot: first row K >= I with the largest |A(K, I)|
        const int K = ln + 1;
        double v = (K >= I && K <= n) ? fabs(AM(K, I)) : -1.0;
        int idx = (K >= I && K <= n) ? K : 0x7fffffff;
        for (int o = 32; o > 0; o >>= 1) {
            const double ov = __shfl_xor(v, o);
            const int oi = __shfl_xor(idx, o);
            if (ov > v || (ov == v && oi < idx)) {
                v = ov;
                idx = oi;
            }
        }
        const int M = idx;
        // (b) swap rows I and M in columns I+1..n
        if (M != I && K > I && K <= n) {
            const double t = AM(I, K);
            AM(I, K) = AM(M, K);
            AM(M, K) = t;
        }
        wsync();
        // (c) reciprocal pivot on the diagonal
        if (ln == 0) {
            double p = AM(M, I);
            if (fabs(p) < TOL) p = copysign(TOL, p);
            const double piv = 1.0 / p;
            AM(M, I) = AM(I, I);
            AM(I, I) = piv;
            ipiv[I] = M;
        }
        wsync();
        // (d) multipliers, (e) elimination; lane K > I owns row K
        const double piv = AM(I, I);
        double mk = 0.0;
        if (K > I && K <= n) {
            mk = AM(K, I) * piv;
            AM(K, I) = mk;
        }
        for (int J = I + 1; J <= n; ++J) {
            const double c = AM(I, J);
            if (c == 0.0) continue;
            if (K > I && K <= n) AM(K, J) = AM(K, J) - mk * c;
        }
        wsync();
    }
    if (ln == 0) {
        double p = AM(n, n);
        if (fabs(p) < TOL) p = copysign(TOL, p);
        AM(n, n) = 1.0 / p;
    }
    wsync();
    // forward substitution
    const int K = ln + 1;
    for (int I = 1; I <= n - 1; ++I) {
        if (ln == 0) {
            const int M = ipiv[I];
            if (M != I) {
                const double t = B[M - 1];
                B[M - 1] = B[I - 1];
                B[I - 1] = t;
            }
        }
        wsync();
        const double c = B[I - 1];
        if (K > I && K <= n) B[K - 1] = B[K - 1] - AM(K, I) * c;
        wsync();
    }
    // back substitution
    for (int J = n; J >= 2; --J) {
        if (ln == 0) B[J - 1] = B[J - 1] * AM(J, J);
        wsync();
        const double c = B[J - 1];
        if (K < J) B[K - 1] = B[K - 1] - AM(K, J) * c;
        wsync();
    }
    if (ln == 0) B[0] = B[0] * AM(1, 1);
    wsync();
#undef AM
}

}  // namespace
