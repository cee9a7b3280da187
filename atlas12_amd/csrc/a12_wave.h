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

// Wavefront reductions of the device library (DPP, no LDS).
extern "C" __device__ double __ockl_wfred_max_f64(double);
extern "C" __device__ int __ockl_wfred_min_i32(int);

__device__ __forceinline__ double readlane_f64(double v, int lane)
{
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                            __builtin_amdgcn_readlane(__double2loint(v), lane));
}

// SOLVIT as solvit_lds, with the matrix in registers: lane r - 1 starts
// with row r (A(r, 1..N) in a[0..N-1]) and B(r) in b; lanes >= N hold
// nothing.  Rows never move between lanes: a lane-to-row map follows the
// reference's row swaps.  Column I stays with the lanes when rows I and M
// swap, which is what the reference's swap of columns I+1..N only leaves
// (the old A(I, I) becomes the new A(M, I), the pivot lane's entry becomes
// the reciprocal pivot).  The forward substitution's step I uses only step
// I's swap and multipliers, so it runs inside the elimination, with the same
// operands in the same order.  Every element update is the reference's
// single expression; the result is the serial routine's, bit for bit.
// Returns X(r) in the lane that ends up holding row r, and r in *row (0 in
// lanes >= N).
template <int N>
__device__ double solvit_reg(double (&a)[N], double b, int *row)
{
#pragma clang fp contract(off)
    static_assert(N >= 2 && N <= 64, "one wavefront");
    const int ln = lane_id();
    const double TOL = 1.0e-30;
    int myrow = ln < N ? ln + 1 : 0;
    double col = a[0];  // this lane's entry in the current pivot column
    for (int I = 1; I <= N - 1; ++I) {
        // (a) pivot: the first row K >= I with the largest |A(K, I)| (a DPP
        // maximum, then the rows that hold it; more than one is rare)
        const bool cand = myrow >= I;
        const double vmax = __ockl_wfred_max_f64(cand ? fabs(col) : -1.0);
        const unsigned long long hit = __ballot(cand && fabs(col) == vmax);
        int M;
        if (hit == 0ull) M = I;  // no comparable candidate (NaN column): no swap
        else if ((hit & (hit - 1ull)) == 0ull) M = __builtin_amdgcn_readlane(myrow, (int)__builtin_ctzll(hit));
        else M = __ockl_wfred_min_i32(((hit >> ln) & 1ull) ? myrow : 0x7fffffff);
        const int pM = (int)__builtin_ctzll(__ballot(myrow == M));
        const int pI = (int)__builtin_ctzll(__ballot(myrow == I));
        // (c) reciprocal pivot
        double p = readlane_f64(col, pM);
        if (fabs(p) < TOL) p = copysign(TOL, p);
        const double piv = 1.0 / p;
        // (b) rows I and M swap: the map only
        if (ln == pM) myrow = I;
        else if (ln == pI) myrow = M;
        const bool below = myrow > I;
        // (d) multipliers; forward substitution step I (B(I) <-> B(M) is the
        // same map swap): B(K) = B(K) - A(K, I) B(I) for K > I
        const double mk = below ? col * piv : 0.0;
        const double bI = readlane_f64(b, pM);
        if (below) b = b - mk * bI;
        // (e) elimination of columns I+1..N with the pivot row's entries
#pragma unroll
        for (int J = 0; J < N; ++J) {  // 0-based column
            if (J < I - 1) continue;
            if (J == I - 1) {
                if (ln == pM) a[J] = piv;
                else if (below) a[J] = mk;
                continue;
            }
            const double c = readlane_f64(a[J], pM);
            if (c != 0.0 && below) a[J] = a[J] - mk * c;
            if (J == I) col = a[J];
        }
    }
    {
        const int pN = (int)__builtin_ctzll(__ballot(myrow == N));
        double p = readlane_f64(a[N - 1], pN);
        if (fabs(p) < TOL) p = copysign(TOL, p);
        if (ln == pN) a[N - 1] = 1.0 / p;
    }
    // back substitution: B(J) = B(J) A(J, J); B(K) = B(K) - A(K, J) B(J), K < J
#pragma unroll
    for (int J = N; J >= 2; --J) {
        const int pJ = (int)__builtin_ctzll(__ballot(myrow == J));
        if (ln == pJ) b = b * a[J - 1];
        const double c = readlane_f64(b, pJ);
        if (myrow >= 1 && myrow < J) b = b - a[J - 1] * c;
    }
    if (myrow == 1) b = b * a[0];
    *row = myrow;
    return b;
}

}  // namespace
