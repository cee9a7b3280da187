// a12_molec.hip -- NMOLEC's gas-phase Newton solve, batched on gfx950.
//
// Restates the Newton iteration of NMOLEC, src/atlas12_modules.f90:
// 9078-9217, in FP64: residual EQ and Jacobian DEQ over NEQUA <= 64
// equations (total particles, one per element, charge), SOLVIT, the damped
// update (x0.69 on a sign flip, the factor-of-100 limiter with its SCALE
// that halves in log on a flip), convergence when every |dXN/XN| <= 1e-4,
// at most 5000 iterations; then XNMOL (:9624-9636).
//
// One wavefront per chain of systems.  A chain is NMOLEC's depth loop: its
// first system starts from the caller's XN, each later one from the
// previous solution times P(J)/P(J-1) (:8879-8884).  A chain of length one
// is an independent solve (a warm start such as XNSAVE, :8888-8892), so a
// batch of models x depths fills the GPU with one wave per system.
//
// Everything of one system lives in LDS.  Lane K-1 owns Jacobian column K
// and EQ(K): it walks, in the reference's molecule order, the molecules with
// a component in equation K, so every entry receives its terms in the
// reference's order; lane 0 also sums EQ(1) over all molecules.  The
// molecule products TERM are formed in parallel beforehand (each product
// in component order).  SOLVIT is the shared one-wave LU.  Results equal
// the serial routine's bit for bit.

#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "a12_common.h"
#include "a12_internal.h"
#include "a12_wave.h"

namespace {

constexpr int kMaxEq = A12_MOL_MAXEQ;
constexpr int kMaxMol = A12_MOL_MAXMOL;
constexpr int kMaxLoc = 3 * A12_MOL_MAXMOL;

// molecule flags
constexpr int kSingle = 1;   // one component: no Newton term
constexpr int kNegIon = 2;   // last component is the electron equation

struct MolLDS {
    double deq[kMaxEq * kMaxEq];
    double eq[kMaxEq], xn[kMaxEq], eqold[kMaxEq], xab[kMaxEq];
    double equilj[kMaxMol], term[kMaxMol];
    int locj[kMaxMol + 1], kcomps[kMaxLoc], mflag[kMaxMol];
    int colptr[kMaxEq + 1], colmol[kMaxLoc];
    int ipiv[kMaxEq];
    int flag[2];
};

// tab layout (int32): locj (nummol+1), kcomps (nloc), mflag (nummol),
// colptr (nequa+1), colmol (colptr[nequa])
__global__ __launch_bounds__(64) void k_molec_newton(
    int ne, int nummol, int nloc, const int *__restrict__ tab, const int *__restrict__ chain_off,
    const double *__restrict__ xntot, const double *__restrict__ pratio,
    const double *__restrict__ xab, const double *__restrict__ equilj, double *__restrict__ xn_io,
    double *__restrict__ xnmol, int *__restrict__ iters, int *__restrict__ status, int maxit,
    int charge)
{
#pragma clang fp contract(off)
    __shared__ MolLDS S;
    const int ln = lane_id();
    const int ne1 = ne + 1;
    {
        const int *t = tab;
        for (int i = ln; i <= nummol; i += 64) S.locj[i] = t[i];
        t += nummol + 1;
        for (int i = ln; i < nloc; i += 64) S.kcomps[i] = t[i];
        t += nloc;
        for (int i = ln; i < nummol; i += 64) S.mflag[i] = t[i];
        t += nummol;
        for (int i = ln; i <= ne; i += 64) S.colptr[i] = t[i];
        const int ncol = t[ne];
        t += ne + 1;
        for (int i = ln; i < ncol; i += 64) S.colmol[i] = t[i];
    }
#define DEQ(i, j) S.deq[((j) - 1) * ne + ((i) - 1)]
    const int s0 = chain_off[blockIdx.x], s1 = chain_off[blockIdx.x + 1];
    const int K = ln + 1;  // this lane's column / equation
    for (int s = s0; s < s1; ++s) {
        wsync();
        if (K <= ne) {
            S.xn[ln] = s == s0 ? xn_io[(size_t)s * ne + ln] : S.xn[ln] * pratio[s];
            S.xab[ln] = xab[(size_t)s * ne + ln];
            S.eqold[ln] = 0.0;
        }
        for (int m = ln; m < nummol; m += 64) S.equilj[m] = equilj[(size_t)s * nummol + m];
        const double xt = xntot[s];
        int it = 0, st = 0;
        for (;;) {
            ++it;
            wsync();
            for (int i = ln; i < ne * ne; i += 64) S.deq[i] = 0.0;
            for (int m = ln; m < nummol; m += 64) {  // TERM = K prod n_k / n_e
                if (S.mflag[m] & kSingle) continue;
                double term = S.equilj[m];
                for (int l = S.locj[m]; l < S.locj[m + 1]; ++l) {
                    const int k = S.kcomps[l - 1];
                    term = k == ne1 ? term / S.xn[ne - 1] : term * S.xn[k - 1];
                }
                S.term[m] = term;
            }
            wsync();
            if (K == 1) {  // EQ(1) and column 1 (:9083-9097, :9118)
                double e1 = -xt;
                for (int k = 2; k <= ne; ++k) {
                    e1 = e1 + S.xn[k - 1];
                    DEQ(k, 1) = -S.xab[k - 1];
                }
                for (int m = 0; m < nummol; ++m)
                    if (!(S.mflag[m] & kSingle)) e1 = e1 + S.term[m];
                S.eq[0] = e1;
            } else if (K <= ne) {
                double ek = S.xn[K - 1] - S.xab[K - 1] * S.xn[0];
                DEQ(1, K) = 1.0;
                DEQ(K, K) = 1.0;
                if (charge && K == ne) {
                    ek = -S.xn[ne - 1];
                    DEQ(ne, ne) = -1.0;
                }
                for (int c = S.colptr[K - 1]; c < S.colptr[K]; ++c) {
                    const int m = S.colmol[c];
                    const double term = S.term[m];
                    const int l1 = S.locj[m], l2 = S.locj[m + 1] - 1;
                    for (int lk = l1; lk <= l2; ++lk) {  // :9121-9140
                        int k = S.kcomps[lk - 1];
                        double d;
                        if (k >= ne1) {
                            k = ne;
                            d = -term / S.xn[ne - 1];
                        } else {
                            d = term / S.xn[k - 1];
                        }
                        if (k != K) continue;
                        ek = ek + term;
                        DEQ(1, K) = DEQ(1, K) + d;
                        for (int lm = l1; lm <= l2; ++lm) {
                            int mm = S.kcomps[lm - 1];
                            if (mm == ne1) mm = ne;
                            DEQ(mm, K) = DEQ(mm, K) + d;
                        }
                    }
                    if (S.mflag[m] & kNegIon) {  // :9145-9160
                        for (int lk = l1; lk <= l2; ++lk) {
                            if (S.kcomps[lk - 1] != K) continue;
                            const double d = term / S.xn[K - 1];
                            if (K == ne) ek = ek - term - term;
                            for (int lm = l1; lm <= l2; ++lm)
                                if (S.kcomps[lm - 1] == ne) DEQ(ne, K) = DEQ(ne, K) - d - d;
                        }
                    }
                }
                S.eq[K - 1] = ek;
            }
            wsync();
            solvit_lds(S.deq, S.eq, S.ipiv, ne);  // :9165
            if (ln == 0) {                         // damped update :9168-9189
                int iferr = 0;
                double scale = 100.0;
                for (int k = 0; k < ne; ++k) {
                    double e = S.eq[k];
                    const double x = S.xn[k];
                    if (fabs(e / x) > 1.0e-4) iferr = 1;
                    if (S.eqold[k] * e < 0.0) e = e * 0.69;
                    const double xneq = x - e;
                    const double xn100 = x / 100.0;
                    if (fabs(xneq) >= xn100) {
                        S.xn[k] = fabs(xneq);
                    } else {
                        S.xn[k] = x / scale;
                        if (S.eqold[k] * e < 0.0) scale = sqrt(scale);
                    }
                    S.eqold[k] = e;
                    S.eq[k] = e;
                }
                S.flag[0] = iferr;
                S.flag[1] = (S.xn[0] != S.xn[0] || S.eq[0] != S.eq[0]) ? 1 : 0;
            }
            wsync();
            if (S.flag[1]) {
                st = 1;
                break;
            }
            if (S.flag[0] == 0) break;
            if (it >= maxit) {
                st = 2;
                break;
            }
        }
        if (K <= ne) xn_io[(size_t)s * ne + ln] = S.xn[ln];
        if (xnmol)
            for (int m = ln; m < nummol; m += 64) {  // :9624-9636
                double v = S.equilj[m];
                for (int l = S.locj[m]; l < S.locj[m + 1]; ++l) {
                    const int k = S.kcomps[l - 1];
                    v = k == ne1 ? v / S.xn[ne - 1] : v * S.xn[k - 1];
                }
                xnmol[(size_t)s * nummol + m] = v;
            }
        if (ln == 0) {
            iters[s] = it;
            status[s] = st;
        }
    }
#undef DEQ
}

}  // namespace

int a12i_molec_set_table(a12_ctx *ctx, const a12_moltab *tb)
{
    MolDev &M = ctx->mol;
    M.ready = false;
    if (!tb || !tb->locj || !tb->kcomps || !tb->idequa || tb->nequa < 2 ||
        tb->nequa > kMaxEq || tb->nummol < 1 || tb->nummol > kMaxMol) {
        ctx->err = "a12_molec_set_table: need 2 <= nequa <= 64, 1 <= nummol <= 400 and all arrays";
        return A12_ERR_ARG;
    }
    const int ne = tb->nequa, nm = tb->nummol;
    if (tb->locj[0] != 1) {
        ctx->err = "a12_molec_set_table: locj must start at 1";
        return A12_ERR_ARG;
    }
    for (int m = 0; m < nm; ++m)
        if (tb->locj[m + 1] <= tb->locj[m]) {
            ctx->err = "a12_molec_set_table: every entry needs at least one component";
            return A12_ERR_ARG;
        }
    const int nloc = tb->locj[nm] - 1;
    if (nloc > kMaxLoc) {
        ctx->err = "a12_molec_set_table: more than 1200 components";
        return A12_ERR_ARG;
    }
    for (int l = 0; l < nloc; ++l)
        if (tb->kcomps[l] < 2 || tb->kcomps[l] > ne + 1) {
            ctx->err = "a12_molec_set_table: component " + std::to_string(l + 1) +
                       " is not an element or the electron (2..nequa+1)";
            return A12_ERR_ARG;
        }
    std::vector<int> mflag(nm, 0);
    std::vector<std::vector<int>> col(ne + 1);
    for (int m = 0; m < nm; ++m) {
        const int l1 = tb->locj[m], l2 = tb->locj[m + 1] - 1;
        if (l1 == l2) {
            mflag[m] |= kSingle;
            continue;
        }
        const int kl = tb->kcomps[l2 - 1];
        if (kl <= ne && tb->idequa[kl - 1] == 100) {
            mflag[m] |= kNegIon;
            for (int l = l1; l <= l2; ++l)
                if (tb->kcomps[l - 1] == ne + 1) {
                    ctx->err = "a12_molec_set_table: negative ion " + std::to_string(m + 1) +
                               " also divides by the electron density";
                    return A12_ERR_ARG;
                }
        }
        for (int l = l1; l <= l2; ++l) {
            const int k = tb->kcomps[l - 1] == ne + 1 ? ne : tb->kcomps[l - 1];
            if (col[k].empty() || col[k].back() != m) col[k].push_back(m);
        }
    }
    std::vector<int> tab;
    tab.insert(tab.end(), tb->locj, tb->locj + nm + 1);
    tab.insert(tab.end(), tb->kcomps, tb->kcomps + nloc);
    tab.insert(tab.end(), mflag.begin(), mflag.end());
    int run = 0;
    for (int k = 1; k <= ne; ++k) {  // colptr[k-1] = start of column k
        tab.push_back(run);
        run += (int)col[k].size();
    }
    tab.push_back(run);
    for (int k = 1; k <= ne; ++k) tab.insert(tab.end(), col[k].begin(), col[k].end());
    A12_HIP(M.tab.reserve(sizeof(int) * tab.size()));
    A12_HIP(hipMemcpyAsync(M.tab.p, tab.data(), sizeof(int) * tab.size(), hipMemcpyHostToDevice,
                           ctx->stream));
    A12_HIP(hipStreamSynchronize(ctx->stream));
    M.nequa = ne;
    M.nummol = nm;
    M.nloc = nloc;
    M.charge = tb->idequa[ne - 1] >= 100;
    M.ready = true;
    return A12_OK;
}

int a12i_molec_newton(a12_ctx *ctx, int32_t nsys, int32_t nchain, const int32_t *chain_off,
                      const double *xntot, const double *pratio, const double *xab,
                      const double *equilj, double *xn, double *xnmol, int32_t *iters,
                      int32_t *status)
{
    MolDev &M = ctx->mol;
    if (!M.ready) {
        ctx->err = "a12_molec_newton_batched: call a12_molec_set_table first";
        return A12_ERR_STATE;
    }
    if (nsys < 0 || nchain < 0 || (nsys > 0 && (!chain_off || !xntot || !xab || !equilj || !xn ||
                                                 !iters || !status))) {
        ctx->err = "a12_molec_newton_batched: bad sizes or null array";
        return A12_ERR_ARG;
    }
    if (nsys == 0) return A12_OK;
    if (nchain < 1 || chain_off[0] != 0 || chain_off[nchain] != nsys) {
        ctx->err = "a12_molec_newton_batched: chain_off must run from 0 to nsys";
        return A12_ERR_ARG;
    }
    bool chained = false;
    for (int c = 0; c < nchain; ++c) {
        if (chain_off[c + 1] <= chain_off[c]) {
            ctx->err = "a12_molec_newton_batched: empty or decreasing chain " + std::to_string(c);
            return A12_ERR_ARG;
        }
        chained |= chain_off[c + 1] - chain_off[c] > 1;
    }
    if (chained && !pratio) {
        ctx->err = "a12_molec_newton_batched: chains longer than one need pratio";
        return A12_ERR_ARG;
    }
    const int ne = M.nequa, nm = M.nummol;
    const size_t n = (size_t)nsys;
    // device layout (doubles): xntot, pratio, xab (n, ne), equilj (n, nm), xn (n, ne),
    // xnmol (n, nm); then ints: chain_off (nchain + 1), iters, status
    const size_t nd = n * (2 + 2 * (size_t)ne + 2 * (size_t)nm);
    const size_t ni = (size_t)nchain + 1 + 2 * n;
    A12_HIP(M.io.reserve(sizeof(double) * nd + sizeof(int) * ni));
    double *d = M.io.as<double>();
    double *d_xt = d, *d_pr = d + n, *d_xab = d + 2 * n, *d_eq = d_xab + n * ne,
           *d_xn = d_eq + n * nm, *d_xm = d_xn + n * ne;
    int *d_co = reinterpret_cast<int *>(d + nd), *d_it = d_co + nchain + 1, *d_st = d_it + n;
    hipStream_t s = ctx->stream;
    A12_HIP(hipMemcpyAsync(d_xt, xntot, sizeof(double) * n, hipMemcpyHostToDevice, s));
    if (pratio) A12_HIP(hipMemcpyAsync(d_pr, pratio, sizeof(double) * n, hipMemcpyHostToDevice, s));
    A12_HIP(hipMemcpyAsync(d_xab, xab, sizeof(double) * n * ne, hipMemcpyHostToDevice, s));
    A12_HIP(hipMemcpyAsync(d_eq, equilj, sizeof(double) * n * nm, hipMemcpyHostToDevice, s));
    A12_HIP(hipMemcpyAsync(d_xn, xn, sizeof(double) * n * ne, hipMemcpyHostToDevice, s));
    A12_HIP(hipMemcpyAsync(d_co, chain_off, sizeof(int) * (nchain + 1), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_molec_newton, dim3(nchain), dim3(64), 0, s, ne, nm, M.nloc,
                       M.tab.as<int>(), d_co, d_xt, d_pr, d_xab, d_eq, d_xn,
                       xnmol ? d_xm : nullptr, d_it, d_st, 5000, M.charge ? 1 : 0);
    A12_CHECK_LAUNCH();
    A12_HIP(hipMemcpyAsync(xn, d_xn, sizeof(double) * n * ne, hipMemcpyDeviceToHost, s));
    if (xnmol) A12_HIP(hipMemcpyAsync(xnmol, d_xm, sizeof(double) * n * nm, hipMemcpyDeviceToHost, s));
    A12_HIP(hipMemcpyAsync(iters, d_it, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    A12_HIP(hipMemcpyAsync(status, d_st, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    A12_HIP(hipStreamSynchronize(s));
    return A12_OK;
}
