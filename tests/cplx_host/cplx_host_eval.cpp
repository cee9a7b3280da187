// Host build of a12_cplx_core.h: compute_line_opacity's BUFFER for chosen
// depths, with the same plan / convolution / walk / per-point functions the
// device kernels run, summed per point in line order.  Called by
// tests/test_cplx.py through ctypes.
#define A12_HD
#include <string.h>

#include <vector>

#include "a12_cplx_core.h"

using namespace cplx;

extern "C" {

// depth_xf (ndep, 9): xf_t, xf_xne, xf_hckt, xf_rho, xf_xnfh, xf_xnfhe1,
// xf_xnfh2, qxnfpel(1), qxnfpel(2) of each depth.  xnfdop, dopple, xnfpel
// (ndep, mw6): the driver's per-depth arrays (xnfpel divided by RHO).
// cont (ndep, length).  grid_d: wbegin, ratio, ratiolg, wlend, cutoff,
// dopratio; grid_i: length, ixwlbeg, ifvac.  depth_out (ndep, 4): txnxn,
// xnfph1, xnfph2, emerge.  buffer (ndep, length).  Returns the number of
// convolutions built.
// flags (planted mistakes, for the tests that show the fixtures catch
// them): 1 = every H line builds its own convolution (no cache sharing),
// 2 = wing wavelengths as wbegin * ratio**(p-1) instead of the running
// product.
int ccpu_run(const a12_cplx_line *lines, int nlines, const double *grid_d, const int *grid_i, int ndep,
             const double *depth_xf, int mw6, const double *xnfdop, const double *dopple,
             const double *xnfpel, const float *cont, const a12_kapp_stark *stark, double *depth_out,
             int *extents, float *buffer, int flags)
{
    Grid G;
    G.wbegin = grid_d[0];
    G.ratio = grid_d[1];
    G.ratiolg = grid_d[2];
    G.wlend = grid_d[3];
    G.cutoff = grid_d[4];
    G.dopratio = grid_d[5];
    G.length = grid_i[0];
    G.ixwlbeg = grid_i[1];
    G.ifvac = grid_i[2];
    const int L = G.length;
    double dg[4 * NDA];
    stark_pow10(stark, dg);
    Stark S;
    stark_view(stark, dg, S);
    std::vector<HLine> hl(nlines);
    for (int l = 0; l < nlines; ++l)
        if (lines[l].itype == -1 || lines[l].itype == -2) hline_setup(lines[l].nblo, lines[l].nbup, hl[l]);
    std::vector<Plan> plan(nlines);
    std::vector<HConv> conv(nlines);
    std::vector<int> owner(nlines);
    std::vector<double> da(NCONV_MAX), pr(NCONV_MAX), step(NCONV_MAX);
    int nconv = 0;
    for (int j = 0; j < ndep; ++j) {
        const double *x = depth_xf + (size_t)j * 9;
        Depth D;
        depth_setup(x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7], x[8], D);
        depth_out[j * 4 + 0] = D.txnxn;
        depth_out[j * 4 + 1] = D.xnfph1;
        depth_out[j * 4 + 2] = D.xnfph2;
        depth_out[j * 4 + 3] = D.emerge;
        const float *cj = cont + (size_t)j * L;
        for (int l = 0; l < nlines; ++l)
            plan_line(lines[l], G, D, xnfdop + (size_t)j * mw6, dopple + (size_t)j * mw6,
                      xnfpel[(size_t)j * mw6 + lines[l].nelion - 1], cj, plan[l]);
        // the per-transition cache: a line reuses the convolution of the
        // previous line that called the profile when it has the same (n, m)
        int last = -1;
        for (int l = 0; l < nlines; ++l) {
            owner[l] = l;
            const Plan &P = plan[l];
            const double wr0 = chain_start(G, 0, 1, P.minred, P.ib0);
            const double wb0 = chain_start(G, 0, 0, P.minred, P.ib0);
            if (!calls_profile(P, G, wr0, wb0)) continue;
            if (!(flags & 1) && last >= 0 && lines[last].nblo == lines[l].nblo &&
                lines[last].nbup == lines[l].nbup)
                owner[l] = owner[last];
            last = l;
        }
        for (int l = 0; l < nlines; ++l) {
            const Plan &P = plan[l];
            if (owner[l] != l || (P.kind != K_HISO && P.kind != K_HBETA && P.kind != K_HMERGE)) continue;
            if (!use_stehle(S, lines[l].nblo, lines[l].nbup, D.xne)) continue;
            conv_build(hl[l], D, S, P.hwdop, conv[l], da.data(), pr.data(), step.data());
            ++nconv;
        }
        float *b = buffer + (size_t)j * L;
        memset(b, 0, sizeof(float) * L);
        for (int l = 0; l < nlines; ++l) {
            const Plan &P = plan[l];
            const a12_cplx_line &ln = lines[l];
            const bool h = P.kind == K_HISO || P.kind == K_HBETA || P.kind == K_HMERGE;
            const bool st = h && use_stehle(S, ln.nblo, ln.nbup, D.xne);
            const HConv *cv = st ? &conv[owner[l]] : nullptr;
            const Extent e = walk_extent(P, ln, h ? &hl[l] : nullptr, D, S, st, cv, G, cj);
            int *ex = extents + ((size_t)j * nlines + l) * 4;
            ex[0] = e.rlo;
            ex[1] = e.rhi;
            ex[2] = e.blo;
            ex[3] = e.bhi;
            if (P.kind == K_CONT) {
                for (int p = e.rlo; p <= e.rhi; ++p) b[p - 1] = b[p - 1] + cont_kappa(P, p);
                continue;
            }
            const int aut = P.kind == K_AUT;
            if (e.rlo <= e.rhi) {
                double v = chain_at(G, aut, 1, P.minred, P.ib0, e.rlo - P.minred);
                for (int p = e.rlo; p <= e.rhi; ++p) {
                    const double u = (flags & 2) && !aut ? G.wbegin * powi_rt(G.ratio, p - 1) : v;
                    b[p - 1] = b[p - 1] + point_kappa(P, ln, h ? &hl[l] : nullptr, D, S, st, cv, u);
                    v = chain_next(G, aut, 1, v);
                }
            }
            if (e.blo <= e.bhi) {
                double v = chain_at(G, aut, 0, P.minred, P.ib0, P.ib0 - 1 - e.bhi);
                for (int p = e.bhi; p >= e.blo; --p) {
                    const double u = (flags & 2) && !aut ? G.wbegin * powi_rt(G.ratio, p - 1) : v;
                    b[p - 1] = b[p - 1] + point_kappa(P, ln, h ? &hl[l] : nullptr, D, S, st, cv, u);
                    v = chain_next(G, aut, 0, v);
                }
            }
        }
    }
    return nconv;
}

// voigt_profile and hydrogen_line_profile pieces for direct checks.
float ccpu_voigt(float v, float a) { return voigt_profile(v, a); }
float ccpu_stark_qs(float b, float p, int n, int m) { return stark_qs(b, p, n, m); }
float ccpu_expint1(float x) { return expint1(x); }
}
