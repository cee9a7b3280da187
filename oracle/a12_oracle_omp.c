/*
 * a12_oracle_omp.c -- multi-threaded drivers over the oracle routines, for
 * the CPU baseline only (test infrastructure).  Each thread owns whole
 * depths (line opacity) or whole wavelength ranges (transfer), so results
 * are identical to the single-threaded routines.
 */
#include "a12_oracle.h"
#include "a12_oracle_atlas.h"

#include <omp.h>
#include <stdlib.h>
#include <string.h>

int or_omp_threads(int n)
{
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
}

void or_synthe_line_opacity_omp(const or_synthe_atm *atm, const or_line_t *lines,
                                int nlines, int prec, int j0, int j1, float *out32,
                                double *out64, int64_t *evals)
{
    int64_t tot = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : tot)
    for (int j = j0; j < j1; ++j) {
        int64_t ev = 0;
        or_synthe_line_opacity_range(atm, lines, nlines, prec, j, j + 1, out32, out64, &ev,
                                     NULL);
        tot += ev;
    }
    if (evals) *evals = tot;
}

void or_synthe_transfer_omp(const or_synthe_rt *rt, const float *opac,
                            const double *opac64, double *h_line,
                            double *h_cont, int i0, int i1)
{
    int chunk = 256;
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = i0; b < i1; b += chunk) {
        int e = b + chunk < i1 ? b + chunk : i1;
        or_synthe_transfer(rt, opac, opac64, h_line + (b - i0), h_cont + (b - i0), b, e);
    }
}

void or_synthe_line_opacity_omp_list(const or_synthe_atm *atm, const or_line_t *lines,
                                     int nlines, int prec, const int *js, int nj,
                                     float *out32, double *out64, int64_t *evals)
{
    int64_t tot = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : tot)
    for (int k = 0; k < nj; ++k) {
        int64_t ev = 0;
        or_synthe_line_opacity_range(atm, lines, nlines, prec, js[k], js[k] + 1, out32, out64,
                                     &ev, NULL);
        tot += ev;
    }
    if (evals) *evals = tot;
}

void or_linop1_omp(const or_atlas_lines_in *in, double *xlines, int64_t *evals)
{
    const size_t plane = (size_t)in->nrhox * in->numnu;
    const int nt = omp_get_max_threads();
    double *part = (double *)calloc(plane * (size_t)nt, sizeof(double));
    int64_t tot = 0;
#pragma omp parallel reduction(+ : tot)
    {
        const int t = omp_get_thread_num(), T = omp_get_num_threads();
        const int64_t l0 = in->nlines * (int64_t)t / T, l1 = in->nlines * (int64_t)(t + 1) / T;
        int64_t ev = 0;
        or_linop1_range(in, l0, l1, part + plane * t, &ev);
        tot += ev;
    }
    memset(xlines, 0, sizeof(double) * plane);
    for (int t = 0; t < nt; ++t)
        for (size_t i = 0; i < plane; ++i) xlines[i] += part[plane * t + i];
    free(part);
    if (evals) *evals = tot;
}

/* One depth j of the SYNTHE line loop with the lines split into contiguous
 * chunks, one per thread, each summed into its own row (prec OR_PREC_F64:
 * FP64 rows into row64; OR_PREC_REF: the reference's FP32 arithmetic into
 * FP32 rows, added into row32 in FP32); the rows are then added in chunk
 * order.  The contribution count and every profile value are those of the
 * serial routine; only the order of the sums differs (rounding level).
 * For checks at list sizes where one serial depth takes minutes. */
void or_synthe_depth_split(const or_synthe_atm *atm, const or_line_t *lines, int nlines,
                           int prec, int j, float *row32, double *row64, int64_t *evals)
{
    const int L = atm->length;
    or_synthe_atm one = *atm;  /* a one-layer view of depth j */
    one.nrhox = 1;
    one.continall = atm->continall + (size_t)j * atm->nf;
    one.xnfpel = atm->xnfpel + (size_t)j * atm->mw6;
    one.dopple = atm->dopple + (size_t)j * atm->mw6;
    one.rho = atm->rho + j;
    one.t = atm->t + j;
    one.xne = atm->xne + j;
    one.xnfh = atm->xnfh + j;
    one.xnfhe1 = atm->xnfhe1 + j;
    one.xnfh2 = atm->xnfh2 + j;
    one.hckt = atm->hckt + j;
    const int nt = omp_get_max_threads();
    double *p64 = prec == OR_PREC_F64 ? (double *)calloc((size_t)L * nt, sizeof(double)) : NULL;
    float *p32 = prec == OR_PREC_F64 ? NULL : (float *)calloc((size_t)L * nt, sizeof(float));
    int64_t *ev = (int64_t *)calloc((size_t)nt, sizeof(int64_t));
#pragma omp parallel
    {
        const int t = omp_get_thread_num(), T = omp_get_num_threads();
        const int64_t l0 = (int64_t)nlines * t / T, l1 = (int64_t)nlines * (t + 1) / T;
        or_synthe_line_opacity_range(&one, lines + l0, (int)(l1 - l0), prec, 0, 1,
                                     p32 ? p32 + (size_t)L * t : NULL,
                                     p64 ? p64 + (size_t)L * t : NULL, &ev[t], NULL);
    }
    int64_t tot = 0;
    for (int t = 0; t < nt; ++t) tot += ev[t];
    if (p64) {
        for (int i = 0; i < L; ++i) row64[i] = 0.0;
        for (int t = 0; t < nt; ++t)
            for (int i = 0; i < L; ++i) row64[i] += p64[(size_t)L * t + i];
    } else {
        for (int i = 0; i < L; ++i) row32[i] = 0.0f;
        for (int t = 0; t < nt; ++t)
            for (int i = 0; i < L; ++i) row32[i] += p32[(size_t)L * t + i];
    }
    free(p64);
    free(p32);
    free(ev);
    if (evals) *evals = tot;
}
