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
