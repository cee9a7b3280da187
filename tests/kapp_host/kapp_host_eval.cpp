// The library's KAPP source (atlas12_amd/csrc/a12_kapp_core.h, a12_kapp_host.h)
// compiled for the host, so tests can compare its arithmetic with the
// reference's own KAPP outputs without a GPU (tests/test_kapp.py).
#define A12_HD
#include <math.h>
#include <stdio.h>
#include <vector>
#include "a12_kapp_core.h"
#include "a12_kapp_host.h"
using namespace kapp;
extern "C" int kcpu_eval(const a12_kapp_tables *t, const a12_kapp_state *s, int nw, const double *waves,
                         double *out)
{
    KPack P;
    std::string e = kapp_pack(t, P);
    if (!e.empty()) { fprintf(stderr, "pack: %s\n", e.c_str()); return -1; }
    e = kapp_check_state(s);
    if (!e.empty()) { fprintf(stderr, "state: %s\n", e.c_str()); return -1; }
    KTab T = kapp_tab(P, P.d.data(), P.i.data(), P.f.data());
    std::vector<double> st;
    kapp_pack_state(s, st);
    KState S = kapp_state_view(st.data(), s->nrhox, s->mion);
    const int n = s->nrhox;
    std::vector<double> pl((size_t)kapp_nplanes(T) * n);
    for (int j = 0; j < n; ++j) {
        kp_state(T, S, Q_HOLTSMARK, j, pl.data());
        for (int l = 0; l < T.tb_nlev_tot; ++l) pl[(size_t)(KS_TB + l) * n + j] = kp_state_tb(T, S, l, j);
    }
    std::vector<double> row(KF_N), sig((size_t)T.tb_nlev_tot * nw + 1);
    for (int l = 0; l < T.tb_nlev_tot; ++l)
        for (int w = 0; w < nw; ++w) sig[(size_t)l * nw + w] = kp_tb_sigma(T, l, waves[w]);
    double cell[23];
    for (int w = 0; w < nw; ++w) {
        kp_freq(T, waves[w], row.data());
        for (int j = 0; j < n; ++j) {
            const double tb = kp_cell_tb(T, S, pl.data(), sig.data(), nw, w, waves[w], j);
            kp_cell(T, S, pl.data(), row.data(), waves[w], j, tb, cell);
            for (int k = 0; k < 23; ++k) out[((size_t)w * 23 + k) * n + j] = cell[k];
        }
    }
    return 0;
}
// What a12_kapp_set_tables would say about these tables ("" when usable).
extern "C" const char *kcpu_check_tables(const a12_kapp_tables *t)
{
    static std::string msg;
    KPack P;
    msg = kapp_pack(t, P);
    return msg.c_str();
}
