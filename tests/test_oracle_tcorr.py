"""The oracle's end of TCORR(3) -- ROSSTAB, TTAUP, CZ_CONSTRUCT and
sections (G2)-(K) -- against the reference's own outputs
(tests/golden/tcorr_ref.npz and tcorr_conv_ref.npz, written by
oracle/ref/ref_atlas_dump on a radiative and a convective deck after a full
30,000-frequency pass), and against what the mathematics fixes.

Reference: ROSSTAB src/atlas12_modules.f90:5937-6268, TTAUP :6303-6523,
CZ_CONSTRUCT :3012-3358, TCORR(3) :2816-2967."""
import math

import numpy as np
import pytest

DECKS = ("tcorr_ref", "tcorr_conv_ref")


@pytest.fixture(scope="module", params=DECKS)
def deck(request, golden):
    return golden(request.param + ".npz")


def _taustd(z):
    n = int(z["n"])
    # the reference's 10.0D0**x is libm's pow (numpy's power differs in the
    # last bit at some points)
    return np.array([math.pow(10.0, float(z["fin_tau1lg"]) + j * float(z["fin_steplg"]))
                     for j in range(n)])


def test_rosstab_store_is_the_reference_table(orc, deck):
    """ROSSTAB(0,0,0) after ROSS(3): the normalisation and all 72 entries
    bit for bit."""
    tab = orc.Rosstab()
    assert tab.store(deck["t"], deck["fin_in_p"], deck["abross"]) == int(deck["n"])
    assert np.array_equal(tab.zero(), deck["fin_rt_zero"])
    assert np.array_equal(tab.table(), deck["fin_rt_tab"])
    assert int(deck["fin_nross"]) == tab.nross and int(deck["fin_irosstab"]) == 1


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_rosstab_lookup_matches_reference(orc, deck, mode):
    """864 queries around the model's layers on a two-profile table, in the
    bilinear, Shepard and MLS modes: every value bit for bit."""
    tab = orc.Rosstab(deck["fin_rt_zero"], deck["fin_lk_tab"])
    q = deck["fin_lk_q"]
    got = tab.lookup(q[0], q[1], mode)
    assert np.array_equal(got, deck["fin_lk_r"][mode - 1])


def test_rosstab_lookup_takes_both_mode1_branches(orc, deck):
    """The queries reach the bilinear blend (all four quadrants filled) and
    the inverse-distance fallback (a quadrant empty); both agree above."""
    tab = orc.Rosstab(deck["fin_rt_zero"], deck["fin_lk_tab"])
    zt, st, zp, sp = deck["fin_rt_zero"]
    t, p = deck["fin_lk_tab"][0], deck["fin_lk_tab"][1]
    q = deck["fin_lk_q"]
    tl, pl = (np.log10(q[0]) - zt) / st, (np.log10(q[1]) - zp) / sp
    full = [all(np.any(((t - a >= 0) == sa) & ((p - b >= 0) == sb))
                for sa in (True, False) for sb in (True, False)) for a, b in zip(tl, pl)]
    assert 0 < sum(full) < len(full)


def test_ttaup_matches_reference(orc, deck):
    """TTAUP on T and PRAD mapped onto TAUSTD (as section (H) calls it):
    ABSTD, PTOTAL and P bit for bit."""
    tab = orc.Rosstab(deck["fin_rt_zero"], deck["fin_rt_tab"])
    ab, pt, pg, nf = orc.ttaup(tab, deck["fin_tt_t"], _taustd(deck), deck["fin_tt_prad"],
                               deck["fin_in_pturb"], float(deck["fin_grav"]))
    assert nf == 0
    assert np.array_equal(ab, deck["fin_tt_ab"])
    assert np.array_equal(pt, deck["fin_tt_ptot"])
    assert np.array_equal(pg, deck["fin_tt_p"])


def _finish_inputs(orc, z):
    out, _, _ = orc.tcorr3(z, z["rdabh"])
    n = int(z["n"])
    sc = dict(n=n, ifconv=int(z["ifconv"]), iter=int(z["iter"]), use4pt=1, use_constructor=1,
              j1smooth=int(z["fin_j1smooth"]), j2smooth=int(z["fin_j2smooth"]),
              flux=float(z["flux"]), mixlth=float(z["mixlth"]), grav=float(z["fin_grav"]),
              tau1lg=float(z["fin_tau1lg"]), steplg=float(z["fin_steplg"]),
              pradk0=float(z["fin_pradk0"]), tfloor=float(z["fin_tfloor"]),
              wtjm1=float(z["fin_wtjm1"]), wtj=float(z["fin_wtj"]), wtjp1=float(z["fin_wtjp1"]))
    arr = {k: z[k] for k in ("flxcnv", "flxrad", "dltdlp", "grdadb", "ptotal", "rho", "heatcp",
                             "dlrdlt", "hscale", "tauros", "rhox", "t", "abross")}
    arr["flxerr"] = out["flxerr"]
    for k in ("p", "xne", "prad", "pturb", "vturb", "bmin", "accrad"):
        arr[k] = z["fin_in_" + k]
    arr["bhyd"] = z["fin_in_bhyd"]
    return sc, arr, out


FIN_KEYS = ("rhox", "t", "tk", "hkt", "hckt", "tkev", "tlog", "p", "xne", "abross", "prad",
            "pradk", "vturb", "bmin", "pturb", "tauros")


def test_tcorr3_whole_matches_reference_state(orc, deck):
    """Sections A-F then (G2)-(K): the module arrays TCORR(3) leaves (the
    structure remapped onto TAUSTD, the thermodynamic vectors, BHYD) and
    FLXERR equal the reference's bit for bit."""
    sc, arr, out = _finish_inputs(orc, deck)
    assert np.array_equal(out["flxerr"], deck["fin_out_flxerr"])
    tab = orc.Rosstab(deck["fin_rt_zero"], deck["fin_rt_tab"])
    st = orc.czc_state()
    got, _, drhox, nf = orc.tcorr3_finish(tab, sc, arr, out["t1"], st)
    for k in FIN_KEYS:
        assert np.array_equal(got[k], deck["fin_out_" + k]), k
    assert np.array_equal(got["bhyd"], deck["fin_out_bhyd"])
    assert st.run_engaged == int(deck["fin_out_czc_engaged"])
    assert np.any(drhox != 0.0)


def test_tcorr3_finish_invariants(orc, deck):
    """What the mathematics fixes, independent of the reference's numbers:
    TAUROS becomes TAUSTD; PRADK = PRAD + PRADK0; TK, HKT, HCKT, TKEV, TLOG
    are the constants' functions of T; with T1 = 0 and DRHOX = 0 the state is
    only remapped, so on a model already on TAUSTD nothing but the floors
    changes."""
    sc, arr, out = _finish_inputs(orc, deck)
    tab = orc.Rosstab(deck["fin_rt_zero"], deck["fin_rt_tab"])
    got, _, _, _ = orc.tcorr3_finish(tab, sc, arr, out["t1"], orc.czc_state())
    assert np.array_equal(got["tauros"], _taustd(deck))
    assert np.array_equal(got["pradk"], got["prad"] + float(deck["fin_pradk0"]))
    kb, h, c = 1.380649e-16, 6.62607015e-27, 2.99792458e10
    assert np.array_equal(got["tk"], kb * got["t"])
    assert np.array_equal(got["hkt"], h / got["tk"])
    assert np.array_equal(got["hckt"], got["hkt"] * c)
    assert np.array_equal(got["tkev"], (kb / 1.602176634e-12) * got["t"])
    assert np.array_equal(got["tlog"], np.log(got["t"]))
    # a second pass from the remapped state with T1 = 0: T and T + T1 are
    # the same profile, so both TTAUP marches agree and DRHOX is exactly 0;
    # MAP1 onto the abscissae it is given returns the ordinates (to
    # rounding), and the monotonicity rule leaves T rising by >= 1 K
    arr2 = dict(arr)
    for k in ("tauros", "rhox", "t", "p", "xne", "abross", "prad", "vturb", "bmin", "pturb"):
        arr2[k] = got[k]
    arr2["bhyd"] = got["bhyd"]
    arr2["accrad"] = got["prad"]
    sc2 = dict(sc, ifconv=0)
    again, _, drhox, _ = orc.tcorr3_finish(tab, sc2, arr2, np.zeros(int(deck["n"])),
                                          orc.czc_state())
    assert np.array_equal(drhox, np.zeros_like(drhox))
    for k in ("rhox", "p", "xne", "abross", "prad", "vturb", "bmin", "pturb"):
        assert np.allclose(again[k], got[k], rtol=1e-12, atol=0), k
    assert np.all(np.diff(again["t"]) >= 1.0 - 1e-9)
    assert np.all(again["t"] >= float(deck["fin_tfloor"]))


def test_cz_construct_matches_reference(orc, deck):
    """CZ_CONSTRUCT on planted inputs that engage it (convective flux in the
    deepest 24 layers, flux errors 8, 4, 0.5 %): T1, the engagement and the
    block top over three calls with its SAVE state carried, bit for bit --
    engage, stay engaged with the sticky block, release."""
    tab = orc.Rosstab(deck["fin_rt_zero"], deck["fin_lk_tab"])
    z = {k: deck["fin_cz_" + k] for k in ("flxcnv", "flxrad", "dltdlp", "grdadb", "t", "rhox",
                                           "ptotal", "rho", "heatcp", "dlrdlt", "p", "vturb",
                                           "abross", "hscale")}
    z.update(flux=float(deck["fin_cz_flux"]), grav=float(deck["fin_cz_grav"]),
             mixlth=float(deck["fin_cz_mixlth"]))
    st = orc.czc_state()
    n = int(deck["n"])
    assert int(deck["fin_cz_engaged"][0]) == 1
    for k in range(3):
        z["flxerr"] = deck["fin_cz_flxerr"][k]
        t1, e = orc.cz_construct(tab, z, np.zeros(n), st, k + 1, use4pt=int(deck["fin_cz_4pt"]))
        assert np.array_equal(t1, deck["fin_cz_t1"][k]), k
        assert st.run_engaged == int(deck["fin_cz_engaged"][k])
        assert st.run_jc == int(deck["fin_cz_jc"][k])
    assert np.any(deck["fin_cz_t1"][0] != 0.0)


def test_rosstab_bilinear_reproduces_a_plane(orc):
    """Mode 1 where log kappa is a plane in (log T, log P): both blends are
    linear interpolations, so any four quadrant neighbours return the plane
    (to rounding).  A diagonal profile fixes the normalisation, then a
    lattice of profiles fills the quadrants of every interior query."""
    def plane(lt, lp):
        return 0.3 + 2.0 * lt - 0.5 * lp
    tab = orc.Rosstab()
    lt = np.linspace(3.5, 4.2, 15)
    lp = np.linspace(1.0, 6.0, 15)
    tab.store(10.0 ** lt, 10.0 ** lp, 10.0 ** plane(lt, lp))
    for q in np.linspace(1.0, 6.0, 11):
        tab.store(10.0 ** lt, np.full(15, 10.0 ** q), 10.0 ** plane(lt, q))
    rng = np.random.default_rng(3)
    qt = rng.uniform(3.56, 4.14, 200)
    qp = rng.uniform(1.6, 5.4, 200)
    got = np.log10(tab.lookup(10.0 ** qt, 10.0 ** qp, 1))
    assert np.max(np.abs(got - plane(qt, qp))) < 1e-12


def test_ttaup_constant_opacity_closed_form(orc):
    """With kappa constant and no radiation or turbulent pressure,
    d ln P / d ln tau = 1 exactly: P = g tau / kappa, which the predictor-
    corrector integrates to its convergence tolerance."""
    tab = orc.Rosstab()
    t = np.linspace(4000.0, 9000.0, 72)
    p = np.logspace(1.0, 6.0, 72)
    tab.store(t, p, np.full(72, 0.37))
    tau = np.array([math.pow(10.0, -6.875 + 0.125 * j) for j in range(72)])
    z = np.zeros(72)
    ab, pt, pg, nf = orc.ttaup(tab, t, tau, z, z, 2.75e4)
    assert nf == 0
    assert np.allclose(ab, 0.37, rtol=1e-13, atol=0)
    assert np.max(np.abs(np.log(pt / (2.75e4 * tau / 0.37)))) < 2e-4
    assert np.array_equal(pg, pt)


CNV_IN = ("t", "tk", "hkt", "hckt", "tkev", "tlog", "p", "ptotal", "rho", "pradk", "tauros",
          "rhox", "abross", "vturb", "height", "flxrad", "flxcnv1")
CNV_OUT = ("dltdlp", "heatcp", "velsnd", "dlrdlt", "grdadb", "hscale", "vconv", "flxcnv",
           "flxcnv0", "flxcnv1")


def convec_inputs(z, overwt=None):
    g = lambda k: z["cnv_" + k]  # noqa: E731
    sc = dict(ifconv=1, ifcorr=int(g("ifcorr")), nconv=int(g("nconv")), use4pt=int(g("use4pt")),
              smooth=int(g("smooth")), grav=float(g("grav")), mixlth=float(g("mixlth")),
              flux=float(g("flux")), overwt=float(g("overwt") if overwt is None else overwt))
    arr = {k: g("in_" + k) for k in CNV_IN}
    return sc, arr, g("edens"), g("rho4")


def test_convec_matches_reference(orc, deck):
    """CONVEC on the EOS evaluations it made (recorded by the harness from
    the same snapshot): the perturb-and-restore drift of T, TK, HKT, HCKT,
    TKEV, TLOG, P, the thermodynamic derivatives, and the MLT solve with its
    post-processing, bit for bit (26 MLT layers, gap fill, smoothing and
    surface zeroing on the convective deck)."""
    sc, arr, e, r = convec_inputs(deck)
    tab = orc.Rosstab(deck["cnv_rt_zero"], deck["cnv_rt_tab"])
    out = orc.convec(tab, sc, arr, e, r)
    for k in ("t", "tk", "hkt", "hckt", "tkev", "tlog", "p"):
        assert np.array_equal(out[k], deck["cnv_st_" + k]), k
    for k in CNV_OUT:
        assert np.array_equal(out[k], deck["cnv_out_" + k]), k


def test_convec_overshooting_matches_reference(orc, deck):
    """The same call with OVERWT = 0.5 (the reference's default is 0): the
    overshoot flux FLXCNV1 from INTEG and MAP1 over HEIGHT, FLXCNV =
    max(FLXCNV0, FLXCNV1), bit for bit."""
    sc, arr, e, r = convec_inputs(deck, overwt=0.5)
    tab = orc.Rosstab(deck["cnv_rt_zero"], deck["cnv_rt_tab"])
    out = orc.convec(tab, sc, arr, e, r)
    for k in ("flxcnv", "flxcnv0", "flxcnv1"):
        assert np.array_equal(out[k], deck["cnv_ovr_" + k]), k


def test_convec_thermodynamics_closed_form(orc, deck):
    """For an ideal monatomic gas without radiation pressure (E = 3/2 P/rho,
    rho = P mu / (k T)) the finite differences give the textbook values:
    d ln rho / d ln T = -1 at constant P, HEATCP = 5/2 P/(rho T), so
    nabla_ad = -P/(rho T) DLRDLT / HEATCP = 2/5, to the differences'
    O(1e-6); with MIXLTH = 0 no depth convects."""
    sc, arr, _, _ = convec_inputs(deck)
    n = len(arr["t"])
    t0, p0 = arr["t"], arr["p"]
    mu_k = 1.3e-8  # mu / k in g/(erg/K) units: any positive constant
    rho_at = lambda t, p: p * mu_k / t  # noqa: E731
    # the four states of CONVEC's perturbations, from the T and P it starts at
    tp, tm, pp, pm = t0 * 1.001, t0 * 0.999, p0 * 1.001, p0 * 0.999
    r4 = np.array([rho_at(tp, p0), rho_at(tm, p0), rho_at(t0, pp), rho_at(t0, pm)])
    e4 = np.array([1.5 * p0 / r4[0], 1.5 * p0 / r4[1], 1.5 * pp / r4[2], 1.5 * pm / r4[3]])
    a = dict(arr, pradk=np.zeros(n), rho=rho_at(t0, p0), ptotal=p0.copy())
    tab = orc.Rosstab(deck["cnv_rt_zero"], deck["cnv_rt_tab"])
    out = orc.convec(tab, dict(sc, mixlth=0.0), a, e4, r4)
    assert np.allclose(out["grdadb"], 0.4, rtol=0, atol=1e-5)
    assert np.allclose(out["dlrdlt"], -1.0, rtol=0, atol=1e-5)
    assert np.all(out["flxcnv"] == 0.0)
