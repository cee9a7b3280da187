"""The device end of TCORR(3) through the C-ABI -- a12_rosstab_*, a12_ttaup,
a12_cz_construct and a12_atlas_tcorr3_finish -- against the reference's
outputs (tests/golden/tcorr_ref.npz, tcorr_conv_ref.npz) and the oracle.

Tolerances: the device's log10, exp, log and pow may differ from libm's in
the last bit, so the table, the lookups and everything downstream agree to
a few ulps rather than bit for bit; discrete decisions (quadrant choice,
TTAUP's convergence count, CZ_CONSTRUCT's gates and block) agree exactly.
Reference: ROSSTAB src/atlas12_modules.f90:5937-6268, TTAUP :6303-6523,
CZ_CONSTRUCT :3012-3358, TCORR(3) :2816-2967."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
DECKS = ("tcorr_ref", "tcorr_conv_ref")


@pytest.fixture(scope="module", params=DECKS)
def deck(request, golden):
    return golden(request.param + ".npz")


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))


def _taustd(z):
    return np.array([math.pow(10.0, float(z["fin_tau1lg"]) + j * float(z["fin_steplg"]))
                     for j in range(int(z["n"]))])


def test_rosstab_store_on_device(gpu_ctx, deck):
    from atlas12_amd import native
    native.rosstab_reset(gpu_ctx)
    native.rosstab_store(gpu_ctx, deck["t"], deck["fin_in_p"], deck["abross"])
    z, tab = native.rosstab_fetch(gpu_ctx)
    assert tab.shape == deck["fin_rt_tab"].shape
    assert np.max(np.abs(z - deck["fin_rt_zero"])) <= 4e-16 * np.max(np.abs(deck["fin_rt_zero"]))
    assert np.max(np.abs(tab - deck["fin_rt_tab"])) <= 1e-14


# Tolerance per IROSSTAB mode: the oracle's lookups move by 2e-14 (modes 1
# and 2) and by 2e-8 (mode 3: the moving-least-squares normal equations,
# ridge 1e-6, are ill-conditioned on a table of two profiles) when the
# queries move by one ulp -- the size of the device log10's last-bit
# differences in the normalised coordinates.
LOOKUP_TOL = {1: 1e-12, 2: 1e-12, 3: 1e-7}


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_rosstab_lookup_on_device(gpu_ctx, deck, mode):
    from atlas12_amd import native
    native.rosstab_load(gpu_ctx, deck["fin_rt_zero"], deck["fin_lk_tab"])
    q = deck["fin_lk_q"]
    got = native.rosstab_lookup(gpu_ctx, q[0], q[1], mode)
    assert _rel(got, deck["fin_lk_r"][mode - 1]) <= LOOKUP_TOL[mode]


def test_rosstab_capacity_and_reset(gpu_ctx, deck):
    """Stores stop at ROSSTAB_MAXTAB (4,800) entries as the reference's do;
    a reset empties the table and lookups then refuse."""
    from atlas12_amd import native
    native.rosstab_reset(gpu_ctx)
    n = int(deck["n"])
    for _ in range(native.ROSSTAB_MAXTAB // n + 2):
        native.rosstab_store(gpu_ctx, deck["t"], deck["fin_in_p"], deck["abross"])
    _, tab = native.rosstab_fetch(gpu_ctx)
    assert tab.shape[1] == native.ROSSTAB_MAXTAB
    assert np.array_equal(tab[:, n:2 * n], tab[:, :n])
    native.rosstab_reset(gpu_ctx)
    with pytest.raises(native.A12Error):
        native.rosstab_lookup(gpu_ctx, [5000.0], [1e4], 1)
    native.rosstab_load(gpu_ctx, deck["fin_rt_zero"], deck["fin_rt_tab"])
    with pytest.raises(native.A12Error):
        native.rosstab_lookup(gpu_ctx, [5000.0], [1e4], 4)


def test_ttaup_on_device(gpu_ctx, deck):
    from atlas12_amd import native
    native.rosstab_load(gpu_ctx, deck["fin_rt_zero"], deck["fin_rt_tab"])
    ab, pt, pg, nf = native.ttaup(gpu_ctx, deck["fin_tt_t"], _taustd(deck), deck["fin_tt_prad"],
                                  deck["fin_in_pturb"], float(deck["fin_grav"]))
    assert nf == 0
    assert _rel(ab, deck["fin_tt_ab"]) <= 1e-12
    assert _rel(pt, deck["fin_tt_ptot"]) <= 1e-12
    assert _rel(pg, deck["fin_tt_p"]) <= 1e-12


def test_cz_construct_on_device(gpu_ctx, deck):
    """The planted engage / stay / release sequence: the same engagement and
    block top on every call, T1 within 1e-9 K of the reference's."""
    from atlas12_amd import native
    native.rosstab_load(gpu_ctx, deck["fin_rt_zero"], deck["fin_lk_tab"])
    z = {k: deck["fin_cz_" + k] for k in ("flxcnv", "flxrad", "dltdlp", "grdadb", "t", "rhox",
                                           "ptotal", "rho", "heatcp", "dlrdlt", "p", "vturb",
                                           "abross", "hscale")}
    z.update(flux=float(deck["fin_cz_flux"]), grav=float(deck["fin_cz_grav"]),
             mixlth=float(deck["fin_cz_mixlth"]))
    st = native.czc_state()
    n = int(deck["n"])
    for k in range(3):
        z["flxerr"] = deck["fin_cz_flxerr"][k]
        t1, e = native.cz_construct(gpu_ctx, z, np.zeros(n), st, k + 1,
                                    use4pt=int(deck["fin_cz_4pt"]))
        # e: this call changed T1; run_engaged: the run-level module flag
        assert e == int(np.any(deck["fin_cz_t1"][k] != 0.0))
        assert st.run_engaged == int(deck["fin_cz_engaged"][k])
        assert st.run_jc == int(deck["fin_cz_jc"][k])
        assert np.max(np.abs(t1 - deck["fin_cz_t1"][k])) <= 1e-9


def _finish_inputs(z, t1, flxerr):
    n = int(z["n"])
    sc = dict(ifconv=int(z["ifconv"]), iter=int(z["iter"]), j1smooth=int(z["fin_j1smooth"]),
              j2smooth=int(z["fin_j2smooth"]), flux=float(z["flux"]), mixlth=float(z["mixlth"]),
              grav=float(z["fin_grav"]), tau1lg=float(z["fin_tau1lg"]),
              steplg=float(z["fin_steplg"]), pradk0=float(z["fin_pradk0"]),
              tfloor=float(z["fin_tfloor"]), wtjm1=float(z["fin_wtjm1"]), wtj=float(z["fin_wtj"]),
              wtjp1=float(z["fin_wtjp1"]))
    arr = {k: z[k] for k in ("flxcnv", "flxrad", "dltdlp", "grdadb", "ptotal", "rho", "heatcp",
                             "dlrdlt", "hscale", "tauros", "rhox", "t", "abross")}
    arr["flxerr"] = flxerr
    for k in ("p", "xne", "prad", "pturb", "vturb", "bmin", "accrad"):
        arr[k] = z["fin_in_" + k]
    arr["bhyd"] = z["fin_in_bhyd"]
    assert len(t1) == n
    return sc, arr


FIN_KEYS = ("rhox", "t", "tk", "hkt", "hckt", "tkev", "tlog", "p", "xne", "abross", "prad",
            "pradk", "vturb", "bmin", "pturb", "tauros")
# DRHOX comes from PPP = (PTOT2 - PTOT1) / PTOT1, the small difference of
# two TTAUP marches: one-ulp changes of T1 and of the ROSSTAB entries move
# the oracle's DRHOX by up to 2.3e-11 of its largest value and RHOX (after
# RHOX + DRHOX and the remap) by up to 3.7e-9 (measured on both decks);
# everything else moves by a few ulps.
FIN_TOL = dict({k: 1e-11 for k in FIN_KEYS}, rhox=1e-7)
DRHOX_TOL = 1e-9


def test_tcorr3_on_device_matches_reference_state(gpu_ctx, deck):
    """Sections A-F (a12_atlas_tcorr3) and (G2)-(K) (a12_atlas_tcorr3_finish)
    on the device, the ROSSTAB table stored on the device: the module state
    the reference's TCORR(3) leaves, every array within 1e-11."""
    from atlas12_amd import native
    native.rosstab_reset(gpu_ctx)
    native.rosstab_store(gpu_ctx, deck["t"], deck["fin_in_p"], deck["abross"])
    out, _, _ = native.atlas_tcorr3(gpu_ctx, deck, rdabh=deck["rdabh"])
    sc, arr = _finish_inputs(deck, out["t1"], out["flxerr"])
    st = native.czc_state()
    got, _, drhox, nf = native.atlas_tcorr3_finish(gpu_ctx, sc, arr, out["t1"], st)
    for k in FIN_KEYS:
        assert _rel(got[k], deck["fin_out_" + k]) <= FIN_TOL[k], k
    assert _rel(got["bhyd"], deck["fin_out_bhyd"]) <= 1e-11
    assert st.run_engaged == int(deck["fin_out_czc_engaged"])


def test_tcorr3_finish_matches_oracle(gpu_ctx, deck, orc):
    """The same inputs to the oracle and the device (the oracle's T1): every
    output within FIN_TOL, DRHOX within DRHOX_TOL of its largest value, and
    the same count of TTAUP depths left unconverged."""
    from atlas12_amd import native
    out, _, _ = orc.tcorr3(deck, deck["rdabh"])
    sc, arr = _finish_inputs(deck, out["t1"], out["flxerr"])
    native.rosstab_load(gpu_ctx, deck["fin_rt_zero"], deck["fin_rt_tab"])
    got, t1g, drg, nfg = native.atlas_tcorr3_finish(gpu_ctx, sc, arr, out["t1"],
                                                     native.czc_state())
    tab = orc.Rosstab(deck["fin_rt_zero"], deck["fin_rt_tab"])
    sco = dict(sc, n=int(deck["n"]), use4pt=1, use_constructor=1)
    want, t1o, dro, nfo = orc.tcorr3_finish(tab, sco, arr, out["t1"], orc.czc_state())
    assert nfg == nfo
    for k in FIN_KEYS:
        assert _rel(got[k], want[k]) <= FIN_TOL[k], k
    assert _rel(got["bhyd"], want["bhyd"]) <= 1e-11
    assert np.max(np.abs(drg - dro)) <= DRHOX_TOL * np.max(np.abs(dro))
    assert np.array_equal(t1g, t1o)


def test_tcorr3_finish_rejects_bad_input(gpu_ctx, deck):
    """Calls that the reference would not make return an error before any
    kernel runs: no table, a smoothing range that reads outside the model,
    arrays of the wrong shape."""
    from atlas12_amd import native
    out = {"t1": np.zeros(int(deck["n"]))}
    sc, arr = _finish_inputs(deck, out["t1"], np.zeros(int(deck["n"])))
    native.rosstab_reset(gpu_ctx)
    with pytest.raises(native.A12Error):
        native.atlas_tcorr3_finish(gpu_ctx, sc, arr, out["t1"], native.czc_state())
    native.rosstab_load(gpu_ctx, deck["fin_rt_zero"], deck["fin_rt_tab"])
    with pytest.raises(native.A12Error):
        native.atlas_tcorr3_finish(gpu_ctx, dict(sc, j1smooth=1, j2smooth=5), arr, out["t1"],
                                   native.czc_state())
    with pytest.raises(ValueError):
        native.atlas_tcorr3_finish(gpu_ctx, sc, dict(arr, p=arr["p"][:-1]), out["t1"],
                                   native.czc_state())


# CONVEC's bubble-opacity iteration stops once a step is below 0.5 K (or
# 1e-3 relative) and takes Steffensen and cycle-mean exits on tests of its
# iterates: one-ulp changes of its inputs (the EOS energies, the ROSSTAB
# entries) move the oracle's converged DELTAT by up to 1.8 % and FLXCNV by up
# to 1.2 % of its largest value on the convective deck (measured).  The
# device's log10 and pow differ from libm in the last bit, so the iterated
# quantities are held to that band (5 %); the thermodynamic derivatives,
# which nothing iterates, to 1e-11.
CNV_DIRECT = ("dltdlp", "heatcp", "velsnd", "dlrdlt", "grdadb", "hscale")
CNV_ITER = ("vconv", "flxcnv", "flxcnv0", "flxcnv1")


def _convec_inputs(z, overwt=None):
    g = lambda k: z["cnv_" + k]  # noqa: E731
    sc = dict(ifconv=1, ifcorr=int(g("ifcorr")), nconv=int(g("nconv")), use4pt=int(g("use4pt")),
              smooth=int(g("smooth")), grav=float(g("grav")), mixlth=float(g("mixlth")),
              flux=float(g("flux")), overwt=float(g("overwt") if overwt is None else overwt))
    arr = {k: g("in_" + k) for k in ("t", "tk", "hkt", "hckt", "tkev", "tlog", "p", "ptotal",
                                     "rho", "pradk", "tauros", "rhox", "abross", "vturb",
                                     "height", "flxrad", "flxcnv1")}
    return sc, arr, g("edens"), g("rho4")


@pytest.mark.parametrize("overwt", [None, 0.5])
def test_convec_on_device_matches_reference(gpu_ctx, deck, overwt):
    from atlas12_amd import native
    sc, arr, e, r = _convec_inputs(deck, overwt)
    native.rosstab_load(gpu_ctx, deck["cnv_rt_zero"], deck["cnv_rt_tab"])
    out = native.convec(gpu_ctx, sc, arr, e, r)
    for k in ("t", "tk", "hkt", "hckt", "tkev", "tlog", "p"):
        assert np.array_equal(out[k], deck["cnv_st_" + k]), k
    for k in CNV_DIRECT:
        assert _rel(out[k], deck["cnv_out_" + k]) <= 1e-11, k
    pre = "cnv_ovr_" if overwt else "cnv_out_"
    for k in CNV_ITER:
        want = deck[pre + k] if pre + k in deck.files else deck["cnv_out_" + k]
        scale = max(np.max(np.abs(want)), 1e-300)
        assert np.max(np.abs(out[k] - want)) <= 0.05 * scale, k
    # the same layers convect (overshooting adds flux near zero at the
    # zone's edges, where the sign of a tiny value is not held)
    assert np.array_equal(out["vconv"] > 0.0, deck["cnv_out_vconv"] > 0.0)
    if not overwt:
        assert np.array_equal(out["flxcnv"] > 0.0, deck["cnv_out_flxcnv"] > 0.0)


def test_convec_mlt_only_leaves_history(gpu_ctx, deck):
    """mlt_only (the Newton probe's call): the MLT solve without the
    post-processing -- no smoothing, FLXCNV0 and FLXCNV1 untouched."""
    from atlas12_amd import native
    sc, arr, e, r = _convec_inputs(deck)
    native.rosstab_load(gpu_ctx, deck["cnv_rt_zero"], deck["cnv_rt_tab"])
    arr = dict(arr, flxcnv1=np.full(len(arr["t"]), 7.0))
    full = native.convec(gpu_ctx, sc, arr, e, r)
    raw = native.convec(gpu_ctx, sc, arr, e, r, mlt_only=True)
    assert np.all(raw["flxcnv1"] == 7.0)
    assert np.all(raw["flxcnv0"] == 0.0)
    for k in CNV_DIRECT:
        assert np.array_equal(raw[k], full[k]), k
    assert np.all(full["flxcnv"][:int(sc["nconv"])] == 0.0)
