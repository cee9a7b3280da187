"""The CPU oracle against the reference Fortran's own outputs (golden
fixtures written by tests/golden/make_golden.py from the amdflang build of
the reference): routine by routine, then the whole SYNTHE run."""
import numpy as np


def test_voigt_bitwise(orc, golden):
    z = golden("numerics_ref.npz")
    h = orc.voigt_ref(z["voigt_v"], z["voigt_a"])
    assert np.array_equal(h, z["voigt_h"])


def test_solvit_bitwise(orc, golden):
    z = golden("numerics_ref.npz")
    for a, b, x in zip(z["solvit_a"], z["solvit_b"], z["solvit_x"]):
        assert np.array_equal(orc.solvit(a, b), x)


def test_map1_bitwise(orc, golden):
    z = golden("numerics_ref.npz")
    for xo, fo, xn, fn, r in zip(z["map1_xold"], z["map1_fold"], z["map1_xnew"],
                                 z["map1_fnew"], z["map1_ret"]):
        out, ret = orc.map1(xo, fo, xn)
        assert ret == r
        assert np.array_equal(out, fn)


def test_integ_parcoe_deriv_bitwise(orc, golden):
    z = golden("numerics_ref.npz")
    for k in range(len(z["integ_x"])):
        x, f = z["integ_x"][k], z["integ_f"][k]
        assert np.array_equal(orc.integ(x, f, z["integ_start"][k]), z["integ_out"][k])
        assert np.array_equal(np.array(orc.parcoe(f, x)), z["parcoe_out"][k])
        assert np.array_equal(orc.deriv(x, f), z["deriv_out"][k])


def test_expi_bitwise(orc, golden):
    z = golden("numerics_ref.npz")
    e = np.array([orc.expi(n, x) for n, x in zip(z["expi_n"], z["expi_x"])])
    assert np.array_equal(e, z["expi_out"])


def test_josh_matches_reference(orc, golden, coef):
    """40 JOSH calls: IFSCAT 0/1, IFSURF 0/1, scattering and line opacity of
    several strengths, the deep-only branch (TAUNU(1) > 20)."""
    g = golden("josh_ref.npz")
    n = g["inputs"].shape[2]
    for k in range(len(g["ifscat"])):
        x = g["inputs"][k]
        r = orc.josh(int(g["ifscat"][k]), int(g["ifsurf"][k]), *x, *coef)
        o = g["out"][k]
        assert r["hnu1"] == o[0], k
        if g["ifsurf"][k] == 0:
            assert r["knu1"] == o[1]
            for i, name in enumerate(("taunu", "snu", "jnu", "hnu", "jmins")):
                assert np.array_equal(r[name], o[2 + i * n: 2 + (i + 1) * n]), (k, name)


def test_synthe_spectrum_every_printed_digit(orc, small_state, small_ref, coef):
    """Line opacity + transfer from the reference's own set-up state
    reproduce synthe.exe's .spec (F11.4, 2E15.6) at every printed digit."""
    op, ev = orc.synthe_line_opacity(small_state, small_ref["lines"], "ref")
    hl, hc = orc.synthe_transfer(small_state, op, *coef)
    spec = small_ref["spec"]
    wave = small_state.wbegin * small_state.ratio ** np.arange(small_state.length) * 10.0
    assert np.all(np.abs(np.round(wave, 4) - spec[:, 0]) <= 1.5e-4)
    e156 = lambda a: np.array([float("%.5E" % v) for v in a])  # noqa: E731
    assert np.array_equal(e156(hl), spec[:, 1])
    assert np.array_equal(e156(hc), spec[:, 2])


def _e124(a):
    """Values as synthe.exe prints them in .linform (1PE12.4)."""
    return np.array([float("%.4E" % v) for v in np.ravel(a)]).reshape(np.shape(a))


def test_synthe_opacity_every_depth_pins_reference_taunu(orc, small_state, small_ref, coef):
    """The reference's per-depth optical depths (synthe.exe more_output=yes:
    TAUNU of every depth at every wavelength point, .linform, src/synthe.f90:
    888-890) pin the line opacity matrix depth by depth, not only through
    the emergent flux: the oracle's opacity (reference arithmetic) through
    the oracle's transfer reproduces all 4,975 x 72 printed values.  A
    1e-3 scaling of the opacity or one depth's opacity dropped changes them
    (planted in the oracle's output)."""
    op, _ = orc.synthe_line_opacity(small_state, small_ref["lines"], "ref")
    ref = small_ref["taunu"]
    assert ref.shape == (small_state.length, small_state.nrhox)
    assert np.array_equal(_e124(orc.synthe_taunu(small_state, op, *coef)), ref)
    scaled = (op * np.float32(1.001)).astype(np.float32)
    assert np.mean(_e124(orc.synthe_taunu(small_state, scaled, *coef)) != ref) > 0.9
    dropped = op.copy()
    dropped[30] = 0.0
    bad = _e124(orc.synthe_taunu(small_state, dropped, *coef)) != ref
    # TAUNU integrates from the surface with parabolic coefficients that
    # reach one depth ahead: depths above 29 are untouched
    assert bad[:, 30:].mean() > 0.5 and not bad[:, :29].any()


def test_record_prep_bitwise(small_ref):
    """gfall -> lte_line_t record preparation (read_gfall restated in
    atlas12_amd.workloads) equals the reference reader's records."""
    from atlas12_amd import workloads as wl
    mine = wl.gfall_to_lte(small_ref["gfall"], wl.Grid.make(500.0, 505.0, 500000.0))
    ref = small_ref["lines"]
    for f in wl.LTE_DTYPE.names:
        assert np.array_equal(mine[f], ref[f]), f


def test_powf10_matches_libm():
    import ctypes
    from atlas12_amd.workloads import powf10
    libm = ctypes.CDLL("libm.so.6")
    libm.powf.restype = ctypes.c_float
    libm.powf.argtypes = [ctypes.c_float, ctypes.c_float]
    x = np.round(np.random.default_rng(0).uniform(-8, 2, 40000), 3).astype(np.float32)
    ref = np.array([libm.powf(10.0, float(v)) for v in x], np.float32)
    assert np.array_equal(powf10(x), ref)


# ---------------------------------------------------------------- ATLAS12
def test_tabvoigt_bitwise(orc, golden):
    z = golden("atlas_small.npz")
    for a, k in zip(orc.tabvoigt(), ("h0", "h1", "h2")):
        assert np.array_equal(a, z[k])


def test_linop1_bitwise(orc, golden):
    """LINOP1's XLINES from the reference's first-iteration state and a
    3000-line store, coarse/fine depth passes and both Voigt regimes."""
    z = golden("atlas_small.npz")
    xl, ev = orc.linop1(z)
    assert ev > 0
    assert np.array_equal(xl, z["xlines"])


def test_rt_pass_accumulators_bitwise(orc, golden, coef):
    """The frequency loop of RUN_RT_PASS from tabulated KAPP outputs
    reproduces the reference's own accumulators (TCORR/RADIAP/ROSS modes
    1-2) and the last frequency's flux bit for bit."""
    z = golden("atlas_small.npz")
    acc, h1 = orc.rt_pass(z, z["xlines"], *coef)
    for k in ("rjmins", "flxrad", "raden", "accrad", "abross"):
        assert np.array_equal(acc[k], z["acc_" + k]), k
    assert acc["pradk0"] == z["acc_pradk0"]
    assert h1[-1] == z["last_hnu"][0]
    # the last frequency's per-depth terms the reference leaves in its module
    # arrays: TCORR(2)'s RDIAGJNU (RDIAGJ's summand, :2507-2510; RDIAGJ and
    # RDABH themselves are local to TCORR), HNU and TAUNU
    for k in ("rdiagjnu", "hnu", "taunu"):
        assert np.array_equal(acc["last_" + k], z["last_" + k]), k
    # H_TOTAL is FLXRAD's twin (the same sum in RADIAP)
    assert np.array_equal(acc["h_total"], acc["flxrad"])



def test_selectlines_bitwise(orc, select_case):
    """SELECTLINES on 72,000 predicted (atomic + diatomic) and 8,000 TiO
    MOL2 records keeps exactly the reference's store, record for record,
    including the TiO repacking (IWL, IELO, IGFLOG + isotope, IGR)."""
    z, pred, tio = select_case
    sel = orc.atlas_select(z, pred, tio)
    assert sel.shape == z["selected"].shape
    assert np.array_equal(sel, z["selected"])


def test_selectlines_planted_mistake_is_caught(orc, select_case):
    """A dropped Boltzmann factor (HCKT(NRHOX) = 0) keeps lines the
    reference rejects: the fixture distinguishes it."""
    z, pred, tio = select_case
    zz = dict(z)
    zz["hckt"] = np.array(z["hckt"]) * 0.0
    assert len(orc.atlas_select(zz, pred, tio)) > len(z["selected"])


def _path_of(store):
    """Source list of each kept record, from its species field: H2O
    -9400..-9403, TiO 8950..8954, diatomic MOLCODEs 8410..8960, H3+ (kept
    as read) 8950 -- the atomic paths are told apart by position only."""
    w = store[:, 1:].copy().view(np.int16).reshape(len(store), 6)
    return w[:, 0].astype(int)


def test_selectlines_all_lists_bitwise(orc, select_all_case):
    """SELECTLINES over all seven lists (predicted, observed low lines,
    hilines, ASCII diatomics, TiO, H2O, H3+: 82,000 raw records) keeps the
    reference's store record for record.  The fixture exercises the
    observed-low-line walk continuing from the predicted one, the empty
    species slot of hilines (XNFDOPMAX = 0), 732 diatomic wavelength keys
    shared by several records (whose order only the reference's heapsort
    gives), H2O isotopes from the signs and kept lines with E_lo = 0."""
    z, lists, _ = select_all_case
    sel = orc.atlas_select_all(z, lists)
    assert sel.shape == z["selected"].shape
    assert np.array_equal(sel, z["selected"])
    ie = _path_of(sel)
    assert np.sum((ie <= -9400) & (ie >= -9403)) > 1000     # H2O kept
    assert np.sum((ie >= 8410) & (ie <= 8960) & (ie != 8950)) > 1000   # diatomics kept
    w = sel[:, 1:].copy().view(np.int16).reshape(len(sel), 6)
    assert np.sum((ie <= -9400) & (w[:, 1] == 0)) > 0        # E_lo = 0 H2O lines


def test_selectlines_all_lists_planted_mistakes_are_caught(orc, select_all_case):
    """Mistakes a restatement could make, each planted in the oracle's
    inputs, change the store the reference fixture holds: the walk of the
    observed low lines restarted (dropped carry-over), the diatomic keys
    sorted stably instead of by the heapsort, and the H3+ log gf correction
    dropped."""
    z, lists, _ = select_all_case
    ref = z["selected"]
    # (a) restart the observed-low-line walk: run it as its own list after an
    #     empty predicted list -- a different store from the reference's
    alt = dict(lists, pred=lists["pred"][:0])
    alone = orc.atlas_select_all(z, alt)
    with_pred = orc.atlas_select_all(z, dict(lists))
    n_pred_kept = len(orc.atlas_select_all(z, {"pred": lists["pred"]}))
    assert not np.array_equal(alone, with_pred[n_pred_kept:])
    # (b) a stable sort of the diatomic records changes the order of ties
    keys = np.array([int(round(np.log(x) / np.log(1.0 + 1.0 / 2.0e6))) for x in lists["dia"]["wlvac_nm"]])
    _, heap_order = orc.sort_indirect_by_int(keys)
    stable = np.argsort(keys, kind="stable")
    assert not np.array_equal(heap_order - 1, stable)
    resorted = orc.atlas_select_all(z, dict(lists, dia=lists["dia"][heap_order - 1]))
    assert not np.array_equal(resorted, ref)   # re-sorting a sorted list is not idempotent
    # (c) H3+ without its -1.272 dex: more H3+ lines kept
    h3 = lists["h3p"].copy()
    w = h3[:, 1:].copy().view(np.int16).reshape(len(h3), 6)
    w[:, 2] = np.minimum(w[:, 2].astype(int) + 1272, 32767).astype(np.int16)
    h3[:, 1:] = w.view(np.int32).reshape(len(h3), 3)
    assert len(orc.atlas_select_all(z, dict(lists, h3p=h3))) > len(ref)


def test_heapsort_properties(orc):
    """sort_indirect_by_int: keys non-decreasing, order a permutation that
    carries the keys, and on distinct keys the same result as any sort."""
    rng = np.random.default_rng(3)
    for n in (0, 1, 2, 7, 100, 5000):
        keys = rng.integers(0, max(n // 4, 1), n).astype(np.int32)
        ks, order = orc.sort_indirect_by_int(keys)
        assert np.all(np.diff(ks) >= 0)
        assert np.array_equal(np.sort(order), np.arange(1, n + 1))
        assert np.array_equal(keys[order - 1], ks)
        distinct = rng.permutation(n).astype(np.int32)
        _, o2 = orc.sort_indirect_by_int(distinct)
        assert np.array_equal(o2 - 1, np.argsort(distinct))


def test_scale_model_bitwise(orc, golden):
    """SCALE_MODEL after READIN: regrid only, an unchanged (Teff, log g) and
    three rescalings reproduce the reference's TAUSTD, remapped structure,
    PRADK/PTOTAL and scalars bit for bit."""
    z = golden("scale_model.npz")
    for k in range(len(z["params"])):
        par, inp = z["params"][k], z["inp"][k]
        n = inp.shape[1]
        out, so, resc = orc.scale_model(n, 80, par[5], par[6], par[:5], par[7], par[8], inp)
        assert np.array_equal(out[0], z["taustd"][k]), k
        assert np.array_equal(out[1:9], z["out"][k][:8]), k
        if resc:
            assert np.array_equal(out[9], z["out"][k][8]), k
            assert np.array_equal(so, z["scal"][k]), k
        else:
            assert np.array_equal(so[[0, 1, 4]], z["scal"][k][[0, 1, 4]]), k
    # the regrid-only and unchanged cases do not rescale; the others do
    assert [orc.scale_model(z["inp"][k].shape[1], 80, *z["params"][k][5:7], z["params"][k][:5],
                            *z["params"][k][7:9], z["inp"][k])[2] for k in range(5)] == \
        [False, False, True, True, True]


def _printed(v, fmt):
    return np.array([float(fmt % x) for x in v])


def test_tcorr3_matches_reference_table(orc, golden):
    """TCORR(3) after a full 30,000-frequency pass: DTLAMB, DTSURF, DTFLUX
    and T1 equal the reference's per-iteration table at every printed digit
    (F8.1), ERROR and DERIV too (ES11.2), CONV/TOT (F8.3).  The convective
    branch is not reached in this model (CONVEC gives no convective flux)."""
    z = golden("tcorr_ref.npz")
    out, _, _ = orc.tcorr3(z, z["rdabh"])
    for k in ("dtlamb", "dtsurf", "dtflux", "t1"):
        assert np.array_equal(_printed(out[k], "%.1f") + 0.0, z["iter_" + k] + 0.0), k
    for k in ("flxerr", "flxdrv"):
        assert np.array_equal(_printed(out[k], "%.2E"), z["iter_" + k]), k
    assert np.array_equal(_printed(out["hratio"], "%.3f"), z["iter_hratio"])


def test_tcorr3_convective_matches_reference_table(orc, golden):
    """TCORR(3) on a deck whose deep layers convect (CONVEC's flux at 20
    layers, CONV/TOT up to 0.59): the convective branches -- the DDEL
    factor, the convective gradient terms, the DTLAMB halving of the five
    layers above a convective one, no damping where CONV/TOT > 0 -- give the
    reference's printed DTLAMB, DTSURF, DTFLUX, T1 (F8.1), ERROR, DERIV and
    CONV/TOT (ES11.2) at every printed digit."""
    z = golden("tcorr_conv_ref.npz")
    assert int(z["ifconv"]) == 1 and np.sum(z["flxcnv"] > 0) >= 10
    assert np.max(z["iter_hratio"]) > 0.5
    out, _, _ = orc.tcorr3(z, z["rdabh"])
    for k in ("dtlamb", "dtsurf", "dtflux", "t1"):
        assert np.array_equal(_printed(out[k], "%.1f") + 0.0, z["iter_" + k] + 0.0), k
    for k in ("flxerr", "flxdrv", "hratio"):
        assert np.array_equal(_printed(out[k], "%.2E"), z["iter_" + k]), k


def test_tcorr3_damping_rules(orc, golden):
    """Iteration 2 against a planted previous correction: same sign and
    shrinking -> x1.25, sign flip -> x0.5, same sign growing -> unchanged
    (src/atlas12_modules.f90:2735-2751); the undamped sum otherwise equals
    iteration 1's."""
    z = golden("tcorr_ref.npz")
    base, _, _ = orc.tcorr3(z, z["rdabh"])
    t1 = base["t1"]
    old = np.where(np.arange(len(t1)) % 3 == 0, 2.0 * t1, np.where(np.arange(len(t1)) % 3 == 1,
                   -t1, 0.5 * t1))
    zz = dict(z)
    zz["ifconv"] = 0  # damping at every depth
    out, _, new_old = orc.tcorr3(zz, z["rdabh"], oldt1=old, iter_=2)
    k = np.arange(len(t1)) % 3
    nz = t1 != 0
    assert np.array_equal(out["t1"][(k == 0) & nz], t1[(k == 0) & nz] * 1.25)
    assert np.array_equal(out["t1"][(k == 1) & nz], t1[(k == 1) & nz] * 0.5)
    assert np.array_equal(out["t1"][(k == 2) & nz], t1[(k == 2) & nz])
    assert np.array_equal(new_old, out["t1"])
