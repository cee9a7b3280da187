"""SYNTHE's record preparation (atlas12_amd.records, workloads.gfall_to_lte)
against the reference reader: the reference's run_mklinelist over a gfall
list, a predicted-format list and a MOL2 list on one 5 nm window
(tests/golden/records_ref.npz, written by oracle/ref/ref_synthe_dump) --
every lte_line_t field bit for bit, in the reference's order (gfall,
predict, mol).  Reference: src/mod_mklinelist.f90 read_gfall :780-1103,
read_predict :1110-1329, read_molec_bin :1371-1529, molec_dispatch
:2455-2586, load_mol_broad :414-485."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def rec(golden):
    return golden("records_ref.npz")


def _parts(rec):
    from atlas12_amd import records as R
    from atlas12_amd import workloads as wl
    grid = wl.Grid.make(*(float(x) for x in rec["grid"]))
    broad = R.load_mol_broad(list(rec["mol_broad"]))
    a = wl.gfall_to_lte(rec["gfall"], grid)
    b = R.predict_to_lte(rec["pred"], grid)
    c = R.mol2_to_lte(rec["mol2"].view(wl.MOL2_DTYPE), grid, broad)
    return a, b, c


def test_all_readers_match_reference_bitwise(rec):
    from atlas12_amd import workloads as wl
    a, b, c = _parts(rec)
    ref = rec["lines"].view(wl.LTE_DTYPE)
    assert len(a) > 0 and len(b) > 0 and len(c) > 0
    mine = np.concatenate([a, b, c])
    assert len(mine) == len(ref)
    for f in wl.LTE_DTYPE.names:
        assert np.array_equal(mine[f], ref[f]), f


def test_predict_species_map(rec):
    """NELIONOLD: the atoms' old index Z(Z+1)/2 + ion - 1 goes to 6 Z - 6 +
    ion for the ions the format keeps; the predicted list exercises mapped,
    unmapped (skipped) and molecular indices."""
    from atlas12_amd import records as R
    m = R.PREDICT_SPECIES
    for z, ion in ((1, 1), (2, 3), (26, 1), (26, 2), (20, 1), (99, 1), (31, 3)):
        old = z * (z + 1) // 2 + ion - 1 if z <= 30 else 496 + 5 * (z - 31) + ion - 1
        assert m[old] == 6 * z - 6 + ion, (z, ion)
    assert m[26 * 27 // 2 + 7 - 1] == 299 + 6 * (26 - 20)  # Fe VII's extra slot
    words = np.ascontiguousarray(rec["pred"])[:, 1:].copy().view(np.int16).reshape(-1, 6)
    idx = np.abs(words[:, 0].astype(np.int32)) // 10
    assert np.any(m[idx] == 0) and np.any(idx >= 841) and np.any((m[idx] != 0) & (idx < 841))


def test_mol2_covers_every_branch(rec):
    """The MOL2 list reaches FeH's 156 -> 126 renaming, unknown isotopes
    (skipped), classical and tabulated radiative damping, both label flags
    and the tabulated van der Waals widths of mol_broad.dat."""
    from atlas12_amd import records as R
    from atlas12_amd import workloads as wl
    m = rec["mol2"].view(wl.MOL2_DTYPE)
    assert np.any(m["code"] == 156) and np.any(m["gr100"] == 0) and np.any(m["flag"] == 1)
    known = np.array([R.mol_dispatch(int(i), int(c))[0] for i, c in zip(m["iso"], m["code"])])
    assert np.any(known == 0) and np.any(known != 0)
    broad = R.load_mol_broad(list(rec["mol_broad"]))
    assert 822 in broad and 606 in broad and broad[822] > 0.0


def test_mol_broad_population_weighting():
    """For a species with no J-dependence the weighted width is the J = 0
    width; with a slope it lies between the floor (0.1 g0) and g0, and a
    larger rotational constant (fewer populated J) keeps it closer to g0."""
    from atlas12_amd import records as R
    row = "  {c} X  0.1000 0.500 0.0600 0.500 {d} 0.0 {b} 1 X"
    flat = R.load_mol_broad([row.format(c=1, d=0.0, b=0.5)])[1]
    slow = R.load_mol_broad([row.format(c=2, d=0.002, b=0.5)])[2]
    fast = R.load_mol_broad([row.format(c=3, d=0.002, b=5.0)])[3]
    assert 0.1 * flat < slow < fast < flat


def test_diatomic_ascii_reader_matches_reference_bitwise(select_all_case):
    """read_diatomics_for_atlas (src/mod_mklinelist.f90:2244-2427) on seven
    generated Kurucz-format lists (CH, OH, C2, CN, CO, NH and AlO, the last
    outside ATLAS12's set): every field of every record equals the
    reference reader's, including terms listed in either order, negative
    (predicted) energies, classical damping, 'X' labels, blank and
    malformed lines and the early end past the window."""
    z, lists, _ = select_all_case
    ref, got = z["diatomic"], lists["dia"]
    assert len(ref) == len(got) > 10_000
    for f in ("wlvac_nm", "gflog_dex", "elo_cm", "gammar_log", "gammaw_log", "molcode"):
        assert np.array_equal(got[f], ref[f]), f
    assert 8470 in set(got["molcode"].tolist()) and 0 not in set(got["molcode"].tolist())


def test_diatomic_ascii_reader_rules():
    """The reader's per-line rules on hand-made lines: the window, the
    MOLCODE-0 skip, the J of the lower term for a tabulated width, the
    stop past wlend + 2 and Fortran F editing without a decimal point."""
    import math

    from atlas12_amd import records
    fmt = "{:10.4f}{:7.3f}{:5.1f}{:10.3f}{:5.1f}{:11.3f}{:4d}{:8s}{:8s}{:2d}{:4d}"
    lines = [fmt.format(500.0, -1.0, 3.5, 1000.0, 4.5, 21000.0, 106, "A", "X1", 12, 0),
             fmt.format(500.0, -1.0, 3.5, 21000.0, 4.5, 1000.0, 813, "A", "A1", 16, 800),
             fmt.format(600.0, -1.0, 3.5, 21000.0, 4.5, -1000.0, 106, "A", "A1", 12, 800),
             fmt.format(9000.0, -1.0, 3.5, 1000.0, 4.5, 1100.0, 106, "A", "A1", 12, 0),
             fmt.format(700.0, -1.0, 3.5, 1000.0, 4.5, 15285.714, 106, "A", "A1", 12, 0)]
    r = records.read_diatomic_ascii(lines, 400.0, 800.0)
    assert len(r) == 2 and list(r["molcode"]) == [8460, 8460]
    assert r["wlvac_nm"][0] == 1.0e7 / 20000.0 and r["elo_cm"][1] == 1000.0
    assert r["gammaw_log"][0] == math.log10(1.0e-8) and r["gammaw_log"][1] == math.log10(1.0e-7)
    assert r["gammar_log"][0] == math.log10(2.223e13 / 500.0 ** 2)
    assert r["gammar_log"][1] == math.log10(math.pow(10.0, 8.0))
    # x1 + x2 of 12CH (0, -0.005) added in double to the REAL(4) log gf
    assert r["gflog_dex"][0] == float(np.float32(-1.0)) + 0.0 + float(np.float32(-0.005))
    # a tabulated width: J of the lower term (E' = -1000 is the lower one here)
    broad = {106: (2.0e-8, 1.0e-10, 1.0e-8)}
    r2 = records.read_diatomic_ascii(lines[2:3], 400.0, 800.0, broad)
    assert r2["gammaw_log"][0] == math.log10(max(2.0e-8 - 1.0e-10 * 4.5, 0.1 * 2.0e-8))
    assert records._f_field("  5000", 4) == 0.5 and records._f_field("   ", 3) == 0.0
