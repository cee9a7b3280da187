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


@pytest.fixture(scope="module")
def recm(golden):
    """tests/golden/records_mol_ref.npz with its raw lists regenerated."""
    import hashlib
    import json

    from atlas12_amd import workloads as wl
    z = golden("records_mol_ref.npz")
    c = json.loads(str(z["config"]))
    g, mol, poly, pbin, h2o = wl.synthe_mol_lists(**{k: c[k] for k in wl.SYNTHE_MOL_KEYS})
    digest = hashlib.sha256(g.tobytes() + "".join(mol.values()).encode() + poly.encode() +
                            pbin.tobytes() + h2o.tobytes()).hexdigest()
    assert digest == str(z["raw_sha256"]), "line-list generator drifted from the fixture"
    return z, c, (g, mol, poly, pbin, h2o)


def test_molecular_readers_match_reference_bitwise(recm):
    """run_mklinelist's molecular ASCII (read_molec_ascii :1535-1689, FeH
    156 -> 126, AlO, unknown isotopes skipped), polyatomic ASCII
    (read_molec_poly :1690-1838, 6-digit codes), polyatomic binary
    (read_molec_poly_bin :1864-2012) and H2O (read_h2o :2034-2174) readers
    at TEFF 3400 K: every lte_line_t field of every record equals the
    reference's, in its order (gfall, mol files, polymol files, H2O)."""
    from atlas12_amd import records as R
    from atlas12_amd import workloads as wl
    z, c, (g, mol, poly, pbin, h2o) = recm
    grid = wl.Grid.make(*c["window"])
    table = R.load_mol_broad_table(list(z["mol_broad"]))
    gwef = {k: v[2] for k, v in table.items()}
    parts = [wl.gfall_to_lte(g, grid)]
    parts += [R.molec_ascii_to_lte(t.split("\n"), grid, table) for t in mol.values()]
    parts += [R.molec_poly_to_lte(poly.split("\n"), grid, table),
              R.molec_poly_bin_to_lte(pbin, grid, gwef), R.h2o_to_lte(h2o, grid, gwef)]
    assert all(len(p) > 100 for p in parts)
    mine = np.concatenate(parts)
    ref = z["lines"].view(wl.LTE_DTYPE)
    assert len(mine) == len(ref)
    for f in wl.LTE_DTYPE.names:
        assert np.array_equal(mine[f], ref[f]), f
    codes = set(np.concatenate(parts[1:-3])["code"].astype(int).tolist())
    assert 126 in codes and 156 not in codes and 813 in codes   # FeH renamed, AlO kept
    # H2O takes the tabulated population-weighted width of code 10108
    assert gwef[10108] > 0.0 and set(parts[-1]["nelion"].tolist()) == {534}


def test_molecular_readers_window_rules():
    """The binary readers' quirks on hand-made lists: read_h2o starts at the
    bisection's last probe with no lower bound after it (a record below
    wlbeg - 1 can be read), read_molec_poly_bin skips those; nothing is
    read when the whole list lies outside the window."""
    import math

    from atlas12_amd import records as R
    from atlas12_amd import workloads as wl
    grid = wl.Grid.make(500.0, 505.0, 500000.0)
    lam = np.linspace(480.0, 520.0, 64)
    iwl = np.rint(np.log(lam) / math.log(1.0 + 1.0 / 2.0e6)).astype(np.int32)
    h = np.zeros(64, wl.H2O_DTYPE)
    h["iwl"], h["ielo"], h["igflog"] = iwl, 100, 15000
    got = R.h2o_to_lte(h, grid)
    lo_wl = np.exp(iwl.astype(float) * math.log(1.0 + 1.0 / 2.0e6))
    inside = np.nonzero((lo_wl >= 499.0) & (lo_wl <= 506.0))[0]
    assert len(got) >= len(inside)
    assert len(got) - len(inside) <= 1           # at most the bisection's last probe below
    words = np.zeros((65, 2), np.int32)
    words[0] = (-1, 10820)
    words[1:, 0] = iwl
    words[1:, 1] = (15000 << 16) | 100
    assert len(R.molec_poly_bin_to_lte(words, grid)) == len(inside)
    words[0, 1] = 10821                           # no species slot: nothing read
    assert len(R.molec_poly_bin_to_lte(words, grid)) == 0
    far = wl.Grid.make(900.0, 905.0, 500000.0)
    assert len(R.h2o_to_lte(h, far)) == 0 and len(R.molec_poly_bin_to_lte(words, far)) == 0
