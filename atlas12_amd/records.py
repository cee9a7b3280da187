"""Line-record preparation on the host (src/mod_mklinelist.f90): the
readers that turn on-disk line lists into the records the two hot paths
read -- SYNTHE's ``lte_line_t`` (LTE_DTYPE, the BIND(C) mirror of
a12_lte_line) and ATLAS12's diatomic_record_t (DIATOMIC_DTYPE, the mirror
of a12_diatomic_rec, which a12_atlas_select_all filters).  The gfall reader
is :func:`atlas12_amd.workloads.gfall_to_lte`.

* :func:`predict_to_lte`        read_predict        (:1110-1329), 0.01502;
* :func:`mol2_to_lte`           read_molec_bin      (:1371-1529) with
                                molec_dispatch      (:2455-2586);
* :func:`molec_ascii_to_lte`    read_molec_ascii    (:1535-1689);
* :func:`molec_poly_to_lte`     read_molec_poly     (:1690-1838);
* :func:`molec_poly_bin_to_lte` read_molec_poly_bin (:1864-2012);
* :func:`h2o_to_lte`            read_h2o            (:2034-2174);
* :func:`read_diatomics_for_atlas` / :func:`read_diatomic_ascii`
                                (:2244-2427), ATLAS12's diatomic lists;
* :func:`load_mol_broad_table`  load_mol_broad      (:414-485) and
                                molbroad_gw_elo     (:396-411).

Expression types follow the reference (REAL(4) where the reference computes
or stores REAL(4)); the records equal the reference readers' bit for bit
(tests/test_records.py)."""
from __future__ import annotations

import math

import numpy as np

from .workloads import LTE_DTYPE, Grid, powf10

# ---------------------------------------------------------------------------
# predicted lines (16-byte records: int32 IWL, int16 x 6)
# ---------------------------------------------------------------------------


def _predict_species_map() -> np.ndarray:
    """read_predict's NELIONOLD (:1125-1227): old species index (1..1005,
    |iwords(3)| / 10) -> SYNTHE's nelion slot (6 Z - 6 + ion), 0 = skipped.

    Built from its structure: old indices count Z(Z+1)/2 + ion - 1 for the
    atoms Z = 1..30 (ions up to 6 map to 6 Z - 6 + ion -- up to 5 for Cl,
    Ar, K; Ca..Ni keep ions 7..10 in four extra slots 299 + 6 (Z - 20) +
    60 (ion - 7); Cu and Zn only ions 1..3), five per element from Z = 31
    to 99 (ions 1..3), and the molecules from 841 on (an explicit list)."""
    m = np.zeros(1006, np.int32)  # 1-based
    for z in range(1, 31):
        base = z * (z + 1) // 2
        for ion in range(1, z + 2):
            old = base + ion - 1
            new = 0
            if z <= 19 or z >= 29:
                top = 3 if z >= 29 else (5 if z >= 17 else 6)
                if ion <= top:
                    new = 6 * z - 6 + ion
            else:  # Ca .. Ni
                if ion <= 6:
                    new = 6 * z - 6 + ion
                elif ion <= 10:
                    new = 299 + 6 * (z - 20) + 60 * (ion - 7)
            m[old] = new
    for z in range(31, 100):
        for ion in range(1, 4):
            m[496 + 5 * (z - 31) + ion - 1] = 6 * z - 6 + ion
    m[841:1006] = _PREDICT_MOLECULES
    return m


# NELIONOLD(841:1005), the molecules of the old predicted format
# (src/mod_mklinelist.f90:1212-1227): a table the format defines
_PREDICT_MOLECULES = np.array([
    240, 0, 378, 384, 390, 246, 252, 258, 396, 0, 300, 306, 312, 402, 336, 408, 0, 342, 414, 420,
    426, 432, 438, 444, 558, 564, 570, 264, 270, 276, 0, 0, 0, 0, 282, 288, 0, 0, 0, 0,
    0, 492, 498, 294, 0, 0, 318, 324, 330, 504, 348, 510, 354, 360, 366, 372, 516, 522, 528, 576,
    582, 588, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 450, 456, 462, 0, 0,
    468, 474, 480, 0, 0, 0, 486, 0, 0, 0, 594, 0, 0, 0, 0, 0, 0, 0, 0, 534,
    540, 546, 0, 0, 552] + [0] * 60, np.int32)
PREDICT_SPECIES = _predict_species_map()
_RATIOLG_R2E6 = math.log(1.0 + 1.0 / 2000000.0)


def _tablog_f32(idx: np.ndarray) -> np.ndarray:
    """tablog(i) = 10.0**((i - 16384) * 0.001) in REAL(4) (:1232-1234)."""
    x = ((idx.astype(np.int32) - 16384).astype(np.float32) * np.float32(0.001)).astype(np.float32)
    return powf10(x)


def predict_to_lte(rec: np.ndarray, grid: Grid) -> np.ndarray:
    """read_predict (:1110-1329) on an array of predicted records (int32
    (n, 4) or the raw 16-byte records, ascending IWL): the records from the
    first with IWL >= INT(ln(wlbeg - 1)/ln(1 + 1/2e6) + 0.5) up to the last
    with IWL <= the matching stop, species mapped through NELIONOLD
    (unmapped ones skipped), CGF = 0.01502 gf / nu, damping tablog / (4 pi
    nu) with 12.5664, nbuff on the SYNTHE grid, code = Z + ion / 100."""
    r = np.ascontiguousarray(rec).view(np.int32).reshape(-1, 4)
    if len(r) == 0:
        return np.zeros(0, LTE_DTYPE)
    istart = int(math.log(grid.wlbeg - 1.0) / _RATIOLG_R2E6 + 0.5)
    istop = int(math.log(grid.wlend + 1.0) / _RATIOLG_R2E6 + 0.5)
    iwl = r[:, 0]
    if iwl[0] > istop or iwl[-1] < istart:
        return np.zeros(0, LTE_DTYPE)
    # the reader's bisection: the first record >= istart, searched over
    # records 2..n (record 1 is never a candidate unless bisection ends there)
    lo, hi = 1, len(r)
    while hi - lo > 1:
        mid = (hi + lo) // 2
        if iwl[mid - 1] < istart:
            lo = mid
        else:
            hi = mid
    nstart = hi
    sel = r[nstart - 1:]
    past = np.nonzero(sel[:, 0] > istop)[0]
    if len(past):
        sel = sel[:past[0]]
    words = sel[:, 1:].copy().view(np.int16).reshape(-1, 6)
    nel_new = np.abs(words[:, 0].astype(np.int32)) // 10
    ok = (nel_new >= 1) & (nel_new <= 1005)
    nel = np.where(ok, PREDICT_SPECIES[np.clip(nel_new, 0, 1005)], 0)
    keep = nel != 0
    sel, words, nel = sel[keep], words[keep], nel[keep]
    wlvac = np.exp(sel[:, 0].astype(np.float64) * _RATIOLG_R2E6)
    freq = 2.99792458e17 / wlvac
    gf = _tablog_f32(words[:, 2]).astype(np.float64)
    congf = 0.01502 * gf / freq
    elo = _tablog_f32(words[:, 1]).astype(np.float64)
    frq4pi = freq * 12.5664
    out = np.zeros(len(sel), LTE_DTYPE)
    ixwl = np.floor(np.log(wlvac) / grid.ratiolg + 0.5).astype(np.int64)
    out["nbuff"] = (ixwl - grid.ixwlbeg + 1).astype(np.int32)
    out["cgf"] = congf.astype(np.float32)
    out["nelion"] = nel
    out["elo"] = elo.astype(np.float32)
    for k, dst in ((3, "gamrf"), (4, "gamsf"), (5, "gamwf")):
        out[dst] = (_tablog_f32(words[:, k]).astype(np.float64) / frq4pi).astype(np.float32)
    z = ((nel - 1) // 6 + 1).astype(np.float32)
    ion = ((nel - 1) % 6).astype(np.float32)
    out["code"] = (z + (ion / np.float32(100.0)).astype(np.float32)).astype(np.float32)
    return out


# ---------------------------------------------------------------------------
# MOL2 binary molecular lists (32-byte records)
# ---------------------------------------------------------------------------

# molec_dispatch (:2455-2586): isotope code [and molecule code] -> nelion,
# x1 + x2 (the isotopic log-abundance corrections added to log gf), and the
# ATLAS12 MOLCODE (0: a species outside ATLAS12's diatomic set).  A table
# the list format defines.  Entries keyed (iso, icode) override (iso, None).
# fudge is 0 in every case.
_MOL_DISPATCH = {
    (1, None): (240, 0.0, -5.0, 8410), (2, None): (240, 0.0, -4.469, 8411),
    (12, 606): (264, -.005, -.005, 8680), (12, 608): (276, -.005, -.001, 8700),
    (12, 106): (246, 0.0, -.005, 8460), (12, None): (270, -.005, -.002, 8690),
    (13, 606): (264, -.005, -1.955, 8681), (13, 608): (276, -1.955, -.001, 8701),
    (13, 106): (246, 0.0, -1.955, 8461), (13, None): (270, -1.955, -.002, 8691),
    (14, None): (252, 0.0, -.002, 8470),
    (15, 607): (270, -.005, -2.444, 8692), (15, None): (252, 0.0, -2.444, 8471),
    (16, 813): (324, 0.0, -.001, 0), (16, None): (258, 0.0, -.001, 8480),
    (17, 813): (324, 0.0, -3.398, 0), (17, None): (276, -.005, -3.398, 8702),
    (18, 814): (330, -.035, -2.690, 8896), (18, 608): (276, -.005, -2.690, 8704),
    (18, 813): (324, 0.0, -2.690, 0), (18, None): (258, 0.0, -2.690, 8482),
    (23, None): (492, 0.0, 0.0, 0),
    (24, 812): (318, -.001, -.102, 0), (24, None): (300, 0.0, -.105, 8510),
    (25, 812): (318, -.001, -1.000, 0), (25, None): (300, 0.0, -.996, 8511),
    (26, 812): (318, -.001, -.958, 0), (26, None): (300, 0.0, -.947, 8512),
    (28, 814): (330, -.035, -.001, 8890), (28, None): (312, 0.0, -.035, 8530),
    (29, 814): (330, -1.328, -.001, 8891), (29, None): (312, 0.0, -1.331, 8531),
    (30, 814): (330, -1.510, -.001, 8892), (30, None): (312, 0.0, -1.516, 8532),
    (33, None): (264, -1.955, -1.955, 8682), (39, None): (498, -.030, 0.0, 0),
    (40, 820): (354, -.013, -.001, 0), (40, None): (342, -.013, 0.0, 8580),
    (41, None): (498, -1.172, 0.0, 0), (42, None): (342, -2.189, 0.0, 8581),
    (43, None): (342, -2.870, 0.0, 8582), (44, None): (342, -1.681, 0.0, 8583),
    (46, 120): (342, -4.398, 0.0, 8584), (46, None): (366, 0.0, -1.101, 0),
    (47, None): (366, 0.0, -1.138, 0),
    (48, 120): (342, -2.728, 0.0, 0), (48, None): (366, 0.0, -.131, 0),
    (49, None): (366, 0.0, -1.259, 0),
    (50, 124): (432, -1.362, 0.0, 8620), (50, None): (366, 0.0, -1.272, 0),
    (51, None): (372, 0.0, -.001, 8960), (52, None): (432, -.077, 0.0, 8621),
    (53, None): (432, -1.022, 0.0, 8622),
    (54, 124): (432, -1.626, 0.0, 8623), (54, None): (444, -1.237, 0.0, 8640),
    (56, None): (444, -.038, 0.0, 8641), (57, None): (444, -1.658, 0.0, 8642),
    (58, None): (444, -2.553, 0.0, 8643),
}


def mol_dispatch(iso: int, icode: int):
    """(nelion, x1, x2) of molec_dispatch; nelion 0 = not a known isotope.
    x1, x2 are REAL(4) literals."""
    v = _MOL_DISPATCH.get((iso, icode)) or _MOL_DISPATCH.get((iso, None))
    if v is None:
        return 0, np.float32(0.0), np.float32(0.0)
    return v[0], np.float32(v[1]), np.float32(v[2])


MOLBROAD_TREF = 3000.0
MOLBROAD_JFLOOR = 0.1


def load_mol_broad_table(lines) -> dict:
    """load_mol_broad (:414-485) over the text lines of data/mol_broad.dat:
    code -> (MOLBROAD_GW, MOLBROAD_DG, MOLBROAD_GWEF), the J = 0 width, its
    slope in J and the population-weighted width."""
    kbol, clight, pi = 1.380649e-16, 2.99792458e10, 3.14159265358979
    pref, wh2, tref = 1.01325e6, 0.85, MOLBROAD_TREF
    out = {}
    for line in lines:
        if not line.strip() or line.startswith("#"):
            continue
        f = line.split()
        try:
            code = int(f[0])
            g2, n2, _, _, d2, _, brot = (float(x) for x in f[2:9])
            int(f[9])
        except (ValueError, IndexError):
            continue
        if code < 0 or code > 11000:
            continue
        conv = (296.0 / tref) ** n2 * (kbol * tref / pref) * 4.0 * pi * clight / \
            (wh2 * (tref / 1.0e4) ** 0.3)
        gw, dg = g2 * conv, d2 * conv
        if brot > 0.0 and dg > 0.0:
            wsum = gsum = 0.0
            for jj in range(301):
                x = float(jj)
                wj = (2.0 * x + 1.0) * math.exp(-brot * x * (x + 1.0) * 1.4388 / tref)
                gj = max(gw - dg * x, MOLBROAD_JFLOOR * gw)
                wsum = wsum + wj
                gsum = gsum + wj * gj
            gwef = gsum / wsum if wsum > 0.0 else -1.0
        else:
            gwef = gw
        out[code] = (gw, dg, gwef)
    return out


def load_mol_broad(lines) -> dict:
    """code -> population-weighted gamma_w (MOLBROAD_GWEF of load_mol_broad,
    :414-485), the value every MOL2 record of that molecule takes
    (molbroad_gw_elo with MOLBROAD_JFROM_ELO = .FALSE., :396-411)."""
    return {c: v[2] for c, v in load_mol_broad_table(lines).items()}


def mol2_to_lte(rec: np.ndarray, grid: Grid, molbroad: dict | None = None) -> np.ndarray:
    """read_molec_bin (:1371-1529) on MOL2 records (workloads.MOL2_DTYPE,
    ascending wavelength): from the first record with wl >= REAL(wlbeg - 1)
    up to the last with wl <= wlend + 1; FeH code 156 read as 126; isotopes
    through molec_dispatch (unknown ones skipped); gf = exp((log gf + x1 +
    x2) ln 10), CGF = 0.026538/1.77245 gf / nu, gamma_rad classical
    (2.223e13 / lambda^2) when the record carries 0, Stark 3e-5 (3e-8 for
    label 1), van der Waals 1e-7 (1e-8) or the species' tabulated width."""
    r = np.ascontiguousarray(rec)
    molbroad = molbroad or {}
    if len(r) == 0:
        return np.zeros(0, LTE_DTYPE)
    lo_wl = np.float32(grid.wlbeg - 1.0)
    wl32 = r["wl"].astype(np.float32)
    lo, hi = 1, len(r)
    while lo < hi:  # the reader's bisection over records 1..nrec
        mid = (lo + hi) // 2
        if wl32[mid - 1] < lo_wl:
            lo = mid + 1
        else:
            hi = mid
    rows = []
    for i in range(lo - 1, len(r)):
        x = r[i]
        icode = int(x["code"])
        if icode == 156:
            icode = 126
        wlvac = float(np.float32(x["wl"]))
        if wlvac > grid.wlend + 1.0:
            break
        nel, x1, x2 = mol_dispatch(int(x["iso"]), icode)
        if nel == 0:
            continue
        gf = math.exp((float(np.float32(x["loggf"])) + float(x1) + float(x2) + 0.0) *
                      2.30258509299405)
        elo = min(abs(float(np.float32(x["e1"]))), abs(float(np.float32(x["e2"]))))
        ixwl = math.floor(math.log(wlvac) / grid.ratiolg + 0.5)
        freq = 2.99792458e17 / wlvac
        congf = 0.026538 / 1.77245 * gf / freq
        frq4pi = freq * 12.5664
        loggr = int(x["gr100"])
        gammar = 2.223e13 / (wlvac * wlvac) if loggr == 0 else math.pow(10.0, loggr * 0.01)
        if int(x["flag"]) == 1:
            gammas, gammaw = 3.0e-8, 1.0e-8
        else:
            gammas, gammaw = 3.0e-5, 1.0e-7
        if 0 <= icode <= 11000 and molbroad.get(icode, -1.0) > 0.0:
            gammaw = molbroad[icode]
        rows.append((ixwl - grid.ixwlbeg + 1, np.float32(congf), nel, np.float32(elo),
                     np.float32(gammar / frq4pi), np.float32(gammas / frq4pi),
                     np.float32(gammaw * 1.0 / frq4pi), np.float32(icode)))
    return np.array(rows, LTE_DTYPE) if rows else np.zeros(0, LTE_DTYPE)


# ---------------------------------------------------------------------------
# ASCII diatomic lists for ATLAS12 (read_diatomics_for_atlas)
# ---------------------------------------------------------------------------

# diatomic_record_t (src/mod_mklinelist.f90:71-78), the BIND(C) mirror of
# a12_diatomic_rec
DIATOMIC_DTYPE = np.dtype([("wlvac_nm", "<f8"), ("gflog_dex", "<f8"), ("elo_cm", "<f8"),
                           ("gammar_log", "<f8"), ("gammaw_log", "<f8"), ("molcode", "<i4"),
                           ("pad", "<i4")])


def mol_molcode(iso: int, icode: int) -> int:
    """The ATLAS12 MOLCODE molec_dispatch returns (:2455-2586); 0 for a
    species outside ATLAS12's diatomic set."""
    v = _MOL_DISPATCH.get((iso, icode)) or _MOL_DISPATCH.get((iso, None))
    return 0 if v is None else v[3]


def _f_field(s: str, d: int) -> float:
    """Fw.d input editing of one field with blanks ignored (an all-blank
    field is 0); without a decimal point the last d digits are the
    fraction.  Raises ValueError where the Fortran read reports an error."""
    t = s.replace(" ", "")
    if not t:
        return 0.0
    t = t.replace("D", "E").replace("d", "e")
    mant, _, ex = t.partition("E") if "E" in t else t.partition("e")
    if not mant or mant in "+-":
        raise ValueError(s)
    if "." not in mant:
        return float(f"{mant}e{int(ex or 0) - d}")
    return float(t)


def _i_field(s: str) -> int:
    t = s.replace(" ", "")
    return int(t) if t else 0


def _mol_line(raw, width: int, code_w: int):
    """One line of a Kurucz molecular list read as the readers do: into a
    width-character buffer (padded, longer lines cut), a trailing CR
    blanked, then (F10.4, F7.3, F5.1, F10.3, F5.1, F11.3, I<code_w>, A8, A8,
    I2, I4).  Returns (lambda, log gf, J, E, J', E', code, upper label,
    isotope, 100 log gamma_rad) with the REAL(4) fields as float32, or None
    for a blank line or one the format rejects."""
    if isinstance(raw, bytes):
        raw = raw.decode("latin-1")
    buf = raw.rstrip("\n")[:width].ljust(width)
    if not buf.strip():
        return None
    lt = len(buf.rstrip())
    if buf[lt - 1] == "\r":
        buf = buf[:lt - 1] + " " + buf[lt:]
    c = 48 + code_w
    try:
        return (_f_field(buf[0:10], 4), np.float32(_f_field(buf[10:17], 3)),
                np.float32(_f_field(buf[17:22], 1)), _f_field(buf[22:32], 3),
                np.float32(_f_field(buf[32:37], 1)), _f_field(buf[37:48], 3),
                _i_field(buf[48:c]), buf[c + 8:c + 16], _i_field(buf[c + 16:c + 18]),
                _i_field(buf[c + 18:c + 22]))
    except ValueError:
        return None


def read_diatomic_ascii(lines, wlbeg_nm: float, wlend_nm: float, broad: dict | None = None,
                        gammaw_scale: float = 1.0) -> np.ndarray:
    """read_one_diatomic_ascii (:2312-2427) over the text lines of one
    Kurucz molecular list: fixed columns (F10.4, F7.3, F5.1, F10.3, F5.1,
    F11.3, I4, A8, A8, I2, I4) = lambda, log gf (REAL(4)), J, E, J', E',
    code, labels, isotope, 100 log gamma_rad.  Blank lines and lines the
    format rejects are skipped; the file ends at the first lambda beyond
    wlend + 2.  Vacuum wavelength 1e7 / ||E'| - |E||, kept within
    [wlbeg, wlend]; species through molec_dispatch (MOLCODE 0 skipped);
    E_lo = min(|E|, |E'|); gamma_rad classical (2.223e13 / lambda^2) for 0;
    gamma_w 1e-8 for an 'X' upper label else 1e-7, or the species'
    J-dependent width from broad (load_mol_broad_table; ATLAS12 runs
    without it), times gammaw_scale.  Returns DIATOMIC_DTYPE records in
    file order."""
    broad = broad or {}
    rows = []
    for raw in lines:
        f = _mol_line(raw, 160, 4)
        if f is None:
            continue
        wl, gflog, xj, e, xjp, ep, icode, clabelp, iso, loggr = f
        if abs(wl) > wlend_nm + 2.0:
            break
        de = abs(abs(ep) - abs(e))
        wlvac = 1.0e7 / de if de != 0.0 else math.inf
        if wlvac < wlbeg_nm or wlvac > wlend_nm:
            continue
        molcode = mol_molcode(iso, icode)
        if molcode == 0:
            continue
        _, x1, x2 = mol_dispatch(iso, icode)
        elo = min(abs(e), abs(ep))
        gammar = 2.223e13 / (wlvac * wlvac) if loggr == 0 else math.pow(10.0, loggr * 0.01)
        gammaw = 1.0e-8 if clabelp[0] == "X" else 1.0e-7
        if 0 <= icode <= 11000 and broad.get(icode, (-1.0, 0.0, 0.0))[0] > 0.0:
            gw, dg, _ = broad[icode]
            xjlow = float(xj if abs(e) <= abs(ep) else xjp)
            gammaw = max(gw - dg * xjlow, MOLBROAD_JFLOOR * gw)
        gammaw = gammaw * gammaw_scale
        rows.append((wlvac, float(gflog) + float(x1) + float(x2) + 0.0, elo,
                     math.log10(gammar), math.log10(gammaw), molcode, 0))
    return np.array(rows, DIATOMIC_DTYPE) if rows else np.zeros(0, DIATOMIC_DTYPE)


def read_diatomics_for_atlas(files, wlbeg_nm: float, wlend_nm: float,
                             broad: dict | None = None) -> np.ndarray:
    """read_diatomics_for_atlas (:2244-2299): the ASCII molecular lists
    named by lines.list ``mol`` entries (``.bin`` entries excluded, missing
    files skipped), each read by :func:`read_diatomic_ascii` with the
    SELECTLINES window [WAVETAB(1), WAVETAB(NWAVE)], concatenated in
    manifest order, unsorted."""
    import os
    parts = []
    for f in files:
        if len(f) >= 4 and f[-4:] == ".bin":
            continue
        if not os.path.exists(f):
            continue
        with open(f, "rb") as fh:
            parts.append(read_diatomic_ascii(fh.read().split(b"\n"), wlbeg_nm, wlend_nm, broad))
    return np.concatenate(parts) if parts else np.zeros(0, DIATOMIC_DTYPE)


# ---------------------------------------------------------------------------
# SYNTHE's molecular ASCII, polyatomic and H2O readers
# ---------------------------------------------------------------------------

TEFF_COOL_LIMIT = 8000.0      # TiO and H2O lists read at or below (:251)
TEFF_POLYMOL_LIMIT = 3500.0   # polyatomic lists read at or below (:247)


def _lte_row(wlvac, grid, gf, elo, gammar, gammas, gammaw, nelion, code, congf_pref):
    ixwl = math.floor(math.log(wlvac) / grid.ratiolg + 0.5)
    freq = 2.99792458e17 / wlvac
    congf = congf_pref * gf / freq
    frq4pi = freq * 12.5664
    return (ixwl - grid.ixwlbeg + 1, np.float32(congf), nelion, np.float32(elo),
            np.float32(gammar / frq4pi), np.float32(gammas / frq4pi),
            np.float32(gammaw * 1.0 / frq4pi), np.float32(code))


def _ascii_widths(clabelp, icode, xj, xjp, e, ep, broad):
    """gamma_s, gamma_w of the ASCII molecular readers: 3e-8, 1e-8 for an
    'X' upper label else 3e-5, 1e-7; a tabulated width (broad: code ->
    (GW, DG, GWEF)) at the J of the lower term, floored at JFLOOR GW."""
    gammas, gammaw = (3.0e-8, 1.0e-8) if clabelp[0] == "X" else (3.0e-5, 1.0e-7)
    if 0 <= icode <= 11000 and broad.get(icode, (-1.0, 0.0, 0.0))[0] > 0.0:
        gw, dg, _ = broad[icode]
        xjlow = float(xj if abs(e) <= abs(ep) else xjp)
        gammaw = max(gw - dg * xjlow, MOLBROAD_JFLOOR * gw)
    return gammas, gammaw


def molec_ascii_to_lte(lines, grid: Grid, broad: dict | None = None) -> np.ndarray:
    """read_molec_ascii (src/mod_mklinelist.f90:1535-1689) over the text
    lines of one Kurucz molecular list (the ``mol`` entries without a .bin
    sidecar): FeH code 156 read as 126; the file ends at the first lambda
    beyond wlend + 2; vacuum wavelength 1e7 / ||E'| - |E|| kept within
    [wlbeg - 1, wlend + 1]; isotopes through molec_dispatch (unknown ones
    skipped); gf = exp((log gf + x1 + x2) ln 10) with the sum in REAL(4);
    CGF 0.026538/1.77245 gf / nu; widths as :func:`_ascii_widths`.  broad:
    :func:`load_mol_broad_table`."""
    broad = broad or {}
    rows = []
    for raw in lines:
        f = _mol_line(raw, 160, 4)
        if f is None:
            continue
        wl, gflog, xj, e, xjp, ep, icode, clabelp, iso, loggr = f
        if icode == 156:
            icode = 126
        if abs(wl) > grid.wlend + 2.0:
            break
        de = abs(abs(ep) - abs(e))
        wlvac = 1.0e7 / de if de != 0.0 else math.inf
        if wlvac < grid.wlbeg - 1.0 or wlvac > grid.wlend + 1.0:
            continue
        nel, x1, x2 = mol_dispatch(iso, icode)
        if nel == 0:
            continue
        s4 = np.float32(np.float32(np.float32(gflog + x1) + x2) + np.float32(0.0))
        gf = math.exp(float(s4) * 2.30258509299405)
        elo = min(abs(e), abs(ep))
        gammar = 2.223e13 / (wlvac * wlvac) if loggr == 0 else math.pow(10.0, loggr * 0.01)
        gammas, gammaw = _ascii_widths(clabelp, icode, xj, xjp, e, ep, broad)
        rows.append(_lte_row(wlvac, grid, gf, elo, gammar, gammas, gammaw, nel, icode,
                             0.026538 / 1.77245))
    return np.array(rows, LTE_DTYPE) if rows else np.zeros(0, LTE_DTYPE)


def poly_nelion(icode: int) -> int:
    """poly_nelion (:2018-2027): CaOH (10820) -> slot 636, else 0."""
    return 636 if icode == 10820 else 0


def molec_poly_to_lte(lines, grid: Grid, broad: dict | None = None) -> np.ndarray:
    """read_molec_poly (:1690-1838): as :func:`molec_ascii_to_lte` with a
    6-digit molecule code, a 256-character line, the species from
    :func:`poly_nelion` and gf = exp(log gf ln 10) (no isotope terms)."""
    broad = broad or {}
    rows = []
    for raw in lines:
        f = _mol_line(raw, 256, 6)
        if f is None:
            continue
        wl, gflog, xj, e, xjp, ep, icode, clabelp, iso, loggr = f
        if abs(wl) > grid.wlend + 2.0:
            break
        de = abs(abs(ep) - abs(e))
        wlvac = 1.0e7 / de if de != 0.0 else math.inf
        if wlvac < grid.wlbeg - 1.0 or wlvac > grid.wlend + 1.0:
            continue
        nel = poly_nelion(icode)
        if nel == 0:
            continue
        gf = math.exp(float(gflog) * 2.30258509299405)
        elo = min(abs(e), abs(ep))
        gammar = 2.223e13 / (wlvac * wlvac) if loggr == 0 else math.pow(10.0, loggr * 0.01)
        gammas, gammaw = _ascii_widths(clabelp, icode, xj, xjp, e, ep, broad)
        rows.append(_lte_row(wlvac, grid, gf, elo, gammar, gammas, gammaw, nel, icode,
                             0.026538 / 1.77245))
    return np.array(rows, LTE_DTYPE) if rows else np.zeros(0, LTE_DTYPE)


def _bisect_start(iwl: np.ndarray, lo: int, hi: int, bound: float) -> int | None:
    """The readers' bisection over 1-based records lo..hi on
    exp(IWL ln(1 + 1/2e6)) < bound; returns the last probed record (the
    readers start there), None when no probe is made."""
    newlimit = None
    while hi - lo > 1:
        newlimit = (hi + lo) // 2
        if math.exp(int(iwl[newlimit - 1]) * _RATIOLG_R2E6) < bound:
            lo = newlimit
        else:
            hi = newlimit
    return newlimit


def molec_poly_bin_to_lte(words: np.ndarray, grid: Grid, gwef: dict | None = None) -> np.ndarray:
    """read_molec_poly_bin (:1864-2012) on a polyatomic binary list as int32
    (n, 2): header (-1, code), then (IWL, int16 E_lo | int16 IGFLOG << 16)
    ascending.  Records from the bisection's last probe up to lambda >
    wlend + 1, those below wlbeg - 1 skipped; CGF 0.026538/1.77245
    tablog(|IGFLOG|) / nu (REAL(4) tablog); classical gamma_rad, gamma_s
    3e-5, gamma_w 1e-7 or the species' population-weighted width (gwef:
    :func:`load_mol_broad`)."""
    w = np.ascontiguousarray(words, np.int32).reshape(-1, 2)
    gwef = gwef or {}
    if len(w) == 0 or w[0, 0] != -1:
        raise ValueError("molbin header missing")
    icode = int(w[0, 1])
    nel = poly_nelion(icode)
    n = len(w)
    if nel == 0 or n < 2:
        return np.zeros(0, LTE_DTYPE)
    iwl = w[:, 0]
    if math.exp(int(iwl[1]) * _RATIOLG_R2E6) > grid.wlend + 1.0:
        return np.zeros(0, LTE_DTYPE)
    if grid.wlbeg - 1.0 > math.exp(int(iwl[n - 1]) * _RATIOLG_R2E6):
        return np.zeros(0, LTE_DTYPE)
    istart = _bisect_start(iwl, 2, n, grid.wlbeg - 1.0)
    if istart is None:
        raise ValueError("molbin list too short for the reader's bisection")
    halves = w[:, 1].copy().view(np.int16).reshape(-1, 2)
    tab = _tablog_f32(np.arange(32769)).astype(np.float64)
    gw = gwef.get(icode, -1.0) if 0 <= icode <= 11000 else -1.0
    gamw = gw if gw > 0.0 else 1.0e-7
    rows = []
    for i in range(istart - 1, n):
        wlvac = math.exp(int(iwl[i]) * _RATIOLG_R2E6)
        if wlvac > grid.wlend + 1.0:
            break
        if wlvac < grid.wlbeg - 1.0:
            continue
        ielo, igf = int(halves[i, 0]), int(halves[i, 1])
        elo = float(np.int16(abs(ielo)))
        rows.append(_lte_row(wlvac, grid, tab[np.int16(abs(igf))], elo, 2.223e13 / (wlvac * wlvac),
                             3.0e-5, gamw, nel, icode, 0.026538 / 1.77245))
    return np.array(rows, LTE_DTYPE) if rows else np.zeros(0, LTE_DTYPE)


H2O_XISO = np.array([0.9976, 0.0004, 0.0020, 0.00001], np.float32)


def h2o_to_lte(rec: np.ndarray, grid: Grid, gwef: dict | None = None) -> np.ndarray:
    """read_h2o (:2034-2174) on 8-byte H2O records (workloads.H2O_DTYPE,
    ascending IWL): nothing when the first record lies beyond wlend + 1 or
    the last below wlbeg - 1; else the records from the bisection's last
    probe (with no lower bound after it) up to lambda > wlend + 1.  Isotope
    from the signs of E_lo and IGFLOG; CGF 0.01502 tablog(|IGFLOG|) / nu
    times the isotope fraction; gamma_rad 2.223e13 / lambda^2 / 1000,
    gamma_s tablog(1), gamma_w tablog(9384) or H2O's population-weighted
    width (code 10108 in gwef) rounded to REAL(4); species 534, code 10108."""
    r = np.ascontiguousarray(rec)
    gwef = gwef or {}
    n = len(r)
    if n == 0:
        return np.zeros(0, LTE_DTYPE)
    iwl = r["iwl"]
    if math.exp(int(iwl[0]) * _RATIOLG_R2E6) > grid.wlend + 1.0:
        return np.zeros(0, LTE_DTYPE)
    if grid.wlbeg - 1.0 > math.exp(int(iwl[n - 1]) * _RATIOLG_R2E6):
        return np.zeros(0, LTE_DTYPE)
    istart = _bisect_start(iwl, 1, n, grid.wlbeg - 1.0)
    if istart is None:
        raise ValueError("H2O list too short for the reader's bisection")
    tab = _tablog_f32(np.arange(32769)).astype(np.float64)
    gw = gwef.get(10108, -1.0)
    gamw0 = float(np.float32(gw)) if gw > 0.0 else tab[9384]
    rows = []
    for i in range(istart - 1, n):
        wlvac = math.exp(int(iwl[i]) * _RATIOLG_R2E6)
        if wlvac > grid.wlend + 1.0:
            break
        ielo, igf = int(r["ielo"][i]), int(r["igflog"][i])
        iso = (1 if igf > 0 else 2) if ielo > 0 else (3 if igf > 0 else 4)
        ixwl = math.floor(math.log(wlvac) / grid.ratiolg + 0.5)
        freq = 2.99792458e17 / wlvac
        congf = 0.01502 * tab[np.int16(abs(igf))] / freq * float(H2O_XISO[iso - 1])
        frq4pi = freq * 12.5664
        gammar = 2.223e13 / (wlvac * wlvac) * 0.001
        rows.append((ixwl - grid.ixwlbeg + 1, np.float32(congf), 534,
                     np.float32(np.int16(abs(ielo))), np.float32(gammar / frq4pi),
                     np.float32(tab[1] / frq4pi), np.float32(gamw0 * 1.0 / frq4pi),
                     np.float32(10108.0)))
    return np.array(rows, LTE_DTYPE) if rows else np.zeros(0, LTE_DTYPE)
