"""Seeded synthetic workloads for the SYNTHE/ATLAS12 hot path.

Builds, from a seed alone, the inputs the line-opacity build and the
transfer solve consume:

* model-atmosphere decks in the Kurucz ``READ DECK6`` card format read by
  READIN (src/atlas12_modules.f90:6644-6944), e.g. the 72-layer solar-like
  structure of the plain-SYNTHE workload;
* line lists written as gfall ASCII records (the fixed format read by
  read_gfall, src/mod_mklinelist.f90:860-866) and, for the GPU, the
  ``lte_line_t`` records that reader produces from them
  (src/mod_mklinelist.f90:888-1064, restated in :func:`gfall_to_lte`).

Nothing here needs the reference tree: the GPU box regenerates the same
inputs from the same seeds.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

# lte_line_t (src/mod_mklinelist.f90:30-41), 32 bytes
LTE_DTYPE = np.dtype([("nbuff", "<i4"), ("cgf", "<f4"), ("nelion", "<i4"),
                      ("elo", "<f4"), ("gamrf", "<f4"), ("gamsf", "<f4"),
                      ("gamwf", "<f4"), ("code", "<f4")])

# Species of the plain-SYNTHE workload: Fe I/II, Ca I/II, Ti I, Na I, Mg I
SYNTHE_SPECIES = (26.00, 26.01, 20.00, 20.01, 22.00, 11.00, 12.00)

EV_TO_CM = 8065.544  # 1 eV in cm^-1


@dataclass
class Grid:
    """SYNTHE wavelength grid, src/synthe.f90:316-324."""
    wlbeg: float
    wlend: float
    resolu: float
    ratio: float
    ratiolg: float
    length: int
    ixwlbeg: int
    wbegin: float

    @staticmethod
    def make(wlbeg: float, wlend: float, resolu: float) -> "Grid":
        ratio = 1.0 + 1.0 / resolu
        ratiolg = math.log(ratio)
        length = int(math.log(wlend / wlbeg) / ratiolg)
        ixwlbeg = int(math.log(wlbeg) / ratiolg)
        wbegin = math.exp(ixwlbeg * ratiolg)
        if wbegin < wlbeg:
            ixwlbeg += 1
            wbegin = math.exp(ixwlbeg * ratiolg)
        return Grid(wlbeg, wlend, resolu, ratio, ratiolg, length, ixwlbeg, wbegin)


# ---------------------------------------------------------------------------
# Model atmospheres
# ---------------------------------------------------------------------------

def _saha_h_electrons(t: np.ndarray, p: np.ndarray, donor: float = 1.0e-4) -> np.ndarray:
    """Electron density from pure-hydrogen Saha ionisation plus a metal
    donor floor (fraction of nuclei), for a starting deck only; SYNTHE
    recomputes n_e with its own equation of state."""
    k = 1.380649e-16
    h = 6.62607015e-27
    me = 9.1093837139e-28
    chi = 13.598 * 1.602176634e-12
    ntot = p / (k * t)
    s = (2.0 * math.pi * me * k * t / h**2) ** 1.5 * np.exp(-chi / (k * t))
    # x^2 n / (1 - x) = s  with n the hydrogen density (electrons neglected)
    x = (-s + np.sqrt(s * s + 4.0 * ntot * s)) / (2.0 * ntot)
    x = np.maximum(x, donor)
    return x * ntot / (1.0 + x)


@dataclass
class Atmosphere:
    teff: float
    logg: float
    rhox: np.ndarray
    t: np.ndarray
    p: np.ndarray
    xne: np.ndarray
    vturb: np.ndarray  # cm/s
    abross: np.ndarray | None = None   # Rosseland opacity (default 1)
    accrad: np.ndarray | None = None   # radiative acceleration (default 0)
    pradk: float = 0.0                 # PRADK card

    @property
    def nrhox(self) -> int:
        return len(self.rhox)

    def deck(self) -> str:
        """Kurucz card deck (READ DECK6 RHOX,T,P,XNE,ABROSS,ACCRAD,VTURB)."""
        out = [f"TEFF   {self.teff:.0f}.  GRAVITY {self.logg:.5f} LTE",
               "TITLE SYNTHETIC",
               f"READ DECK6 {self.nrhox} RHOX,T,P,XNE,ABROSS,ACCRAD,VTURB"]
        ab = np.ones(self.nrhox) if self.abross is None else self.abross
        ac = np.zeros(self.nrhox) if self.accrad is None else self.accrad
        for j in range(self.nrhox):
            out.append(f" {self.rhox[j]:.8E} {self.t[j]:8.1f} {self.p[j]:.3E} {self.xne[j]:.3E}"
                       f" {ab[j]:.3E} {ac[j]:.3E} {self.vturb[j]:.3E}")
        out.append(f"PRADK {self.pradk:.4E}")
        out.append("BEGIN                    ITERATION   1 COMPLETED")
        return "\n".join(out) + "\n"


def scurve_atmosphere(nrhox: int = 72, t_top: float = 4400.0, t_bot: float = 9500.0,
                      logrhox0: float = -3.0, logrhox1: float = 2.0, logg: float = 4.44,
                      teff: float = 5777.0, vturb_kms: float = 1.5) -> Atmosphere:
    """RHOX log-spaced, T rising on the smooth s-curve T = t_top + dT s^2 (3-2s)
    (the structure of the plain-SYNTHE and large-SYNTHE workloads)."""
    s = np.arange(nrhox) / (nrhox - 1)
    rhox = 10.0 ** (logrhox0 + (logrhox1 - logrhox0) * s)
    t = t_top + (t_bot - t_top) * s * s * (3.0 - 2.0 * s)
    # the deck carries 1 decimal of T; round so the deck is the model
    t = np.round(t, 1)
    p = 10.0 ** logg * rhox
    xne = _saha_h_electrons(t, p)
    vt = np.full(nrhox, vturb_kms * 1.0e5)
    return Atmosphere(teff, logg, rhox, t, p, xne, vt)


# ---------------------------------------------------------------------------
# Line lists
# ---------------------------------------------------------------------------

GFALL_FIELDS = [("wl", "f8"), ("loggf", "f4"), ("code", "f4"), ("e", "f8"),
                ("ep", "f8"), ("gr", "f4"), ("gs", "f4"), ("gw", "f4")]


def synthetic_gfall(n: int, wl0: float, wl1: float, seed: int,
                    species=SYNTHE_SPECIES, loggf=("uniform", -5.0, 0.0),
                    elo=("uniform", 0.0, 8.0), lgr=(7.0, 9.0), lgs=(-6.0, -4.0),
                    lgw=(-8.0, -7.0)) -> np.ndarray:
    """Seeded atomic lines, centres uniform in ln(lambda) over [wl0, wl1] nm,
    quantised to the gfall print precision (1e-4 nm, 1e-3 dex, 1e-3 cm^-1,
    1e-2 dex) and sorted by wavelength (read_gfall stops at the first record
    past the window, src/mod_mklinelist.f90:896)."""
    rng = np.random.default_rng(seed)
    wl = np.exp(rng.uniform(math.log(wl0), math.log(wl1), n))
    if loggf[0] == "uniform":
        gf = rng.uniform(loggf[1], loggf[2], n)
    else:
        gf = rng.normal(loggf[1], loggf[2], n)
    if elo[0] == "uniform":
        e_ev = rng.uniform(elo[1], elo[2], n)
    else:
        e_ev = rng.exponential(elo[1], n)
    sp = np.asarray(species)[rng.integers(0, len(species), n)]
    gr = rng.uniform(*lgr, n)
    gs = rng.uniform(*lgs, n)
    gw = rng.uniform(*lgw, n)
    rec = np.zeros(n, dtype=GFALL_FIELDS)
    rec["wl"] = np.round(wl, 4)
    rec["loggf"] = np.clip(np.round(gf, 3), -99.0, 99.0)
    rec["code"] = sp
    e_cm = np.round(e_ev * EV_TO_CM, 3)
    rec["e"] = e_cm
    rec["ep"] = np.round(e_cm + 1.0e7 / rec["wl"], 3)
    # a zero log-gamma selects the reader's default widths; keep them non-zero
    for name, v in (("gr", gr), ("gs", gs), ("gw", gw)):
        q = np.round(v, 2)
        q[q == 0.0] = 0.01
        rec[name] = q
    rec.sort(order="wl", kind="stable")
    return rec


def write_gfall(rec: np.ndarray, path: str) -> None:
    """gfall ASCII, FORMAT(F11.4,F7.3,F6.2,F12.3,F5.1,1X,A8,A2,F12.3,F5.1,1X,A8,A2,
    F6.2,F6.2,F6.2,A4,I2,I2,I3,F6.3,I3,F6.3,A8,A2,A8,A2,2I5,I6)."""
    lab = "SYNTH   " + "  "
    with open(path, "w") as fh:
        for r in rec:
            fh.write(f"{r['wl']:11.4f}{r['loggf']:7.3f}{r['code']:6.2f}{r['e']:12.3f}{0.5:5.1f} {lab}"
                     f"{r['ep']:12.3f}{1.5:5.1f} {lab}{r['gr']:6.2f}{r['gs']:6.2f}{r['gw']:6.2f}"
                     f"SYN {0:2d}{0:2d}{0:3d}{0.0:6.3f}{0:3d}{0.0:6.3f}{'':8s}{'':2s}{'':8s}{'':2s}"
                     f"{0:5d}{0:5d}{0:6d}\n")


_LIBM = None


def powf10(x: np.ndarray) -> np.ndarray:
    """10.0**x for REAL(4) x as the reference's runtime evaluates it (C
    ``powf``).  powf works in double internally and rounds once, so it
    equals pow() rounded to FP32 except when that double lands within a
    hair of an FP32 rounding midpoint; only those elements call powf."""
    global _LIBM
    x = np.asarray(x, dtype=np.float32)
    d = np.power(10.0, x.astype(np.float64))
    f = d.astype(np.float32)
    # distance of d from the FP32 midpoint next to f, in FP32 ulps
    up = np.nextafter(f, np.float32(np.inf))
    dn = np.nextafter(f, np.float32(0))
    mid = np.where(d >= f.astype(np.float64), (f.astype(np.float64) + up) * 0.5,
                   (f.astype(np.float64) + dn) * 0.5)
    ulp = np.abs(up.astype(np.float64) - f.astype(np.float64))
    near = np.abs(d - mid) < 2.0e-2 * ulp
    if near.any():
        if _LIBM is None:
            import ctypes
            _LIBM = ctypes.CDLL("libm.so.6")
            _LIBM.powf.restype = ctypes.c_float
            _LIBM.powf.argtypes = [ctypes.c_float, ctypes.c_float]
        idx = np.nonzero(near)[0]
        f[idx] = np.array([_LIBM.powf(10.0, float(v)) for v in x[idx]], dtype=np.float32)
    return f


def gfall_to_lte(rec: np.ndarray, grid: Grid) -> np.ndarray:
    """``lte_line_t`` records as read_gfall builds them for plain LTE lines
    (src/mod_mklinelist.f90:888-1064): window selection with the 0.1 nm
    margin (lim = 7, :846-847, :903-907), nbuff = INT(ln lambda / ln ratio +
    0.5) - ixwlbeg + 1, gf = 10.0**loggf in FP32, cgf = 0.026538/1.77245 gf /
    nu, damping 10**log gamma / (12.5664 nu).  Expression types follow the
    reference (FP32 where it is REAL(4))."""
    delfactor = grid.wlbeg / 500.0 if grid.wlbeg > 500.0 else 1.0
    wl = rec["wl"].astype(np.float64)
    keep = (wl >= grid.wlbeg - 0.1 * delfactor) & (wl <= grid.wlend + 0.1 * delfactor)
    r = rec[keep]
    wl = r["wl"].astype(np.float64)
    out = np.zeros(len(r), dtype=LTE_DTYPE)
    ixwl = np.floor(np.log(wl) / grid.ratiolg + 0.5).astype(np.int64)
    out["nbuff"] = (ixwl - grid.ixwlbeg + 1).astype(np.int32)
    gf = powf10(r["loggf"])
    frelin = 2.99792458e17 / wl
    cgf = 0.026538 / 1.77245 * gf.astype(np.float64) / frelin
    frq4pi = 12.5664 * frelin
    code = r["code"].astype(np.float32)
    nelem = np.floor(code).astype(np.int32)
    icharge = np.rint((code - nelem.astype(np.float32)) * np.float32(100.0)).astype(np.int32)
    out["nelion"] = nelem * 6 - 6 + (icharge + 1)
    out["cgf"] = cgf.astype(np.float32)
    out["elo"] = np.minimum(np.abs(r["e"]), np.abs(r["ep"])).astype(np.float32)
    for src, dst in (("gr", "gamrf"), ("gs", "gamsf"), ("gw", "gamwf")):
        g = np.power(10.0, r[src].astype(np.float32).astype(np.float64))
        out[dst] = (g / frq4pi).astype(np.float32)
    out["code"] = code
    return out


def synthe_plain_lines(grid: Grid, n: int = 200_000, seed: int = 1) -> np.ndarray:
    """The plain-SYNTHE line list: n atomic lines over the grid's window."""
    rec = synthetic_gfall(n, grid.wlbeg, grid.wlend, seed)
    return gfall_to_lte(rec, grid)


# ---------------------------------------------------------------------------
# ATLAS12 opacity-sampling line store
# ---------------------------------------------------------------------------

# ATLAS species slot of atom Z, ion stage i (1 = neutral): Z(Z+1)/2 + i - 1
# (src/atlas12_modules.f90:10148-10150, NOFF :10194-10197)
def atlas_nelion(z: int, ion: int) -> int:
    return z * (z + 1) // 2 + ion - 1


ATLAS_SPECIES = tuple(atlas_nelion(z, i) for z, i in
                      ((26, 1), (26, 2), (20, 1), (20, 2), (22, 1), (11, 1), (12, 1)))
LOG_RATIO_2E6 = math.log(1.0 + 1.0 / 2000000.0)


def _tablog_code(x):
    """Inverse of TABLOG (src/atlas12_modules.f90:2083-2086): log10 value ->
    int16 code 1000 x + 16384."""
    return np.clip(np.rint(np.asarray(x) * 1000.0) + 16384, -32768, 32767).astype(np.int16)


def atlas_linedata(n: int, wl0: float, wl1: float, seed: int, species=ATLAS_SPECIES,
                   loggf=("normal", -3.0, 1.5), elo=("exp", 2.0), lgr=(7.0, 9.0),
                   lgs=(-6.0, -4.0), lgw=(-8.0, -7.0)) -> np.ndarray:
    """Seeded LINEDATA records (4 x int32 per line: IWL, then int16 pairs
    (IELION, IELO), (IGFLOG, IGR), (IGS, IGW)) -- the packed layout that
    SELECTLINES stores and LINOP1 unpacks (src/atlas12_modules.f90:1950-1963).
    Sorted by IWL."""
    rng = np.random.default_rng(seed)
    lam = np.exp(rng.uniform(math.log(wl0), math.log(wl1), n))
    iwl = np.rint(np.log(lam) / LOG_RATIO_2E6).astype(np.int64)
    sp = np.asarray(species)[rng.integers(0, len(species), n)]
    gf = rng.normal(loggf[1], loggf[2], n) if loggf[0] == "normal" else rng.uniform(loggf[1], loggf[2], n)
    e_ev = rng.exponential(elo[1], n) if elo[0] == "exp" else rng.uniform(elo[1], elo[2], n)
    e_cm = np.maximum(e_ev * EV_TO_CM, 1.0e-3)
    rec = np.zeros((n, 4), np.int32)
    rec[:, 0] = iwl
    pairs = np.zeros((n, 6), np.int16)
    pairs[:, 0] = (sp * 10).astype(np.int16)
    pairs[:, 1] = _tablog_code(np.log10(e_cm))
    pairs[:, 2] = _tablog_code(gf)
    pairs[:, 3] = _tablog_code(rng.uniform(*lgr, n))
    pairs[:, 4] = _tablog_code(rng.uniform(*lgs, n))
    pairs[:, 5] = _tablog_code(rng.uniform(*lgw, n))
    rec[:, 1:] = pairs.view(np.int32).reshape(n, 3)
    order = np.argsort(rec[:, 0], kind="stable")
    return np.ascontiguousarray(rec[order])


# MOL2 32-byte molecular record, the one binary molecular list SELECTLINES
# reads (src/atlas12_modules.f90:19540-19647): vacuum wavelength [nm], log gf,
# the two term energies [cm^-1], molecule code, isotope, 100 log10(gamma_rad)
# (0 = classical), upper-label flag.  Header: 8 x int32, magic 'MOL2', count.
MOL2_DTYPE = np.dtype([("wl", "<f4"), ("loggf", "<f4"), ("e1", "<f4"), ("e2", "<f4"),
                       ("code", "<i4"), ("iso", "<i4"), ("gr100", "<i4"), ("flag", "<i4")])
MOL2_MAGIC = 0x4D4F4C32
TIO_CODE = 822
TIO_NELION = 895


def tio_mol2_records(n: int, wl0: float, wl1: float, seed: int,
                     loggf=(-3.0, 1.5), elo_ev=2.0) -> np.ndarray:
    """Seeded TiO (code 822, isotopes 46-50) MOL2 records, ascending in
    wavelength; log gf ~ N(-3, 1.5), lower term energy ~ Exp(2 eV), the upper
    term 1e7/lambda above it, one record in four with classical damping."""
    rng = np.random.default_rng(seed)
    wl = np.sort(np.exp(rng.uniform(math.log(wl0), math.log(wl1), n))).astype(np.float32)
    rec = np.zeros(n, MOL2_DTYPE)
    rec["wl"] = wl
    rec["loggf"] = np.round(rng.normal(loggf[0], loggf[1], n), 3)
    e_lo = np.round(rng.exponential(elo_ev, n) * EV_TO_CM, 3)
    e_up = np.round(e_lo + 1.0e7 / wl.astype(np.float64), 3)
    swap = rng.random(n) < 0.5          # the list does not order the two terms
    rec["e1"] = np.where(swap, e_up, e_lo)
    rec["e2"] = np.where(swap, e_lo, e_up)
    rec["code"] = TIO_CODE
    rec["iso"] = rng.integers(46, 51, n)
    gr = np.rint(rng.uniform(7.0, 9.0, n) * 100.0).astype(np.int32)
    rec["gr100"] = np.where(rng.random(n) < 0.25, 0, gr)
    return rec


def write_mol2(rec: np.ndarray, path: str) -> None:
    hdr = np.zeros(8, "<i4")
    hdr[0] = MOL2_MAGIC
    hdr[1] = len(rec)
    with open(path, "wb") as fh:
        fh.write(hdr.tobytes())
        fh.write(np.ascontiguousarray(rec, MOL2_DTYPE).tobytes())


def atlas_iteration_lists(n: int, wl0: float, wl1: float, seed: int, molecules=(),
                          molecular_share: float = 0.2, tio_loggf=(-3.0, 1.5)):
    """Raw line lists of one ATLAS12 iteration: (predicted 16-byte records,
    TiO MOL2 records).  (1 - share) of n are atomic lines over ATLAS_SPECIES;
    of the molecular share, half are diatomics in the predicted format over
    the given molecule slots and half TiO MOL2 records.  log gf ~ N(-3, 1.5),
    E_lo ~ Exp(2 eV).  Both lists ascending in wavelength, as SELECTLINES
    requires."""
    n_mol = int(round(n * molecular_share))
    n_tio = n_mol // 2 if molecules else n_mol
    n_dia = n_mol - n_tio
    n_atom = n - n_mol
    pred = atlas_linedata(n_atom, wl0, wl1, seed)
    if n_dia:
        dia = atlas_linedata(n_dia, wl0, wl1, seed + 1000, species=tuple(molecules))
        pred = np.concatenate([pred, dia])
        pred = np.ascontiguousarray(pred[np.argsort(pred[:, 0], kind="stable")])
    tio = tio_mol2_records(n_tio, wl0, wl1, seed + 2000, loggf=tio_loggf)
    return pred, tio


# The other raw lists SELECTLINES reads (src/atlas12_modules.f90:19316-19781).
# h2opokazatel.bin: 8-byte records, int32 IWL then int16 IELO, int16 IGFLOG,
# whose signs carry the isotope (:19659-19680); h3plus.dat: 16-byte records.
H2O_DTYPE = np.dtype([("iwl", "<i4"), ("ielo", "<i2"), ("igflog", "<i2")])
H2O_NELION = 940
H3P_NELION = 895


def h2o_records(n: int, wl0: float, wl1: float, seed: int, loggf=(-4.0, 1.5),
                elo_ev: float = 0.5, n_zero_elo: int = 0) -> np.ndarray:
    """Seeded H2O records ascending in IWL: E_lo ~ Exp(elo_ev) in whole
    cm^-1, log gf ~ N(loggf), isotopes 1..4 (H2 16O, 17O, 18O, HD 16O) in
    proportion 7:1:1:1 through the signs of IELO and IGFLOG; n_zero_elo
    records (spread over the list) with E_lo = 0 and log gf = 3."""
    rng = np.random.default_rng(seed)
    lam = np.sort(np.exp(rng.uniform(math.log(wl0), math.log(wl1), n)))
    e = np.clip(np.rint(rng.exponential(elo_ev, n) * EV_TO_CM), 1, 32767).astype(np.int32)
    code = np.maximum(_tablog_code(rng.normal(loggf[0], loggf[1], n)).astype(np.int32), 1)
    iso = rng.choice(4, n, p=(0.7, 0.1, 0.1, 0.1)) + 1
    if n_zero_elo:   # strong lines, so that the selection keeps them
        z = rng.choice(n, n_zero_elo, replace=False)
        e[z] = 0
        code[z] = int(_tablog_code(3.0))
    rec = np.zeros(n, H2O_DTYPE)
    rec["iwl"] = np.rint(np.log(lam) / LOG_RATIO_2E6).astype(np.int32)
    rec["ielo"] = np.where((iso == 1) | (iso == 2), e, -e).astype(np.int16)
    rec["igflog"] = np.where((iso == 1) | (iso == 3), code, -code).astype(np.int16)
    return rec


def h3plus_records(n: int, wl0: float, wl1: float, seed: int, loggf=(0.0, 1.5)) -> np.ndarray:
    """Seeded H3+ records (16-byte LINEDATA layout, species field 8950)."""
    return atlas_linedata(n, wl0, wl1, seed, species=(H3P_NELION,),
                          loggf=("normal", loggf[0], loggf[1]), elo=("exp", 1.0))


# Kurucz molecule code, isotope code of each ASCII diatomic list the
# generator writes (molec_dispatch, src/mod_mklinelist.f90:2455-2586)
DIATOMIC_LISTS = {"ch.asc": ((106, 12), (106, 13)), "oh.asc": ((108, 16), (108, 18)),
                  "c2.asc": ((606, 12), (606, 33)), "cn.asc": ((607, 12), (607, 15)),
                  "co.asc": ((608, 12), (608, 13)), "nh.asc": ((107, 14),),
                  "alo.asc": ((813, 16),)}


def _kurucz_mol_text(rng, n: int, wl0: float, wl1: float, species, code_w: int, loggf,
                     elo_ev: float, dup_share: float) -> str:
    """One Kurucz-format molecular list (see diatomic_ascii_lists), the
    molecule code in a field code_w wide."""
    lam = np.exp(rng.uniform(math.log(wl0), math.log(wl1), n))
    elo = np.round(rng.exponential(elo_ev, n) * EV_TO_CM, 3)
    eup = np.round(elo + 1.0e7 / lam, 3)
    dup = np.nonzero(rng.random(n) < dup_share)[0]
    dup = dup[dup > 0]
    elo[dup], eup[dup] = elo[dup - 1], eup[dup - 1]
    wl = 1.0e7 / np.abs(eup - elo)
    order = np.argsort(wl, kind="stable")
    sp = rng.integers(0, len(species), n)
    gf = rng.normal(loggf[0], loggf[1], n)
    swap = rng.random(n) < 0.3
    neg = rng.random(n) < 0.1
    xj = rng.integers(0, 80, n) + 0.5
    cf = "{:%dd}" % code_w
    lines = []
    for k, i in enumerate(order):
        icode, iso = species[sp[i]]
        e, ep = (eup[i], elo[i]) if swap[i] else (elo[i], eup[i])
        if neg[i]:
            ep = -ep
        xjp = xj[i] + rng.integers(-1, 2)
        labp = ("X" if rng.random() < 0.33 else "A") + f"{int(rng.integers(0, 99)):02d}e"
        loggr = 0 if rng.random() < 0.25 else int(rng.integers(700, 900))
        lines.append(f"{wl[i]:10.4f}{gf[i]:7.3f}{xj[i]:5.1f}{e:10.3f}{max(xjp, 0.5):5.1f}"
                     f"{ep:11.3f}" + cf.format(icode) + f"{'A01e':8s}{labp:8s}{iso:2d}{loggr:4d}")
        if k == n // 3:
            lines.append("")
        if k == n // 2:
            lines.append("  bad line" + "?" * 60)
    lines.append(f"{wl1 * 1.5:10.4f}{-1.0:7.3f}{0.5:5.1f}{100.0:10.3f}{1.5:5.1f}"
                 f"{100.0 + 1.0e7 / wl1:11.3f}" + cf.format(species[0][0]) +
                 f"{'A':8s}{'X':8s}{species[0][1]:2d}{0:4d}")
    lines.append(lines[0])
    return "\n".join(lines) + "\n"


# the SYNTHE molecular lists of the reader fixture: FeH listed under code 156
# (read as 126) and AlO, which SYNTHE keeps and ATLAS12 does not
SYNTHE_MOL_LISTS = {"ch.asc": ((106, 12), (106, 13)), "co.asc": ((608, 12), (608, 13)),
                    "feh.asc": ((156, 56), (126, 57)), "alo.asc": ((813, 16), (813, 17)),
                    "cn.asc": ((607, 12), (607, 15), (607, 99))}


def diatomic_ascii_lists(n: int, wl0: float, wl1: float, seed: int, loggf=(-2.0, 1.5),
                         elo_ev: float = 0.6, dup_share: float = 0.05,
                         lists: dict | None = None) -> dict:
    """Seeded Kurucz-format ASCII molecular lists (columns F10.4, F7.3, F5.1,
    F10.3, F5.1, F11.3, I4, A8, A8, I2, I4: lambda, log gf, J, E, J', E',
    code, labels, isotope, 100 log gamma_rad), n lines per file of lists
    (default DIATOMIC_LISTS: file -> (code, isotope) pairs), each ascending
    in lambda.  Terms listed in either order, some with negative (predicted)
    energies; one in four with classical damping; 'X' upper labels for a
    third.  dup_share of the lines repeat a neighbour's terms (equal
    wavelengths: ties for the wavelength sort).  Each file also holds a
    blank line, a line the format rejects, and ends with a line beyond the
    window and one after it that must not be read.  Returns {file: text}."""
    rng = np.random.default_rng(seed)
    return {f: _kurucz_mol_text(rng, n, wl0, wl1, sp, 4, loggf, elo_ev, dup_share)
            for f, sp in (lists or DIATOMIC_LISTS).items()}


def polymol_ascii_list(n: int, wl0: float, wl1: float, seed: int, loggf=(-2.0, 1.5)) -> str:
    """A seeded polyatomic ASCII list (the molecular columns with a 6-digit
    code, I6): CaOH (10820) and one record in four of an unlisted code."""
    rng = np.random.default_rng(seed)
    return _kurucz_mol_text(rng, n, wl0, wl1, ((10820, 1), (10820, 2), (10820, 3), (10821, 1)),
                            6, loggf, 0.4, 0.02)


def polymol_bin_records(n: int, wl0: float, wl1: float, seed: int, icode: int = 10820,
                        loggf=(-2.0, 1.5)) -> np.ndarray:
    """A seeded polyatomic binary list, int32 (n + 1, 2): header (-1, code),
    then (IWL, int16 E_lo | int16 IGFLOG << 16) ascending in IWL, IGFLOG
    signs mixed (the reader takes |IGFLOG|)."""
    rng = np.random.default_rng(seed)
    lam = np.sort(np.exp(rng.uniform(math.log(wl0), math.log(wl1), n)))
    out = np.zeros((n + 1, 2), np.int32)
    out[0] = (-1, icode)
    out[1:, 0] = np.rint(np.log(lam) / LOG_RATIO_2E6).astype(np.int32)
    half = np.zeros((n, 2), np.int16)
    half[:, 0] = np.clip(np.rint(rng.exponential(0.4, n) * EV_TO_CM), 0, 32767).astype(np.int16)
    code = np.maximum(_tablog_code(rng.normal(loggf[0], loggf[1], n)).astype(np.int32), 1)
    half[:, 1] = np.where(rng.random(n) < 0.2, -code, code).astype(np.int16)
    out[1:, 1] = half.view(np.int32).reshape(n)
    return out


SYNTHE_MOL_KEYS = ("seed", "n_gfall", "n_mol", "n_poly", "n_polybin", "n_h2o", "span")


def synthe_mol_lists(seed: int, n_gfall: int, n_mol: int, n_poly: int, n_polybin: int,
                     n_h2o: int, span):
    """The raw lists of SYNTHE's molecular reader workload over span (nm):
    (gfall records, {molecular ASCII file: text} of SYNTHE_MOL_LISTS,
    polyatomic ASCII text, polyatomic binary words, H2O records); the
    binary lists reach 20 nm beyond span on each side."""
    w0, w1 = span
    g = synthetic_gfall(n_gfall, w0 + 1.5, w1 - 1.5, seed=seed)
    mol = diatomic_ascii_lists(n_mol, w0, w1, seed + 1, lists=SYNTHE_MOL_LISTS)
    poly = polymol_ascii_list(n_poly, w0, w1, seed + 2)
    pbin = polymol_bin_records(n_polybin, w0 - 20.0, w1 + 20.0, seed + 3)
    h2o = h2o_records(n_h2o, w0 - 20.0, w1 + 20.0, seed + 4, loggf=(-2.0, 1.5))
    return g, mol, poly, pbin, h2o


def atlas_all_lists(mols, seed: int, wl_range, n_pred: int, n_lowobs: int, n_hilines: int,
                    n_h2o: int, n_h3p: int, n_dia: int):
    """Every raw list SELECTLINES reads: the predicted and TiO lists (as
    atlas_iteration_lists, TiO log gf ~ N(4, 1.5)), observed low lines,
    high-excitation lines (one species in ten a slot, 1004, that stellar
    atmospheres leave empty), H2O (six with E_lo = 0), H3+ and the ASCII
    diatomic lists (text per file).  Returns (pred, tio, lowobs, hilines,
    h2o, h3p, {file: text})."""
    w0, w1 = wl_range
    s = seed
    pred, tio = atlas_iteration_lists(n_pred, w0, w1, seed=s, molecules=mols, tio_loggf=(4.0, 1.5))
    lowobs = atlas_linedata(n_lowobs, w0, w1, s + 1, loggf=("normal", -1.0, 1.5))
    hi_species = ATLAS_SPECIES + (atlas_nelion(26, 1),) * 2 + (1004,)
    hilines = atlas_linedata(n_hilines, w0, w1, s + 2, species=hi_species,
                             loggf=("normal", 0.0, 1.5), elo=("uniform", 4.0, 10.0))
    h2o = h2o_records(n_h2o, w0, w1, s + 3, loggf=(0.0, 1.5), n_zero_elo=6)
    h3p = h3plus_records(n_h3p, w0, w1, s + 4, loggf=(3.0, 1.5))
    dia = diatomic_ascii_lists(n_dia, w0, w1, s + 5)
    return pred, tio, lowobs, hilines, h2o, h3p, dia


# ---------------------------------------------------------------------------
# Batched SYNTHE: many models over one shared line list
# ---------------------------------------------------------------------------

@dataclass
class BatchModel:
    nrhox: int
    teff: float
    logg: float
    mh: float
    wl0: float
    wl1: float
    offset: int   # first grid point, relative to the shared grid
    length: int


def _depth_resample(a: np.ndarray, n: int) -> np.ndarray:
    """Resample a (nrhox, ...) array onto n layers spanning the same
    column-mass range (linear in layer index; logarithmic where positive)."""
    a = np.asarray(a, np.float64)
    src = np.linspace(0.0, 1.0, a.shape[0])
    dst = np.linspace(0.0, 1.0, n)
    flat = a.reshape(a.shape[0], -1)
    out = np.empty((n, flat.shape[1]))
    for c in range(flat.shape[1]):
        col = flat[:, c]
        if np.all(col > 0):
            out[:, c] = np.exp(np.interp(dst, src, np.log(col)))
        elif np.all(col == 0):
            out[:, c] = 0.0
        else:
            out[:, c] = np.interp(dst, src, col)
    return out.reshape((n,) + a.shape[1:])


def batch_model_state(base, m: BatchModel, grid: Grid, nedge: int = 8):
    """A synthetic model from a reference-derived state: layers resampled
    to m.nrhox, T and Doppler widths scaled with Teff, densities with log g
    and metals with [M/H], and a smooth continuum (log kappa linear in log
    lambda, 90 % absorption) tabulated on nedge edges over the model's
    window of the shared grid.  Not a converged atmosphere: a workload of
    the right shape and opacity level."""
    from .state import SyntheState
    n = m.nrhox
    ts = m.teff / 5777.0
    gs = 10.0 ** (m.logg - 4.44)
    r = lambda k: _depth_resample(getattr(base, k), n)  # noqa: E731
    xnfpel = r("xnfpel")
    xnfpel[:, 12:] *= 10.0 ** m.mh          # elements beyond H and He
    xnfpel *= gs
    wbegin = math.exp((grid.ixwlbeg + m.offset) * grid.ratiolg)
    wend = wbegin * grid.ratio ** (m.length - 1)
    wledge = np.exp(np.linspace(math.log(wbegin * 0.999), math.log(wend * 1.001), nedge))
    mid = base.continall[:, base.continall.shape[1] // 2]
    c0 = np.interp(np.linspace(0.0, 1.0, n), np.linspace(0.0, 1.0, base.nrhox), mid)
    c0 = c0 + 0.5 * (m.logg - 4.44) + 0.3 * m.mh
    cont = np.empty((n, 3 * (nedge - 1)))
    for e in range(nedge - 1):
        w0, w1 = wledge[e], wledge[e + 1]
        for q, w in enumerate((w0, 0.5 * (w0 + w1), w1)):
            cont[:, 3 * e + q] = c0 + 1.5 * math.log10(w / 500.0)
    return SyntheState(
        nrhox=n, length=m.length, ixwlbeg=grid.ixwlbeg + m.offset, wbegin=wbegin,
        ratio=grid.ratio, resolu=grid.resolu, wledge=wledge, continall=cont,
        contabs=cont + math.log10(0.9), contscat=cont + math.log10(0.1), xnfpel=xnfpel,
        dopple=r("dopple") * math.sqrt(ts), t=r("t") * ts, xne=r("xne") * math.sqrt(gs) * ts ** 4,
        rho=r("rho") * gs / ts, xnfh=r("xnfh") * gs, xnfhe1=r("xnfhe1") * gs,
        xnfh2=r("xnfh2") * gs, hckt=r("hckt") / ts, hkt_syn=r("hkt_syn") / ts,
        hkt_atl=r("hkt_atl") / ts, rhox=r("rhox"))


def synthe_batch_workload(base, nmodel: int = 256, nlines: int = 10_000_000, seed: int = 3,
                          wl_lo: float = 300.0, wl_hi: float = 1000.0,
                          resolu: float = 300_000.0):
    """The batched-SYNTHE workload: nmodel models with layers uniform in
    48-96, Teff 3500-8000 K, log g 1-5, [M/H] -2..+0.5, each with its own
    window 10-300 nm wide in wl_lo..wl_hi at resolu, over one list of
    nlines atomic lines (log gf ~ N(-3, 1.5), lower level exponential with a
    2 eV scale) sorted by centre on the shared grid.  Returns (grid, models,
    states, lines, nb_shift, line_lo, line_hi): model k takes
    lines[line_lo[k]:line_hi[k]] with nbuff - nb_shift[k]; the range is
    read_gfall's (centres within the window plus its 0.1 nm margin scaled
    by wlbeg / 500, src/mod_mklinelist.f90:846-847, :903-907)."""
    grid = Grid.make(wl_lo, wl_hi, resolu)
    rec = synthetic_gfall(nlines, wl_lo, wl_hi, seed, loggf=("normal", -3.0, 1.5),
                          elo=("exp", 2.0))
    lines = gfall_to_lte(rec, grid)
    wl = rec["wl"][(rec["wl"] >= grid.wlbeg - 0.1 * grid.wlbeg / 500.0)
                   & (rec["wl"] <= grid.wlend + 0.1 * grid.wlbeg / 500.0)].astype(np.float64)
    order = np.argsort(lines["nbuff"], kind="stable")
    lines, wl = lines[order], wl[order]
    rng = np.random.default_rng(seed + 1)
    models, states, shift, lo, hi = [], [], [], [], []
    for _ in range(nmodel):
        width = rng.uniform(10.0, 300.0)
        w0 = rng.uniform(wl_lo, wl_hi - width)
        ix0 = int(math.ceil(math.log(w0) / grid.ratiolg))
        length = int(math.log((w0 + width) / w0) / grid.ratiolg)
        m = BatchModel(int(rng.integers(48, 97)), rng.uniform(3500.0, 8000.0),
                       rng.uniform(1.0, 5.0), rng.uniform(-2.0, 0.5), w0, w0 + width,
                       ix0 - grid.ixwlbeg, length)
        models.append(m)
        states.append(batch_model_state(base, m, grid))
        margin = 0.1 * (m.wl0 / 500.0 if m.wl0 > 500.0 else 1.0)
        shift.append(m.offset)
        lo.append(int(np.searchsorted(wl, m.wl0 - margin)))
        hi.append(int(np.searchsorted(wl, m.wl1 + margin, side="right")))
    return grid, models, states, lines, np.array(shift), np.array(lo), np.array(hi)


# ---------------------------------------------------------------------------
# Large SYNTHE run: 50M lines, molecular band progressions, cool dwarf
# ---------------------------------------------------------------------------

# SYNTHE species slots of the diatomics in the xnfpelsyn molecule table
# (nelion = 6 (39 + position), src/mod_mklinelist.f90:1706, :2016-2025,
# :2467-2489): H2, CH, NH, C2, CN, CO; and their masses [amu].
LARGE_MOLECULES = ((240, 2.016), (246, 13.019), (252, 15.015), (264, 24.022), (270, 26.018),
                   (276, 28.010))


def cool_dwarf_state(base, grid: Grid, nrhox: int = 72, t_top: float = 2800.0):
    """The plain-SYNTHE state turned into a cool dwarf (T rising from t_top
    to t_top * 9500/4400, log g 5) on the given window, with the diatomics
    of LARGE_MOLECULES populated: number densities a fixed fraction of
    hydrogen falling with temperature, Doppler widths from their masses and
    the atmosphere's turbulence.  A workload of the right shape, not a
    converged model."""
    teff = 5777.0 * t_top / 4400.0
    m = BatchModel(nrhox, teff, 5.0, 0.0, grid.wlbeg, grid.wlend, 0, grid.length)
    st = batch_model_state(base, m, grid, nedge=16)
    k_b, amu, c = 1.380649e-16, 1.66053907e-24, 2.99792458e10
    vt2 = (1.5e5) ** 2
    for q, (slot, mass) in enumerate(LARGE_MOLECULES):
        frac = 10.0 ** (-3.0 - 0.5 * q) * (4000.0 / st.t) ** 6
        st.xnfpel[:, slot - 1] = frac * st.xnfh
        st.dopple[:, slot - 1] = np.sqrt(2.0 * k_b * st.t / (mass * amu) + vt2) / c
    return st


def large_synthe_lines(grid: Grid, n: int = 50_000_000, seed: int = 7,
                       atomic_share: float = 0.3, band_len: int = 200) -> np.ndarray:
    """n lte_line_t records over the grid, sorted by nbuff: atomic_share
    atomic lines (SYNTHE_SPECIES, log gf uniform in [-5, 0], lower level
    uniform in 0-8 eV) and the rest in molecular band progressions of
    band_len lines each -- a band origin uniform in ln(lambda), lines every
    2e-5 in lambda (10 grid points at R = 500,000), log gf ~ N(-5, 1), lower
    level B J (J + 1) with B = 1.5 cm^-1 over a vibrational term uniform in
    0-1 eV -- on the diatomics of LARGE_MOLECULES.  Built as read_gfall
    forms records (nbuff, cgf = 0.026538/1.77245 gf / nu, damping over
    12.5664 nu, src/mod_mklinelist.f90:898-1005), vectorised."""
    rng = np.random.default_rng(seed)
    na = int(n * atomic_share)
    nb = (n - na) // band_len
    nm = nb * band_len
    lw0, lw1 = math.log(grid.wlbeg), math.log(grid.wlend)
    wl = np.empty(na + nm)
    wl[:na] = np.exp(rng.uniform(lw0, lw1, na))
    origin = np.exp(rng.uniform(lw0, lw1 - band_len * 2.0e-5, nb))
    jj = np.arange(band_len, dtype=np.float64)
    wl[na:] = (origin[:, None] * (1.0 + 2.0e-5 * jj[None, :])).ravel()
    loggf = np.empty(na + nm, np.float32)
    loggf[:na] = rng.uniform(-5.0, 0.0, na)
    loggf[na:] = rng.normal(-5.0, 1.0, nm)
    e_cm = np.empty(na + nm)
    e_cm[:na] = rng.uniform(0.0, 8.0, na) * EV_TO_CM
    evib = rng.uniform(0.0, 1.0, nb) * EV_TO_CM
    e_cm[na:] = (evib[:, None] + 1.5 * jj[None, :] * (jj[None, :] + 1.0)).ravel()
    nel = np.empty(na + nm, np.int32)
    sp = np.asarray(SYNTHE_SPECIES)[rng.integers(0, len(SYNTHE_SPECIES), na)]
    nelem = np.floor(sp).astype(np.int32)
    nel[:na] = nelem * 6 - 6 + np.rint((sp - nelem) * 100.0).astype(np.int32) + 1
    slots = np.array([s for s, _ in LARGE_MOLECULES], np.int32)
    nel[na:] = np.repeat(slots[rng.integers(0, len(slots), nb)], band_len)
    code = np.empty(na + nm, np.float32)
    code[:na] = sp
    code[na:] = 100.0 + (nel[na:] // 6 - 39)  # a molecule tag; not read on the hot path
    lg = rng.uniform(7.0, 9.0, na + nm), rng.uniform(-6.0, -4.0, na + nm), \
        rng.uniform(-8.0, -7.0, na + nm)
    keep = (wl >= grid.wlbeg) & (wl <= grid.wlend)
    idx = np.flatnonzero(keep)
    ixwl = np.floor(np.log(wl[idx]) / grid.ratiolg + 0.5).astype(np.int64)
    nbuff = (ixwl - grid.ixwlbeg + 1).astype(np.int32)
    order = np.argsort(nbuff, kind="stable")
    idx, nbuff = idx[order], nbuff[order]
    out = np.zeros(len(idx), dtype=LTE_DTYPE)
    out["nbuff"] = nbuff
    frelin = 2.99792458e17 / wl[idx]
    gf = np.power(np.float32(10.0), loggf[idx]).astype(np.float64)
    out["cgf"] = (0.026538 / 1.77245 * gf / frelin).astype(np.float32)
    out["nelion"] = nel[idx]
    out["elo"] = e_cm[idx].astype(np.float32)
    frq4pi = 12.5664 * frelin
    for g, dst in zip(lg, ("gamrf", "gamsf", "gamwf")):
        out[dst] = (np.power(10.0, g[idx]) / frq4pi).astype(np.float32)
    out["code"] = code[idx]
    return out


def kapp_standin(z, numnu):
    """KAPP outputs at every frequency from the reference's values at every
    100th: log-linear in frequency index where a table is positive, linear
    where it holds zeros."""
    sub = np.asarray(z["kapp_sub"], np.float64)
    x = np.arange(numnu, dtype=np.float64)
    out = []
    for k in ("acont", "sigmac", "scont", "sline", "aline", "sigmal"):
        s = np.asarray(z[k + "_sub"])  # (nsub, nrhox)
        if np.all(s > 0):
            f = np.exp(np.stack([np.interp(x, sub, np.log(s[:, j])) for j in range(s.shape[1])], 1))
        else:
            f = np.stack([np.interp(x, sub, s[:, j]) for j in range(s.shape[1])], 1)
        out.append(np.ascontiguousarray(f))
    return out


def kapp_deck_state(ref, deck: str, mion: int = 1006) -> dict:
    """KAPP's depth inputs of one deck of a KAPP fixture (the layout of
    tests/golden/kapp_ref.npz: per-depth arrays, BHYD (6, nrhox), and the
    XNF/XNFP columns the sources read, listed in xnf_ids/xnfp_ids) as full
    (mion, nrhox) arrays, zero in the columns the sources do not read."""
    n = len(ref[deck + "_t"])
    st = {k: np.asarray(ref[f"{deck}_{k}"]) for k in ("t", "tkev", "tlog", "hkt", "hckt", "xne", "rho",
                                                      "bmin", "equilh2", "bhyd")}
    xnf = np.zeros((mion, n))
    xnf[np.asarray(ref["xnf_ids"]) - 1] = ref[deck + "_xnf"]
    xnfp = np.zeros((mion, n))
    xnfp[np.asarray(ref["xnfp_ids"]) - 1] = ref[deck + "_xnfp"]
    st["xnf"], st["xnfp"] = xnf, xnfp
    return st
