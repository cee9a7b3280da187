"""ctypes wrapper of the CPU oracle (liba12oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, as the checker.  The product library in
atlas12_amd/ never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "liba12oracle.so")

P_D = C.POINTER(C.c_double)
P_F = C.POINTER(C.c_float)


class OrLine(C.Structure):
    _fields_ = [("nbuff", C.c_int32), ("cgf", C.c_float), ("nelion", C.c_int32),
                ("elo", C.c_float), ("gamrf", C.c_float), ("gamsf", C.c_float),
                ("gamwf", C.c_float), ("code", C.c_float)]


class OrJoshIO(C.Structure):
    _fields_ = [("nrhox", C.c_int)] + [(k, P_D) for k in (
        "acont", "aline", "sigmac", "sigmal", "scont", "sline", "rhox", "bnu",
        "coefj", "coefh", "taunu", "snu", "jnu", "hnu", "jmins")] + [
        ("knu1", C.c_double), ("maxj", C.c_int), ("vef_sweeps", C.c_int),
        ("vef_unconverged", C.c_int), ("pivot_floor_hits", C.c_int)]


class OrSyntheAtm(C.Structure):
    _fields_ = [("nrhox", C.c_int), ("length", C.c_int), ("wbegin", C.c_double),
                ("ratio", C.c_double), ("resolu", C.c_double), ("nedge", C.c_int),
                ("wledge", P_D), ("nf", C.c_int), ("continall", P_D), ("mw6", C.c_int),
                ("xnfpel", P_D), ("dopple", P_D)] + [(k, P_D) for k in (
                    "rho", "t", "xne", "xnfh", "xnfhe1", "xnfh2", "hckt")]


class OrSyntheRT(C.Structure):
    _fields_ = [("nrhox", C.c_int), ("length", C.c_int), ("wbegin", C.c_double),
                ("ratio", C.c_double), ("nedge", C.c_int), ("wledge", P_D), ("nf", C.c_int),
                ("contabs", P_D), ("contscat", P_D), ("hkt_syn", P_D), ("hkt_atl", P_D),
                ("rhox", P_D), ("coefj", P_D), ("coefh", P_D)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            build()
        L = C.CDLL(LIBPATH)
        L.or_voigt_ref.restype = C.c_float
        L.or_voigt_ref.argtypes = [C.c_float, C.c_float]
        L.or_voigt_f64.restype = C.c_double
        L.or_voigt_f64.argtypes = [C.c_double, C.c_double]
        L.or_map1.restype = C.c_int
        L.or_solvit.restype = C.c_int
        L.or_expi.restype = C.c_double
        L.or_expi.argtypes = [C.c_int, C.c_double]
        L.or_josh.restype = C.c_double
        L.or_omp_threads.restype = C.c_int
        _lib = L
    return _lib


def _p(a, t=P_D):
    return a.ctypes.data_as(t)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------- numerics
def voigt_ref(v, a):
    L = lib()
    return np.array([L.or_voigt_ref(float(x), float(y)) for x, y in zip(v, a)], np.float32)


def voigt_f64(v, a):
    L = lib()
    return np.array([L.or_voigt_f64(float(x), float(y)) for x, y in zip(v, a)])


def solvit(a, b):
    a = np.asfortranarray(a, dtype=np.float64).copy(order="F")
    b = _f64(b).copy()
    lib().or_solvit(a.ctypes.data_as(P_D), C.c_int(len(b)), _p(b))
    return b


def map1(xold, fold, xnew):
    xold, fold, xnew = _f64(xold), _f64(fold), _f64(xnew)
    out = np.zeros(len(xnew))
    r = lib().or_map1(_p(xold), _p(fold), C.c_int(len(xold)), _p(xnew), _p(out),
                      C.c_int(len(xnew)))
    return out, r


def integ(x, f, start):
    x, f = _f64(x), _f64(f)
    out = np.zeros(len(x))
    lib().or_integ(_p(x), _p(f), _p(out), C.c_int(len(x)), C.c_double(start))
    return out


def parcoe(f, x):
    f, x = _f64(f), _f64(x)
    a, b, c = np.zeros(len(x)), np.zeros(len(x)), np.zeros(len(x))
    lib().or_parcoe(_p(f), _p(x), _p(a), _p(b), _p(c), C.c_int(len(x)))
    return a, b, c


def deriv(x, f):
    x, f = _f64(x), _f64(f)
    out = np.zeros(len(x))
    lib().or_deriv(_p(x), _p(f), _p(out), C.c_int(len(x)))
    return out


def expi(n, x):
    return lib().or_expi(int(n), float(x))


def josh(ifscat, ifsurf, acont, aline, sigmac, sigmal, scont, sline, rhox, bnu, coefj, coefh):
    n = len(rhox)
    keep = [_f64(a) for a in (acont, aline, sigmac, sigmal, scont, sline, rhox, bnu, coefj, coefh)]
    outs = [np.zeros(n) for _ in range(5)]
    io = OrJoshIO()
    io.nrhox = n
    for k, a in zip(("acont", "aline", "sigmac", "sigmal", "scont", "sline", "rhox", "bnu",
                     "coefj", "coefh"), keep):
        setattr(io, k, _p(a))
    for k, a in zip(("taunu", "snu", "jnu", "hnu", "jmins"), outs):
        setattr(io, k, _p(a))
    h1 = lib().or_josh(C.byref(io), C.c_int(ifscat), C.c_int(ifsurf))
    return dict(hnu1=h1, knu1=io.knu1, taunu=outs[0], snu=outs[1], jnu=outs[2], hnu=outs[3],
                jmins=outs[4], maxj=io.maxj, vef_sweeps=io.vef_sweeps,
                vef_unconverged=io.vef_unconverged, pivot_floor_hits=io.pivot_floor_hits)


# ---------------------------------------------------------------- SYNTHE
class _Keep:
    """Holds numpy buffers alive for the lifetime of a ctypes struct."""


def _atm(st) -> tuple:
    k = _Keep()
    k.arrs = {n: _f64(getattr(st, n)) for n in ("wledge", "continall", "xnfpel", "dopple", "rho",
                                               "t", "xne", "xnfh", "xnfhe1", "xnfh2", "hckt")}
    a = OrSyntheAtm()
    a.nrhox, a.length = st.nrhox, st.length
    a.wbegin, a.ratio, a.resolu = st.wbegin, st.ratio, st.resolu
    a.nedge = len(st.wledge)
    a.nf = st.continall.shape[1]
    a.mw6 = st.xnfpel.shape[1]
    for n, arr in k.arrs.items():
        setattr(a, n, _p(arr))
    return a, k


def synthe_continuum(st, j):
    a, keep = _atm(st)
    out = np.zeros(st.length, np.float32)
    out64 = np.zeros(st.length)
    lib().or_synthe_continuum(C.byref(a), C.c_int(j), _p(out, P_F), _p(out64))
    return out, out64


def synthe_line_opacity(st, lines, prec="ref", want_nstep=False):
    """Opacity matrix (nrhox, length); prec 'ref' (FP32, reference types) or
    'f64'.  Returns (matrix, evals[, nstep])."""
    a, keep = _atm(st)
    lines = np.ascontiguousarray(lines)
    nl = len(lines)
    ev = C.c_int64(0)
    ns = np.zeros((st.nrhox, nl), np.int32) if want_nstep else None
    if prec == "ref":
        out = np.zeros((st.nrhox, st.length), np.float32)
        lib().or_synthe_line_opacity(C.byref(a), lines.ctypes.data_as(C.c_void_p), C.c_int(nl),
                                     C.c_int(0), _p(out, P_F), None, C.byref(ev),
                                     ns.ctypes.data_as(C.c_void_p) if want_nstep else None)
    else:
        out = np.zeros((st.nrhox, st.length))
        lib().or_synthe_line_opacity(C.byref(a), lines.ctypes.data_as(C.c_void_p), C.c_int(nl),
                                     C.c_int(1), None, _p(out), C.byref(ev),
                                     ns.ctypes.data_as(C.c_void_p) if want_nstep else None)
    if want_nstep:
        return out, ev.value, ns
    return out, ev.value


def synthe_depth_split(st, lines, j, prec="f64", threads=16):
    """Opacity of one depth j over all points with the lines split across
    threads (or_synthe_depth_split): (row (length,), evals).  prec 'f64':
    FP64; 'ref': the reference's FP32 arithmetic.  Every profile value and
    the count are the serial routine's; only the order of the sums differs."""
    a, keep = _atm(st)
    lines = np.ascontiguousarray(lines)
    f64 = prec != "ref"
    row = np.zeros(st.length, np.float64 if f64 else np.float32)
    ev = C.c_int64(0)
    L = lib()
    L.or_omp_threads(int(threads))
    L.or_synthe_depth_split(C.byref(a), lines.ctypes.data_as(C.c_void_p), C.c_int(len(lines)),
                            C.c_int(1 if f64 else 0), C.c_int(int(j)),
                            None if f64 else _p(row, P_F), _p(row) if f64 else None,
                            C.byref(ev))
    return row, ev.value


def synthe_depths(st, lines, js, prec="f64", threads=16):
    """Opacity rows of the depths js (one depth per thread, each the serial
    routine): (rows (len(js), length), evals over those depths)."""
    a, keep = _atm(st)
    lines = np.ascontiguousarray(lines)
    idx = np.ascontiguousarray(js, np.int32)
    ev = C.c_int64(0)
    L = lib()
    L.or_omp_threads(int(threads))
    if prec == "ref":
        out = np.zeros((st.nrhox, st.length), np.float32)
        args = (_p(out, P_F), None)
    else:
        out = np.zeros((st.nrhox, st.length))
        args = (None, _p(out))
    L.or_synthe_line_opacity_omp_list(C.byref(a), lines.ctypes.data_as(C.c_void_p),
                                      C.c_int(len(lines)), C.c_int(0 if prec == "ref" else 1),
                                      idx.ctypes.data_as(C.c_void_p), C.c_int(len(idx)), *args,
                                      C.byref(ev))
    return out[idx], ev.value


def _rt(st, coefj, coefh):
    k = _Keep()
    k.arrs = {n: _f64(v) for n, v in (("wledge", st.wledge), ("contabs", st.contabs),
                                      ("contscat", st.contscat), ("hkt_syn", st.hkt_syn),
                                      ("hkt_atl", st.hkt_atl), ("rhox", st.rhox),
                                      ("coefj", coefj), ("coefh", coefh))}
    r = OrSyntheRT()
    r.nrhox, r.length, r.wbegin, r.ratio = st.nrhox, st.length, st.wbegin, st.ratio
    r.nedge = len(st.wledge)
    r.nf = st.contabs.shape[1]
    for n, arr in k.arrs.items():
        setattr(r, n, _p(arr))
    return r, k


def synthe_transfer(st, opac, coefj, coefh, i0=0, i1=None, threads=1):
    """H_line, H_cont for wavelength points [i0, i1) from the opacity matrix
    (nrhox, length) FP32 (reference) or FP64."""
    i1 = st.length if i1 is None else i1
    r, keep = _rt(st, coefj, coefh)
    hl = np.zeros(i1 - i0)
    hc = np.zeros(i1 - i0)
    opac = np.ascontiguousarray(opac)
    if opac.dtype == np.float32:
        args = (_p(opac, P_F), None)
    else:
        opac = _f64(opac)
        args = (None, _p(opac))
    L = lib()
    if threads != 1:
        L.or_omp_threads(int(threads))
        L.or_synthe_transfer_omp(C.byref(r), *args, _p(hl), _p(hc), C.c_int(i0), C.c_int(i1))
    else:
        L.or_synthe_transfer(C.byref(r), *args, _p(hl), _p(hc), C.c_int(i0), C.c_int(i1))
    return hl, hc


def synthe_taunu(st, opac, coefj, coefh, i0=0, i1=None):
    """TAUNU (i1 - i0, nrhox) of the line-included JOSH call at wavelength
    points [i0, i1), the table synthe.exe writes to .linform, from the
    opacity matrix (nrhox, length) FP32 (reference) or FP64."""
    i1 = st.length if i1 is None else i1
    r, keep = _rt(st, coefj, coefh)
    out = np.zeros((i1 - i0, st.nrhox))
    opac = np.ascontiguousarray(opac)
    if opac.dtype == np.float32:
        args = (_p(opac, P_F), None)
    else:
        opac = _f64(opac)
        args = (None, _p(opac))
    lib().or_synthe_taunu(C.byref(r), *args, _p(out), C.c_int(i0), C.c_int(i1))
    return out


# ---------------------------------------------------------------- ATLAS12
class OrAtlasLinesIn(C.Structure):
    _fields_ = [("nrhox", C.c_int), ("numnu", C.c_int), ("nlines", C.c_int), ("mion", C.c_int),
                ("ldion", C.c_int), ("nwave1", C.c_int), ("waveset", P_D),
                ("linedata", C.c_void_p), ("tabcont", P_D), ("iwavetab", C.c_void_p),
                ("xnfdop", P_D), ("dopple", P_D)] + [(k, P_D) for k in (
                    "t", "xne", "hckt", "xnf1", "xnf3", "xnf841", "h0", "h1", "h2")]


class OrRtIn(C.Structure):
    _fields_ = [("nrhox", C.c_int), ("nfreq", C.c_int), ("teff", C.c_double),
                ("wave", P_D), ("rcowt", P_D), ("t", P_D), ("hkt", P_D), ("rhox", P_D)] + [
        (k, P_D) for k in ("acont", "sigmac", "scont", "sline", "aline", "sigmal", "xlines",
                           "coefj", "coefh")]


class OrRtAcc(C.Structure):
    _fields_ = [(k, C.c_double * 128) for k in ("rjmins", "rdabh", "rdiagj", "flxrad", "raden",
                                                "h_total", "accrad", "abross")] + [
        ("pradk0", C.c_double)] + [(k, C.c_double * 128) for k in (
            "last_rdiagjnu", "last_hnu", "last_taunu")]


def tabvoigt(vsteps=200.0, n=2001):
    h = [np.zeros(n) for _ in range(3)]
    lib().or_tabvoigt(C.c_double(vsteps), C.c_int(n), *[_p(a) for a in h])
    return h


def voigt_kurucz(v, a, tabs):
    L = lib()
    L.or_voigt_kurucz.restype = C.c_double
    h0, h1, h2 = [_f64(t) for t in tabs]
    return np.array([L.or_voigt_kurucz(C.c_double(x), C.c_double(y), _p(h0), _p(h1), _p(h2))
                     for x, y in zip(v, a)])


def full_ion(sub, idx, mion, n):
    """(mion, n) array with the given species rows filled."""
    out = np.zeros((mion, n))
    out[idx] = sub
    return out


def linop1(z, threads: int = 1):
    """LINOP1 on an ATLAS state dict (as in tests/golden/atlas_small.npz).
    Returns (xlines (numnu, nrhox), evals).  threads > 1: the OpenMP
    timing port (lines in contiguous chunks per thread; different sum
    order)."""
    n = int(len(z["t"]))
    numnu = int(len(z["waveset"]))
    mion = int(z["mion"])
    k = _Keep()
    k.a = {"waveset": _f64(z["waveset"]), "tabcont": _f64(z["tabcont"]),
           "xnfdop": _f64(full_ion(z["xnfdop"], z["ion_idx"], mion, n)),
           "dopple": _f64(full_ion(z["dopple"], z["ion_idx"], mion, n))}
    for kk in ("t", "xne", "hckt", "xnf1", "xnf3", "xnf841", "h0", "h1", "h2"):
        k.a[kk] = _f64(z[kk])
    ld = np.ascontiguousarray(z["linedata"], np.int32)
    iw = np.ascontiguousarray(z["iwavetab"], np.int32)
    s = OrAtlasLinesIn()
    s.nrhox, s.numnu, s.nlines, s.mion, s.ldion = n, numnu, len(ld), mion, n
    s.nwave1 = len(iw)
    s.linedata = ld.ctypes.data_as(C.c_void_p)
    s.iwavetab = iw.ctypes.data_as(C.c_void_p)
    for kk, v in k.a.items():
        setattr(s, kk, _p(v))
    out = np.zeros((numnu, n))
    ev = C.c_int64(0)
    if threads > 1:
        lib().or_linop1_omp(C.byref(s), _p(out), C.byref(ev))
    else:
        lib().or_linop1(C.byref(s), _p(out), C.byref(ev))
    return out, ev.value


def rt_pass(z, xlines, coefj, coefh):
    """RUN_RT_PASS's frequency loop on tabulated KAPP outputs.  Returns
    (accumulator dict, hnu1 per frequency)."""
    n = int(len(z["t"]))
    nf = int(len(z["waveset"]))
    keep = {k: _f64(z[k]) for k in ("waveset", "rcoset", "t", "hkt", "rhox", "acont", "sigmac",
                                    "scont", "sline", "aline", "sigmal")}
    keep["xlines"] = _f64(xlines)
    keep["coefj"], keep["coefh"] = _f64(coefj), _f64(coefh)
    s = OrRtIn()
    s.nrhox, s.nfreq, s.teff = n, nf, float(z["teff"])
    s.wave = _p(keep["waveset"])
    s.rcowt = _p(keep["rcoset"])
    for kk in ("t", "hkt", "rhox", "acont", "sigmac", "scont", "sline", "aline", "sigmal",
               "xlines", "coefj", "coefh"):
        setattr(s, kk, _p(keep[kk]))
    acc = OrRtAcc()
    h1 = np.zeros(nf)
    lib().or_rt_pass(C.byref(s), C.byref(acc), _p(h1))
    out = {k: np.array(getattr(acc, k)[:n]) for k in ("rjmins", "rdabh", "rdiagj", "flxrad",
                                                      "raden", "h_total", "accrad", "abross",
                                                      "last_rdiagjnu", "last_hnu", "last_taunu")}
    out["pradk0"] = acc.pradk0
    return out, h1


class OrSelectIn(C.Structure):
    _fields_ = [("nrhox", C.c_int), ("mion", C.c_int), ("ldion", C.c_int), ("nwave", C.c_int),
                ("teff", C.c_double), ("tabcont", P_D), ("iwavetab", C.c_void_p),
                ("wavetab", P_D), ("xnfdop", P_D), ("hckt", P_D)]


def atlas_select(z, pred, mol2):
    """SELECTLINES' filter on predicted records (n, 4) int32 and MOL2
    records (atlas12_amd.workloads.MOL2_DTYPE) with the state in z (tabcont,
    iwavetab, wavetab, xnfdop rows ion_idx, hckt, teff, mion).  Returns the
    kept store (m, 4) int32."""
    n = int(len(z["hckt"]))
    mion = int(z["mion"])
    k = _Keep()
    k.a = {"tabcont": _f64(z["tabcont"]), "wavetab": _f64(z["wavetab"]),
           "xnfdop": _f64(full_ion(z["xnfdop"], z["ion_idx"], mion, n)), "hckt": _f64(z["hckt"])}
    iw = np.ascontiguousarray(z["iwavetab"], np.int32)
    pr = np.ascontiguousarray(pred, np.int32)
    mo = np.ascontiguousarray(mol2)
    assert mo.dtype.itemsize == 32
    s = OrSelectIn()
    s.nrhox, s.mion, s.ldion, s.nwave = n, mion, n, len(z["wavetab"])
    s.teff = float(z["teff"])
    s.iwavetab = iw.ctypes.data_as(C.c_void_p)
    for kk, v in k.a.items():
        setattr(s, kk, _p(v))
    out = np.zeros((len(pr) + len(mo), 4), np.int32)
    L = lib()
    L.or_atlas_select.restype = C.c_int64
    m = L.or_atlas_select(C.byref(s), pr.ctypes.data_as(C.c_void_p), C.c_int64(len(pr)),
                          mo.ctypes.data_as(C.c_void_p), C.c_int64(len(mo)),
                          out.ctypes.data_as(C.c_void_p))
    if m < 0:
        raise ValueError("or_atlas_select: NELION outside 1..mion")
    return out[:m]


class OrSelectLists(C.Structure):
    _fields_ = [("pred", C.c_void_p), ("npred", C.c_int64), ("lowobs", C.c_void_p),
                ("nlowobs", C.c_int64), ("hilines", C.c_void_p), ("nhilines", C.c_int64),
                ("dia", C.c_void_p), ("ndia", C.c_int64), ("mol2", C.c_void_p),
                ("nmol2", C.c_int64), ("h2o", C.c_void_p), ("nh2o", C.c_int64),
                ("h3p", C.c_void_p), ("nh3p", C.c_int64)]


SELECT_LISTS = ("pred", "lowobs", "hilines", "dia", "mol2", "h2o", "h3p")


def atlas_select_all(z, lists: dict):
    """SELECTLINES over all seven lists: lists maps SELECT_LISTS names to
    records (16-byte lists as (n, 4) int32, dia as DIATOMIC_DTYPE, mol2 as
    MOL2_DTYPE, h2o as H2O_DTYPE); absent ones are empty.  z: tabcont,
    iwavetab, wavetab, hckt, teff, mion, and xnfdop (mion, n) or its rows
    ion_idx.  Returns the kept store (m, 4) int32."""
    n = int(len(z["hckt"]))
    mion = int(z["mion"])
    xd = z["xnfdop"] if "ion_idx" not in z else full_ion(z["xnfdop"], z["ion_idx"], mion, n)
    k = _Keep()
    k.a = {"tabcont": _f64(z["tabcont"]), "wavetab": _f64(z["wavetab"]),
           "xnfdop": _f64(xd), "hckt": _f64(z["hckt"])}
    iw = np.ascontiguousarray(z["iwavetab"], np.int32)
    s = OrSelectIn()
    s.nrhox, s.mion, s.ldion, s.nwave = n, mion, n, len(z["wavetab"])
    s.teff = float(z["teff"])
    s.iwavetab = iw.ctypes.data_as(C.c_void_p)
    for kk, v in k.a.items():
        setattr(s, kk, _p(v))
    sl = OrSelectLists()
    keep, total = [], 0
    for name in SELECT_LISTS:
        a = lists.get(name)
        if a is None or len(a) == 0:
            continue
        a = np.ascontiguousarray(a)
        keep.append(a)
        setattr(sl, name, a.ctypes.data_as(C.c_void_p))
        setattr(sl, "n" + name, len(a))
        total += len(a)
    out = np.zeros((max(total, 1), 4), np.int32)
    L = lib()
    L.or_atlas_select_all.restype = C.c_int64
    m = L.or_atlas_select_all(C.byref(s), C.byref(sl), out.ctypes.data_as(C.c_void_p))
    if m == -1:
        raise ValueError("or_atlas_select_all: NELION outside 1..mion")
    if m == -2:
        raise ValueError("or_atlas_select_all: diatomic MOLCODE outside its table")
    return out[:m]


def sort_indirect_by_int(keys):
    """(sorted keys, 1-based order) of SELECTLINES' heapsort."""
    kk = np.ascontiguousarray(keys, np.int32).copy()
    order = np.arange(1, len(kk) + 1, dtype=np.int32)
    lib().or_sort_indirect_by_int(kk.ctypes.data_as(C.c_void_p), order.ctypes.data_as(C.c_void_p),
                                  C.c_int64(len(kk)))
    return kk, order


def scale_model(n, kw, tau1lg, steplg, sc, teff_new, logg_new, fields):
    """SCALE_MODEL on one model.  fields: (7, n) RHOX, T, P, XNE, ABROSS,
    PRAD, VTURB; sc: TEFF, GRAV, PRADK0, PCON, PTURB0.  Returns (out (10, kw):
    TAUSTD, the 7 fields, PRADK, PTOTAL; so (6): TEFF, GRAV, GLOG, FLUX,
    PRADK0, PZERO; rescaled flag)."""
    fin = [_f64(f) for f in fields]
    out = np.zeros((10, kw))
    so = np.full(6, np.nan)
    scv = _f64(sc)
    ins = (P_D * 7)(*[_p(f) for f in fin])
    outs = (P_D * 10)(*[_p(out[i]) for i in range(10)])
    L = lib()
    L.or_scale_model.restype = C.c_int
    r = L.or_scale_model(C.c_int(n), C.c_int(kw), C.c_double(tau1lg), C.c_double(steplg),
                         _p(scv), C.c_double(teff_new), C.c_double(logg_new), ins, outs, _p(so))
    return out, so, bool(r)


TCORR_FIELDS = ("tauros", "rhox", "t", "dltdlp", "abross", "flxcnv", "flxcnv0", "grdadb",
                "ptotal", "rho", "dlrdlt", "heatcp", "hscale", "flxrad", "rjmins", "rdiagj")


class OrTcorrIn(C.Structure):
    _fields_ = [("n", C.c_int), ("ifconv", C.c_int), ("iter", C.c_int), ("flux", C.c_double),
                ("teff", C.c_double), ("mixlth", C.c_double)] + [(k, P_D) for k in TCORR_FIELDS]


def tcorr3(z, rdabh, oldt1=None, iter_=None):
    """TCORR(3) sections A-F on a state dict (keys as tests/golden/
    tcorr_ref.npz).  Returns (out dict T1, DTFLUX, DTLAMB, DTSURF, FLXERR,
    FLXDRV, HRATIO; updated rdabh; updated oldt1)."""
    n = int(z["n"])
    keep = {k: _f64(z[k]) for k in TCORR_FIELDS}
    s = OrTcorrIn()
    s.n, s.ifconv = n, int(z["ifconv"])
    s.iter = int(z["iter"]) if iter_ is None else int(iter_)
    s.flux, s.teff, s.mixlth = float(z["flux"]), float(z["teff"]), float(z["mixlth"])
    for k in TCORR_FIELDS:
        setattr(s, k, _p(keep[k]))
    rd = _f64(np.array(rdabh, np.float64).copy())
    ot = _f64(np.zeros(n) if oldt1 is None else np.array(oldt1, np.float64).copy())
    out = np.zeros((7, n))
    lib().or_tcorr3(C.byref(s), _p(rd), _p(ot), _p(out))
    names = ("t1", "dtflux", "dtlamb", "dtsurf", "flxerr", "flxdrv", "hratio")
    return {k: out[i] for i, k in enumerate(names)}, rd, ot


ROSSTAB_MAXTAB = 80 * 60  # kw * 60, src/atlas12_modules.f90:1779


class OrRosstab(C.Structure):
    _fields_ = [("nross", C.c_int), ("maxtab", C.c_int), ("zerot", C.c_double),
                ("slopet", C.c_double), ("zerop", C.c_double), ("slopep", C.c_double),
                ("tabt", P_D), ("tabp", P_D), ("ross", P_D)]


class Rosstab:
    """ROSSTAB's table (the module state ROSSTAB_*): store and lookup."""

    def __init__(self, zero=None, tab=None, maxtab=ROSSTAB_MAXTAB):
        self.a = np.zeros((3, maxtab))
        self.s = OrRosstab()
        self.s.maxtab = maxtab
        self.s.nross = 0
        if tab is not None:
            tab = np.asarray(tab, np.float64)
            self.a[:, :tab.shape[1]] = tab
            self.s.nross = tab.shape[1]
            self.s.zerot, self.s.slopet, self.s.zerop, self.s.slopep = (float(x) for x in zero)
        self.s.tabt, self.s.tabp, self.s.ross = (_p(self.a[i]) for i in range(3))

    @property
    def nross(self):
        return int(self.s.nross)

    def table(self):
        return self.a[:, :self.nross].copy()

    def zero(self):
        return np.array([self.s.zerot, self.s.slopet, self.s.zerop, self.s.slopep])

    def store(self, t, p, abross):
        t, p, ab = _f64(t), _f64(p), _f64(abross)
        return lib().or_rosstab_store(C.byref(self.s), len(t), _p(t), _p(p), _p(ab))

    def lookup(self, temp, pressure, mode=1):
        L = lib()
        L.or_rosstab_lookup.restype = C.c_double
        L.or_rosstab_lookup.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double]
        return np.array([L.or_rosstab_lookup(C.addressof(self.s), int(mode), float(t), float(q))
                         for t, q in zip(np.atleast_1d(temp), np.atleast_1d(pressure))])


def ttaup(tab, t, tau, prad, pturb, grav, mode=1):
    """TTAUP on a Rosstab.  Returns (abstd, ptotal, p_gas, n_unconverged)."""
    n = len(t)
    t, tau, prad, pturb = _f64(t), _f64(tau), _f64(prad), _f64(pturb)
    ab, pt, pg = np.zeros(n), np.zeros(n), np.zeros(n)
    L = lib()
    L.or_ttaup.argtypes = [C.c_void_p, C.c_int] + [P_D] * 7 + [C.c_double, C.c_int]
    nf = L.or_ttaup(C.addressof(tab.s), int(mode), _p(t), _p(tau), _p(ab), _p(pt), _p(pg),
                    _p(prad), _p(pturb), float(grav), n)
    return ab, pt, pg, nf


class OrCzcState(C.Structure):
    _fields_ = [("prev_engaged", C.c_int), ("jc_sticky", C.c_int), ("nbot_bad", C.c_int),
                ("run_engaged", C.c_int), ("run_jc", C.c_int), ("last_dtmax", C.c_double),
                ("rel_err", C.c_double)]


CZ_FIELDS = ("flxcnv", "flxrad", "dltdlp", "grdadb", "flxerr", "t", "rhox", "ptotal", "rho",
             "heatcp", "dlrdlt", "p", "vturb", "abross", "hscale")


class OrCzcIn(C.Structure):
    _fields_ = [(k, C.c_int) for k in ("n", "ifconv", "iter", "use4pt", "use_constructor",
                                       "irosstab")] + \
               [(k, C.c_double) for k in ("flux", "mixlth", "grav")] + \
               [(k, P_D) for k in CZ_FIELDS] + [("tab", C.c_void_p)]


def czc_state():
    st = OrCzcState()
    st.last_dtmax = 1.0e30
    return st


def cz_construct(tab, z, t1, state, it, ifconv=1, use4pt=1, mode=1):
    """CZ_CONSTRUCT(T1) on arrays z[k] for k in CZ_FIELDS and scalars
    z["flux"], z["mixlth"], z["grav"]; state (czc_state()) is its SAVE
    state, updated.  Returns (t1 out, engaged)."""
    keep = {k: _f64(z[k]) for k in CZ_FIELDS}
    s = OrCzcIn()
    s.n, s.ifconv, s.iter, s.use4pt, s.use_constructor, s.irosstab = (
        len(keep["t"]), int(ifconv), int(it), int(use4pt), 1, int(mode))
    s.flux, s.mixlth, s.grav = float(z["flux"]), float(z["mixlth"]), float(z["grav"])
    for k in CZ_FIELDS:
        setattr(s, k, _p(keep[k]))
    s.tab = C.addressof(tab.s)
    out = _f64(np.array(t1, np.float64).copy())
    L = lib()
    L.or_cz_construct.argtypes = [C.c_void_p, C.c_void_p, P_D]
    e = L.or_cz_construct(C.byref(s), C.byref(state), _p(out))
    return out, e


FIN_IN = ("flxcnv", "flxrad", "dltdlp", "grdadb", "flxerr", "ptotal", "rho", "heatcp",
          "dlrdlt", "hscale")
FIN_IO = ("tauros", "rhox", "t", "p", "xne", "abross", "prad", "vturb", "bmin", "pturb")
FIN_OUT = ("tk", "hkt", "hckt", "tkev", "tlog", "pradk")


class OrTcorrFinIn(C.Structure):
    _fields_ = [(k, C.c_int) for k in ("n", "ifconv", "iter", "use4pt", "use_constructor",
                                       "irosstab", "j1smooth", "j2smooth")] + \
               [(k, C.c_double) for k in ("flux", "mixlth", "grav", "tau1lg", "steplg", "pradk0",
                                          "tfloor", "wtjm1", "wtj", "wtjp1")] + \
               [(k, P_D) for k in FIN_IN] + [("tab", C.c_void_p)]


class OrTcorrFinIo(C.Structure):
    _fields_ = [(k, P_D) for k in FIN_IO] + [("accrad", P_D), ("bhyd", P_D)] + \
               [(k, P_D) for k in FIN_OUT]


def tcorr3_finish(tab, sc, arr, t1, state, mode=1):
    """TCORR(3) sections (G2)-(K).  sc: dict of the scalars of OrTcorrFinIn;
    arr: dict with FIN_IN, FIN_IO, "accrad" (n) and "bhyd" (6, n).  Returns
    (out dict: FIN_IO + FIN_OUT + bhyd after the remap, t1 after (G2),
    drhox, TTAUP depths not converged)."""
    n = len(arr["t"])
    s = OrTcorrFinIn()
    for k in ("n", "ifconv", "iter", "use4pt", "use_constructor", "j1smooth", "j2smooth"):
        setattr(s, k, int(sc.get(k, n if k == "n" else 0)))
    s.irosstab = int(mode)
    for k in ("flux", "mixlth", "grav", "tau1lg", "steplg", "pradk0", "tfloor", "wtjm1", "wtj",
              "wtjp1"):
        setattr(s, k, float(sc[k]))
    keep = {k: _f64(arr[k]) for k in FIN_IN}
    for k in FIN_IN:
        setattr(s, k, _p(keep[k]))
    s.tab = C.addressof(tab.s)
    io = OrTcorrFinIo()
    out = {k: _f64(np.array(arr[k], np.float64).copy()) for k in FIN_IO}
    out.update({k: np.zeros(n) for k in FIN_OUT})
    acc = _f64(arr["accrad"])
    out["bhyd"] = _f64(np.array(arr["bhyd"], np.float64).reshape(6, n).copy())
    for k in FIN_IO + FIN_OUT:
        setattr(io, k, _p(out[k]))
    io.accrad, io.bhyd = _p(acc), _p(out["bhyd"])
    t1o = _f64(np.array(t1, np.float64).copy())
    drhox = np.zeros(n)
    L = lib()
    L.or_tcorr3_finish.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, P_D, P_D]
    nf = L.or_tcorr3_finish(C.byref(s), C.byref(io), C.byref(state), _p(t1o), _p(drhox))
    return out, t1o, drhox, nf


CNV_IN = ("ptotal", "rho", "pradk", "tauros", "rhox", "abross", "vturb", "height", "flxrad")
CNV_STATE = ("t", "tk", "hkt", "hckt", "tkev", "tlog", "p")
CNV_OUT = ("dltdlp", "heatcp", "velsnd", "dlrdlt", "grdadb", "hscale", "vconv", "flxcnv",
           "flxcnv0", "flxcnv1", "abconv", "deltat")


class OrConvecIn(C.Structure):
    _fields_ = [(k, C.c_int) for k in ("n", "ifconv", "ifcorr", "nconv", "use4pt", "smooth",
                                       "irosstab")] + \
               [(k, C.c_double) for k in ("grav", "mixlth", "flux", "overwt")] + \
               [(k, P_D) for k in CNV_IN] + [("edens", P_D * 4), ("rho4", P_D * 4),
                                             ("tab", C.c_void_p)]


class OrConvecIo(C.Structure):
    _fields_ = [(k, P_D) for k in CNV_STATE + CNV_OUT]


def convec(tab, sc, arr, edens, rho4, mode=1, mlt_only=False):
    """CONVEC on the EOS evaluations edens, rho4 (4, n).  sc: n, ifconv,
    ifcorr, nconv, use4pt, smooth, grav, mixlth, flux, overwt; arr: CNV_IN,
    CNV_STATE and flxcnv1 (n).  Returns a dict of CNV_STATE (after the
    perturb-and-restore) and CNV_OUT."""
    n = len(arr["t"])
    s = OrConvecIn()
    s.n = n
    for k in ("ifconv", "ifcorr", "nconv", "use4pt", "smooth"):
        setattr(s, k, int(sc[k]))
    s.irosstab = int(mode)
    for k in ("grav", "mixlth", "flux", "overwt"):
        setattr(s, k, float(sc[k]))
    keep = {k: _f64(arr[k]) for k in CNV_IN}
    for k in CNV_IN:
        setattr(s, k, _p(keep[k]))
    e = _f64(np.asarray(edens, np.float64).reshape(4, n))
    r = _f64(np.asarray(rho4, np.float64).reshape(4, n))
    for i in range(4):
        s.edens[i] = _p(e[i])
        s.rho4[i] = _p(r[i])
    s.tab = C.addressof(tab.s)
    out = {k: _f64(np.array(arr[k], np.float64).copy()) for k in CNV_STATE}
    out.update({k: np.zeros(n) for k in CNV_OUT})
    out["flxcnv1"] = _f64(np.array(arr["flxcnv1"], np.float64).copy())
    io = OrConvecIo()
    for k in CNV_STATE + CNV_OUT:
        setattr(io, k, _p(out[k]))
    L = lib()
    L.or_convec.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
    L.or_convec(C.byref(s), C.byref(io), int(bool(mlt_only)))
    out["_keep"] = (keep, e, r)
    del out["_keep"]
    return out


class OrMoltab(C.Structure):
    _fields_ = [("nequa", C.c_int), ("nummol", C.c_int), ("locj", C.POINTER(C.c_int)),
                ("kcomps", C.POINTER(C.c_int)), ("idequa", C.POINTER(C.c_int))]


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def molec_newton(z, xntot, xab, equilj, xn0):
    """NMOLEC's gas-phase Newton solve of one depth
    (src/atlas12_modules.f90:9078-9217).  z: the molecule table (locj,
    kcomps, idequa as READMOL, 1-based).  Returns (xn, iterations,
    status 0 converged / 1 NaN / 2 cap)."""
    keep = [_i32(z["locj"]), _i32(z["kcomps"]), _i32(z["idequa"])]
    tb = OrMoltab(len(keep[2]), len(keep[0]) - 1,
                  *[k.ctypes.data_as(C.POINTER(C.c_int)) for k in keep])
    xn = _f64(np.array(xn0, np.float64).copy())
    xa, eq = _f64(xab), _f64(equilj)
    st = C.c_int(0)
    L = lib()
    L.or_molec_newton.restype = C.c_int
    it = L.or_molec_newton(C.byref(tb), C.c_double(xntot), _p(xa), _p(eq), _p(xn), C.byref(st))
    return xn, int(it), int(st.value)


def nmolec_inputs(z, case):
    """Per-depth XNTOT = P/kT and XAB (:8864-8876) of a tests/golden/
    molec_ref.npz case, and the surface start (:8851-8860)."""
    p, tk, t = z[case + "_p"], z[case + "_tk"], z[case + "_t"]
    ide = np.asarray(z["idequa"])
    xab = np.zeros((len(p), len(ide)))
    for k in range(1, len(ide)):
        if ide[k] < 100:
            xab[:, k] = np.maximum(z[case + "_xabund"][ide[k] - 1], 1.0e-20)
    xntot = p / tk
    xn1 = xntot[0] / 2.0 if t[0] >= 4000.0 else xntot[0]
    x = xn1 / 10.0
    start = x * xab[0]
    start[0] = xn1
    if ide[-1] == 100:
        start[-1] = x
    return xntot, xab, start


def nmolec_chain(z, case):
    """NMOLEC's depth loop (:8862-9217 without condensation, which is off
    above COND_TMAX = 2600 K): each depth starts from the previous one's
    solution times P(J)/P(J-1).  Returns (xn (nrhox, nequa), iterations,
    statuses)."""
    xntot, xab, xn = nmolec_inputs(z, case)
    p = z[case + "_p"]
    eqj = z[case + "_equilj"]
    out, its, sts = [], [], []
    for j in range(len(p)):
        if j > 0:
            xn = xn * (p[j] / p[j - 1])
        xn, it, st = molec_newton(z, xntot[j], xab[j], eqj[j], xn)
        out.append(xn)
        its.append(it)
        sts.append(st)
    return np.array(out), np.array(its), np.array(sts)


def nmolec_reference_xn(z, case):
    """The reference's converged solution vector per depth, read back from
    its outputs: XN(1) = XNATOM, XN(NEQUA) = XNE, and each atom equation's
    XN(K) = XNMOL of its one-component entry (EQUILJ = 1, :8937-8939)."""
    ide = np.asarray(z["idequa"])
    locj, kc = np.asarray(z["locj"]), np.asarray(z["kcomps"])
    n = len(z[case + "_p"])
    xn = np.full((n, len(ide)), np.nan)
    xn[:, 0] = z[case + "_xnatom"]
    if ide[-1] == 100:
        xn[:, -1] = z[case + "_xne"]
    for jm in range(len(locj) - 1):
        if locj[jm + 1] - locj[jm] == 1:
            k = kc[locj[jm] - 1]
            if k <= len(ide) and ide[k - 1] < 100:
                xn[:, k - 1] = z[case + "_xnmol"][jm]
    return xn
