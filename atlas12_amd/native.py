"""ctypes binding of liba12gpu.so (include/a12gpu.h).

This is the host-side mirror of the reference's hot-path subroutines:
:func:`synthe_line_opacity` stands for SYNTHE's depth loop
(src/synthe.f90:476-689), :func:`synthe_transfer` for its wavelength loop
(:698-718), :func:`josh` for JOSH (src/atlas12_modules.f90:18398-18776).
There is no CPU fallback: without the built library every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .state import SyntheState
from .workloads import LTE_DTYPE

PKG = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(PKG, "liba12gpu.so")

PREC_MIXED = 0
PREC_FP64 = 1

P_D = C.POINTER(C.c_double)

EXPORTS = ["a12_ctx_create", "a12_ctx_destroy", "a12_last_error", "a12_stream",
           "a12_set_operators", "a12_synthe_line_opacity", "a12_synthe_load",
           "a12_synthe_opacity_run", "a12_synthe_opacity_fetch", "a12_synthe_opacity_device",
           "a12_synthe_evals", "a12_synthe_transfer_run", "a12_synthe_transfer_fetch",
           "a12_synthe_transfer", "a12_josh_batch", "a12_voigt_profile",
           "a12_set_profiling", "a12_stage_ms", "a12_synthe_eval_split",
           "a12_atlas_load", "a12_atlas_linop1_run", "a12_atlas_linop1_fetch",
           "a12_atlas_set_kapp", "a12_atlas_rt_run", "a12_atlas_rt_fetch",
           "a12_atlas_select_lines", "a12_atlas_select_all", "a12_atlas_stage_ms", "a12_atlas_shape_info",
           "a12_atlas_set_shape", "a12_atlas_autotune", "a12_scale_model", "a12_mlp_init",
           "a12_mlp_get_params", "a12_mlp_set_params", "a12_mlp_set_data", "a12_mlp_grads",
           "a12_mlp_train", "a12_mlp_infer", "a12_mlp_gemm", "a12_atlas_tcorr3",
           "a12_molec_set_table", "a12_molec_newton_batched",
           "a12_synthe_shape_info", "a12_synthe_set_shape", "a12_synthe_batch_load",
           "a12_synthe_batch_set_budget", "a12_synthe_batch_run", "a12_synthe_batch_fetch",
           "a12_synthe_batch_info", "a12_solvit_batched", "a12_atlas_load_lines",
           "a12_atlas_set_state", "a12_atlas_rt_pass_noalloc", "a12_model_converge",
           "a12_kapp_set_tables", "a12_kapp_set_state", "a12_atlas_kapp_run", "a12_kapp_eval",
           "a12_kapp_constants", "a12_kapp_ms", "a12_atlas_kapp_fetch", "a12_cplx_load",
           "a12_cplx_run", "a12_cplx_fetch", "a12_cplx_clear", "a12_cplx_ms",
           "a12_rosstab_reset", "a12_rosstab_store", "a12_rosstab_load", "a12_rosstab_fetch",
           "a12_rosstab_lookup", "a12_ttaup", "a12_cz_construct", "a12_atlas_tcorr3_finish",
           "a12_convec"]


class A12Error(RuntimeError):
    pass


class LteLine(C.Structure):
    _fields_ = [("nbuff", C.c_int32), ("cgf", C.c_float), ("nelion", C.c_int32),
                ("elo", C.c_float), ("gamrf", C.c_float), ("gamsf", C.c_float),
                ("gamwf", C.c_float), ("code", C.c_float)]


class SyntheAtm(C.Structure):
    _fields_ = [("nrhox", C.c_int32), ("length", C.c_int32), ("wbegin", C.c_double),
                ("ratio", C.c_double), ("resolu", C.c_double), ("nedge", C.c_int32),
                ("wledge", P_D), ("nf", C.c_int32), ("continall", P_D), ("contabs", P_D),
                ("contscat", P_D), ("mw6", C.c_int32), ("xnfpel", P_D), ("dopple", P_D),
                ("t", P_D), ("xne", P_D), ("rho", P_D), ("xnfh", P_D), ("xnfhe1", P_D),
                ("xnfh2", P_D), ("hckt", P_D), ("hkt", P_D), ("hkt_atlas", P_D), ("rhox", P_D)]


class AtlasAtm(C.Structure):
    _fields_ = [("nrhox", C.c_int32), ("numnu", C.c_int32), ("mion", C.c_int32),
                ("ldion", C.c_int32), ("ldtab", C.c_int32), ("nwave1", C.c_int32),
                ("waveset", P_D), ("rcoset", P_D), ("tabcont", P_D), ("iwavetab", C.c_void_p),
                ("xnfdop", P_D), ("dopple", P_D), ("t", P_D), ("xne", P_D), ("hckt", P_D),
                ("hkt", P_D), ("rhox", P_D), ("xnf1", P_D), ("xnf3", P_D), ("xnf841", P_D),
                ("h0tab", P_D), ("h1tab", P_D), ("h2tab", P_D), ("teff", C.c_double)]


ACC_NAMES = ("rjmins", "rdabh", "rdiagj", "flxrad", "raden", "h_total", "accrad", "abross")

_lib = None


def load_library(path: str = LIBPATH):
    """Load liba12gpu.so; raises A12Error if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise A12Error(f"{path} is missing: build it with `python -m atlas12_amd.build`")
    L = C.CDLL(path)
    for name in EXPORTS:
        if not hasattr(L, name):
            raise A12Error(f"{path} lacks symbol {name}")
    L.a12_last_error.restype = C.c_char_p
    L.a12_stream.restype = C.c_void_p
    L.a12_synthe_opacity_device.restype = C.c_void_p
    L.a12_synthe_evals.restype = C.c_int64
    L.a12_ctx_destroy.argtypes = [C.c_void_p]
    _lib = L
    return L


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a):
    return a.ctypes.data_as(P_D)


class Context:
    """One device context (a12_ctx): its stream and device buffers."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        self.h = C.c_void_p()
        rc = self.lib.a12_ctx_create(C.c_int32(device), C.byref(self.h))
        if rc != 0:
            raise A12Error(f"a12_ctx_create({device}) failed with status {rc}")
        self._keep = []

    def close(self):
        if self.h:
            self.lib.a12_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str):
        if rc != 0:
            msg = self.lib.a12_last_error(self.h)
            raise A12Error(f"{what}: status {rc}: {msg.decode() if msg else ''}")

    @property
    def stream(self) -> int:
        return int(self.lib.a12_stream(self.h) or 0)

    # ---------------------------------------------------------------- JOSH
    def set_operators(self, coefj, coefh):
        cj, ch = _f64(coefj), _f64(coefh)
        if cj.size != 2601 or ch.size != 2601:
            raise A12Error("COEFJ/COEFH must be 51 x 51")
        self._check(self.lib.a12_set_operators(self.h, _ptr(cj), _ptr(ch)), "a12_set_operators")

    def josh_batch(self, ifscat, ifsurf, rhox, acont, aline, sigmac, sigmal, scont, sline, bnu):
        """JOSH for a batch of frequencies; inputs (nbatch, nrhox)."""
        ins = [_f64(np.atleast_2d(a)) for a in (acont, aline, sigmac, sigmal, scont, sline, bnu)]
        nb, n = ins[0].shape
        rx = _f64(rhox)
        hnu1, knu1 = np.zeros(nb), np.zeros(nb)
        outs = [np.zeros((nb, n)) for _ in range(5)]
        self._check(self.lib.a12_josh_batch(
            self.h, C.c_int32(ifscat), C.c_int32(ifsurf), C.c_int32(n), C.c_int32(nb), _ptr(rx),
            *[_ptr(a) for a in ins], _ptr(hnu1), _ptr(knu1), *[_ptr(o) for o in outs]),
            "a12_josh_batch")
        return dict(hnu1=hnu1, knu1=knu1, taunu=outs[0], snu=outs[1], jnu=outs[2],
                    hnu=outs[3], jmins=outs[4])

    def solvit_batched(self, a, b):
        """SOLVIT on the GPU for a batch of systems: a (nsys, n, n) with
        a[s, i, j] = A(i+1, j+1), b (nsys, n); returns x (nsys, n)."""
        a = np.asarray(a, np.float64)
        b = np.asarray(b, np.float64)
        nsys, n = b.shape
        if a.shape != (nsys, n, n):
            raise A12Error("solvit_batched: a must be (nsys, n, n) for b (nsys, n)")
        acm = np.ascontiguousarray(a.transpose(0, 2, 1))  # column-major per system
        bb = np.ascontiguousarray(b)
        x = np.zeros((nsys, n))
        self._check(self.lib.a12_solvit_batched(self.h, C.c_int32(n), C.c_int32(nsys), _ptr(acm),
                                                _ptr(bb), _ptr(x)), "a12_solvit_batched")
        return x

    def voigt(self, v, a, prec: int = 0):
        """H(v, a) on the GPU: prec 0 mixed FP32 path, 1 FP64, 2 the
        reference's FP32-rounded value."""
        v = np.ascontiguousarray(v, np.float32)
        a = np.ascontiguousarray(a, np.float32)
        h = np.zeros(len(v))
        self._check(self.lib.a12_voigt_profile(self.h, C.c_int32(prec), C.c_int32(len(v)),
                                               v.ctypes.data_as(C.c_void_p),
                                               a.ctypes.data_as(C.c_void_p), _ptr(h)),
                    "a12_voigt_profile")
        return h

    # -------------------------------------------------------------- SYNTHE
    def _atm(self, st: SyntheState) -> SyntheAtm:
        keep = {k: _f64(getattr(st, k)) for k in (
            "wledge", "continall", "contabs", "contscat", "xnfpel", "dopple", "t", "xne", "rho",
            "xnfh", "xnfhe1", "xnfh2", "hckt", "hkt_syn", "hkt_atl", "rhox")}
        a = SyntheAtm()
        a.nrhox, a.length = st.nrhox, st.length
        a.wbegin, a.ratio, a.resolu = st.wbegin, st.ratio, st.resolu
        a.nedge = len(st.wledge)
        a.nf = st.continall.shape[1]
        a.mw6 = st.xnfpel.shape[1]
        for k in ("wledge", "continall", "contabs", "contscat", "xnfpel", "dopple", "t", "xne",
                  "rho", "xnfh", "xnfhe1", "xnfh2", "hckt"):
            setattr(a, k, _ptr(keep[k]))
        a.hkt = _ptr(keep["hkt_syn"])
        a.hkt_atlas = _ptr(keep["hkt_atl"])
        a.rhox = _ptr(keep["rhox"])
        self._keep = [keep]
        return a

    def synthe_set_shape(self, shape: int):
        """Tile shape of the line-opacity gather (synthe_shapes())."""
        self._check(self.lib.a12_synthe_set_shape(self.h, C.c_int32(shape)),
                    "a12_synthe_set_shape")

    # ------------------------------------------------------ batched SYNTHE
    def synthe_batch_load(self, states, lines: np.ndarray, nb_shift, line_lo, line_hi):
        """Many models over one shared line list sorted by nbuff: model k
        uses lines[line_lo[k]:line_hi[k]] with nbuff - nb_shift[k]."""
        lines = np.ascontiguousarray(lines, dtype=LTE_DTYPE)
        atms = (SyntheAtm * len(states))()
        kept = []
        for k, st in enumerate(states):
            atms[k] = self._atm(st)
            kept.append(self._keep)
        sh = np.ascontiguousarray(nb_shift, np.int32)
        lo = np.ascontiguousarray(line_lo, np.int64)
        hi = np.ascontiguousarray(line_hi, np.int64)
        self._check(self.lib.a12_synthe_batch_load(
            self.h, C.c_int32(len(states)), atms, sh.ctypes.data_as(C.c_void_p),
            lo.ctypes.data_as(C.c_void_p), hi.ctypes.data_as(C.c_void_p),
            lines.ctypes.data_as(C.c_void_p), C.c_int64(len(lines))), "a12_synthe_batch_load")
        self._batch = [(st.nrhox, st.length) for st in states]

    def synthe_batch_set_budget(self, nbytes: int):
        self._check(self.lib.a12_synthe_batch_set_budget(self.h, C.c_int64(nbytes)),
                    "a12_synthe_batch_set_budget")

    def synthe_batch_run(self, prec: int = PREC_MIXED, transfer: bool = False,
                         graph: bool = False):
        self._check(self.lib.a12_synthe_batch_run(self.h, C.c_int32(prec), C.c_int32(transfer),
                                                  C.c_int32(graph)), "a12_synthe_batch_run")
        self._batch_prec = prec
        self._batch_transfer = transfer

    def synthe_batch_fetch(self, k: int, opacity: bool = True):
        """(opacity (nrhox, length) or None, hnu_line, hnu_cont (None
        without transfer), evals [all, Voigt core]) of model k."""
        n, L = self._batch[k]
        op = None
        if opacity:
            op = np.zeros((n, L), np.float64 if self._batch_prec == PREC_FP64 else np.float32)
        hl = np.zeros(L) if self._batch_transfer else None
        hc = np.zeros(L) if self._batch_transfer else None
        ev = np.zeros(2, np.int64)
        self._check(self.lib.a12_synthe_batch_fetch(
            self.h, C.c_int32(k), None if op is None else op.ctypes.data_as(C.c_void_p),
            None if hl is None else _ptr(hl), None if hc is None else _ptr(hc),
            ev.ctypes.data_as(C.c_void_p)), "a12_synthe_batch_fetch")
        return op, hl, hc, ev

    def synthe_batch_info(self):
        info = np.zeros(5, np.int64)
        self._check(self.lib.a12_synthe_batch_info(self.h, info.ctypes.data_as(C.c_void_p)),
                    "a12_synthe_batch_info")
        return dict(zip(("nmodel", "groups", "work_bytes", "budget", "largest_model_work"),
                        (int(v) for v in info)))

    def synthe_load(self, st: SyntheState, lines: np.ndarray):
        lines = np.ascontiguousarray(lines, dtype=LTE_DTYPE)
        a = self._atm(st)
        self._check(self.lib.a12_synthe_load(self.h, C.byref(a), lines.ctypes.data_as(C.c_void_p),
                                             C.c_int32(len(lines))), "a12_synthe_load")
        self._loaded = (st.nrhox, st.length)

    def synthe_opacity_run(self, prec: int = PREC_MIXED):
        self._check(self.lib.a12_synthe_opacity_run(self.h, C.c_int32(prec)),
                    "a12_synthe_opacity_run")
        self._prec = prec

    def synthe_opacity_fetch(self) -> np.ndarray:
        n, L = self._loaded
        out = np.zeros((n, L), np.float64 if self._prec == PREC_FP64 else np.float32)
        self._check(self.lib.a12_synthe_opacity_fetch(self.h, out.ctypes.data_as(C.c_void_p)),
                    "a12_synthe_opacity_fetch")
        return out

    def synthe_evals(self) -> int:
        return int(self.lib.a12_synthe_evals(self.h))

    def eval_split(self):
        """(all contributions, Voigt-core contributions) of the last run."""
        t, c = C.c_int64(0), C.c_int64(0)
        self._check(self.lib.a12_synthe_eval_split(self.h, C.byref(t), C.byref(c)),
                    "a12_synthe_eval_split")
        return t.value, c.value

    def set_profiling(self, on: bool = True):
        self._check(self.lib.a12_set_profiling(self.h, C.c_int32(1 if on else 0)),
                    "a12_set_profiling")

    def stage_ms(self):
        """[prep+continuum, line set-up, partition, tile gather, transfer] in
        ms from HIP events on the context stream (-1 = not measured)."""
        ms = (C.c_float * 5)()
        self._check(self.lib.a12_stage_ms(self.h, ms), "a12_stage_ms")
        return list(ms)

    def synthe_transfer_run(self):
        self._check(self.lib.a12_synthe_transfer_run(self.h), "a12_synthe_transfer_run")

    def synthe_transfer_fetch(self):
        L = self._loaded[1]
        hl, hc = np.zeros(L), np.zeros(L)
        self._check(self.lib.a12_synthe_transfer_fetch(self.h, _ptr(hl), _ptr(hc)),
                    "a12_synthe_transfer_fetch")
        return hl, hc

    def synchronize(self):
        """Waits for the context's stream (HIP runtime's hipStreamSynchronize)."""
        hip = C.CDLL("libamdhip64.so.7")  # the runtime the library links
        rc = hip.hipStreamSynchronize(C.c_void_p(self.stream))
        if rc != 0:
            raise A12Error(f"hipStreamSynchronize failed with HIP error {rc}")


def synthe_shapes():
    """Gather tile shapes: list of (points per thread, lines per LDS stage)."""
    lib = load_library()
    out, buf = [], (C.c_int32 * 2)()
    n = lib.a12_synthe_shape_info(C.c_int32(0), buf)
    for i in range(n):
        lib.a12_synthe_shape_info(C.c_int32(i), buf)
        out.append((buf[0], buf[1]))
    return out


def synthe_line_opacity(ctx: Context, st: SyntheState, lines: np.ndarray, prec: int = PREC_MIXED):
    """SYNTHE's LTE line-opacity build, src/synthe.f90:476-689.  Returns the
    opacity matrix (nrhox, length) and the number of accumulated
    contributions."""
    lines = np.ascontiguousarray(lines, dtype=LTE_DTYPE)
    a = ctx._atm(st)
    out = np.zeros((st.nrhox, st.length), np.float64 if prec == PREC_FP64 else np.float32)
    ev = C.c_int64(0)
    ctx._check(ctx.lib.a12_synthe_line_opacity(
        ctx.h, C.byref(a), lines.ctypes.data_as(C.c_void_p), C.c_int32(len(lines)),
        C.c_int32(prec), out.ctypes.data_as(C.c_void_p), C.byref(ev)), "a12_synthe_line_opacity")
    ctx._loaded = (st.nrhox, st.length)
    ctx._prec = prec
    return out, ev.value


def synthe_transfer(ctx: Context, st: SyntheState, opacity: np.ndarray):
    """SYNTHE's wavelength loop, src/synthe.f90:698-718: (H_line, H_cont)."""
    prec = PREC_FP64 if opacity.dtype == np.float64 else PREC_MIXED
    op = np.ascontiguousarray(opacity)
    a = ctx._atm(st)
    hl, hc = np.zeros(st.length), np.zeros(st.length)
    ctx._check(ctx.lib.a12_synthe_transfer(ctx.h, C.byref(a), op.ctypes.data_as(C.c_void_p),
                                           C.c_int32(prec), _ptr(hl), _ptr(hc)),
               "a12_synthe_transfer")
    return hl, hc


# ---------------------------------------------------------------- ATLAS12
def atlas_state_from(z, mion: int = 1006):
    """ATLAS12 state dict (keys as in tests/golden/atlas_small.npz: species
    rows of xnfdop/dopple given by ion_idx) -> arrays in the C-ABI layout."""
    n = len(z["t"])
    xnfdop = np.zeros((mion, n))
    dopple = np.zeros((mion, n))
    xnfdop[z["ion_idx"]] = z["xnfdop"]
    dopple[z["ion_idx"]] = z["dopple"]
    d = {k: _f64(z[k]) for k in ("waveset", "rcoset", "tabcont", "t", "xne", "hckt", "hkt",
                                 "rhox", "xnf1", "xnf3", "xnf841", "h0", "h1", "h2")}
    d["xnfdop"], d["dopple"] = xnfdop, dopple
    d["iwavetab"] = np.ascontiguousarray(z["iwavetab"], np.int32)
    d["teff"] = float(z["teff"])
    d["mion"] = mion
    return d


def _atlas_struct(d):
    a = AtlasAtm()
    n = len(d["t"])
    a.nrhox, a.numnu, a.mion, a.ldion, a.ldtab = n, len(d["waveset"]), d["mion"], n, n
    a.nwave1 = len(d["iwavetab"])
    for k in ("waveset", "rcoset", "tabcont", "xnfdop", "dopple", "t", "xne", "hckt", "hkt",
              "rhox", "xnf1", "xnf3", "xnf841"):
        setattr(a, k, _ptr(d[k]))
    a.iwavetab = d["iwavetab"].ctypes.data_as(C.c_void_p)
    a.h0tab, a.h1tab, a.h2tab = _ptr(d["h0"]), _ptr(d["h1"]), _ptr(d["h2"])
    a.teff = d["teff"]
    return a


def _linedata(linedata):
    ld = np.ascontiguousarray(linedata, np.int32)
    if ld.ndim != 2 or ld.shape[1] != 4:
        raise ValueError("linedata must have shape (nlines, 4)")
    return ld


def atlas_load(ctx: Context, d, linedata):
    """Line store and state in one call (a12_atlas_load)."""
    ld = _linedata(linedata)
    a = _atlas_struct(d)
    ctx._check(ctx.lib.a12_atlas_load(ctx.h, C.byref(a), ld.ctypes.data_as(C.c_void_p),
                                      C.c_int32(len(ld))), "a12_atlas_load")
    ctx._atl = (len(d["t"]), len(d["waveset"]))


def atlas_load_lines(ctx: Context, d, linedata):
    """The line store alone, once per store (a12_atlas_load_lines); d
    supplies the grid (waveset, iwavetab, mion)."""
    ld = _linedata(linedata)
    a = _atlas_struct(d)
    ctx._check(ctx.lib.a12_atlas_load_lines(ctx.h, C.byref(a), ld.ctypes.data_as(C.c_void_p),
                                            C.c_int32(len(ld))), "a12_atlas_load_lines")
    ctx._atl = None


def atlas_set_state(ctx: Context, d):
    """The per-iteration state on the loaded store's grid (a12_atlas_set_state)."""
    a = _atlas_struct(d)
    ctx._check(ctx.lib.a12_atlas_set_state(ctx.h, C.byref(a)), "a12_atlas_set_state")
    ctx._atl = (len(d["t"]), len(d["waveset"]))


def _atl_dims(ctx: Context):
    dims = getattr(ctx, "_atl", None)
    if dims is None:
        raise A12Error("no ATLAS12 state set on this context (atlas_set_state / atlas_load)")
    return dims


def atlas_linop1(ctx: Context, prec: int = PREC_FP64, fetch: bool = True):
    """LINOP1 (src/atlas12_modules.f90:18132-18374): XLINES (numnu, nrhox)
    and the contribution count."""
    ctx._check(ctx.lib.a12_atlas_linop1_run(ctx.h, C.c_int32(prec)), "a12_atlas_linop1_run")
    if not fetch:
        return None, None
    n, nf = _atl_dims(ctx)
    out = np.zeros((nf, n))
    ev = C.c_int64(0)
    ctx._check(ctx.lib.a12_atlas_linop1_fetch(ctx.h, _ptr(out), C.byref(ev)),
               "a12_atlas_linop1_fetch")
    return out, ev.value


def atlas_set_kapp(ctx: Context, acont, sigmac, scont, sline, aline, sigmal):
    """KAPP's six tables, each (numnu, nrhox) as the device holds them
    (nrhox fastest); any other size is rejected before the call."""
    n, nf = _atl_dims(ctx)
    ts = [_f64(t) for t in (acont, sigmac, scont, sline, aline, sigmal)]
    for name, t in zip(("acont", "sigmac", "scont", "sline", "aline", "sigmal"), ts):
        if t.size != n * nf:
            raise ValueError(f"{name} has {t.size} elements, the state needs {nf} x {n}")
    ctx._check(ctx.lib.a12_atlas_set_kapp(ctx.h, *[_ptr(t) for t in ts]), "a12_atlas_set_kapp")


def atlas_rt_pass(ctx: Context, xlines=None, fetch: bool = True):
    """RUN_RT_PASS's frequency loop (src/atlas12_modules.f90:3905-3949):
    accumulator dict and HNU(1) per frequency."""
    n, nf = _atl_dims(ctx)
    xl = None if xlines is None else _f64(xlines)
    if xl is not None and xl.size != n * nf:
        raise ValueError(f"xlines has {xl.size} elements, the state needs {nf} x {n}")
    ctx._check(ctx.lib.a12_atlas_rt_run(ctx.h, None if xl is None else _ptr(xl)),
               "a12_atlas_rt_run")
    if not fetch:
        return None, None
    acc = np.zeros(9 * n)
    h1 = np.zeros(nf)
    ctx._check(ctx.lib.a12_atlas_rt_fetch(ctx.h, _ptr(acc), _ptr(h1)), "a12_atlas_rt_fetch")
    out = {k: acc[i * n:(i + 1) * n].copy() for i, k in enumerate(ACC_NAMES)}
    out["pradk0"] = float(acc[8 * n])
    return out, h1


TCORR_STATE = ("tauros", "rhox", "t", "dltdlp", "abross", "flxcnv", "flxcnv0", "grdadb",
               "ptotal", "rho", "dlrdlt", "heatcp", "hscale")
TCORR_SUMS = ("flxrad", "rjmins", "rdiagj")
TCORR_OUT = ("t1", "dtflux", "dtlamb", "dtsurf", "flxerr", "flxdrv", "hratio")


class TcorrIn(C.Structure):
    _fields_ = [("nrhox", C.c_int32), ("ifconv", C.c_int32), ("iter", C.c_int32),
                ("flux", C.c_double), ("teff", C.c_double), ("mixlth", C.c_double)] + \
               [(k, P_D) for k in TCORR_STATE + TCORR_SUMS]


def atlas_tcorr3(ctx: Context, z, rdabh=None, oldt1=None, iter_=None, device_sums=False):
    """TCORR(3) sections A-F (src/atlas12_modules.f90:2528-2752) on the
    device.  z: mapping with n, ifconv, iter, flux, teff, mixlth and the
    TCORR_STATE arrays; with device_sums False also TCORR_SUMS and rdabh
    (RDABH before TCORR), with device_sums True those four come from the
    last atlas_rt_pass on this context.  Returns (out dict of TCORR_OUT,
    updated rdabh, updated oldt1)."""
    n = int(z["n"])
    keep = {k: _f64(z[k]) for k in TCORR_STATE + (() if device_sums else TCORR_SUMS)}
    for k, a in keep.items():
        if a.shape != (n,):
            raise ValueError(f"{k} must have shape ({n},)")
    s = TcorrIn()
    s.nrhox, s.ifconv = n, int(z["ifconv"])
    s.iter = int(z["iter"]) if iter_ is None else int(iter_)
    s.flux, s.teff, s.mixlth = float(z["flux"]), float(z["teff"]), float(z["mixlth"])
    for k, a in keep.items():
        setattr(s, k, _ptr(a))
    rd = np.zeros(n) if rdabh is None else _f64(np.array(rdabh, np.float64).copy())
    if rd.shape != (n,):
        raise ValueError(f"rdabh must have shape ({n},)")
    ot = np.zeros(n) if oldt1 is None else _f64(np.array(oldt1, np.float64).copy())
    out = np.zeros((7, n))
    ctx._check(ctx.lib.a12_atlas_tcorr3(ctx.h, C.byref(s), _ptr(rd), _ptr(ot), _ptr(out)),
               "a12_atlas_tcorr3")
    return {k: out[i] for i, k in enumerate(TCORR_OUT)}, rd, ot


ROSSTAB_MAXTAB = 4800  # kw * 60, src/atlas12_modules.f90:1779


def rosstab_reset(ctx: Context):
    """Empty the context's ROSSTAB table (a new model)."""
    ctx._check(ctx.lib.a12_rosstab_reset(ctx.h), "a12_rosstab_reset")


def rosstab_store(ctx: Context, t, p, abross):
    """ROSSTAB(0, 0, 0): append the layers' (T, P, ABROSS) on the device."""
    t, p, ab = _f64(t), _f64(p), _f64(abross)
    n = len(t)
    if p.shape != (n,) or ab.shape != (n,):
        raise ValueError("t, p and abross must have the same length")
    ctx._check(ctx.lib.a12_rosstab_store(ctx.h, n, _ptr(t), _ptr(p), _ptr(ab)),
               "a12_rosstab_store")


def rosstab_load(ctx: Context, zero, tab):
    """Replace the table: zero (ZEROT, SLOPET, ZEROP, SLOPEP), tab (3, nross)."""
    z = _f64(np.asarray(zero, np.float64).reshape(4))
    t = _f64(np.asarray(tab, np.float64))
    if t.ndim != 2 or t.shape[0] != 3:
        raise ValueError("tab must have shape (3, nross)")
    ctx._check(ctx.lib.a12_rosstab_load(ctx.h, t.shape[1], _ptr(z), _ptr(t)), "a12_rosstab_load")


def rosstab_fetch(ctx: Context):
    """(zero (4), tab (3, nross)) of the context's table."""
    nr = C.c_int32(0)
    ctx._check(ctx.lib.a12_rosstab_fetch(ctx.h, C.byref(nr), None, None), "a12_rosstab_fetch")
    z = np.zeros(4)
    t = np.zeros((3, nr.value))
    if nr.value:
        ctx._check(ctx.lib.a12_rosstab_fetch(ctx.h, C.byref(nr), _ptr(z), _ptr(t)),
                   "a12_rosstab_fetch")
    return z, t


def rosstab_lookup(ctx: Context, temp, pressure, mode: int = 1):
    """ROSSTAB(T, P, V) for each query (mode = IROSSTAB)."""
    t, p = _f64(np.atleast_1d(temp)), _f64(np.atleast_1d(pressure))
    if p.shape != t.shape:
        raise ValueError("temp and pressure must have the same length")
    out = np.zeros(len(t))
    ctx._check(ctx.lib.a12_rosstab_lookup(ctx.h, int(mode), len(t), _ptr(t), _ptr(p), _ptr(out)),
               "a12_rosstab_lookup")
    return out


def ttaup(ctx: Context, t, tau, prad, pturb, grav: float, mode: int = 1):
    """TTAUP on the device.  Returns (abstd, ptotal, p_gas, n_unconverged)."""
    t, tau, prad, pturb = (_f64(a) for a in (t, tau, prad, pturb))
    n = len(t)
    for a in (tau, prad, pturb):
        if a.shape != (n,):
            raise ValueError(f"every array must have shape ({n},)")
    ab, pt, pg = np.zeros(n), np.zeros(n), np.zeros(n)
    nf = C.c_int32(0)
    fn = ctx.lib.a12_ttaup
    fn.argtypes = [C.c_void_p, C.c_int32, C.c_int32] + [P_D] * 4 + [C.c_double] + [P_D] * 3 + \
        [C.POINTER(C.c_int32)]
    ctx._check(fn(ctx.h, int(mode), n, _ptr(t), _ptr(tau), _ptr(prad), _ptr(pturb), float(grav),
                  _ptr(ab), _ptr(pt), _ptr(pg), C.byref(nf)), "a12_ttaup")
    return ab, pt, pg, nf.value


class CzcState(C.Structure):
    _fields_ = [("prev_engaged", C.c_int32), ("jc_sticky", C.c_int32), ("nbot_bad", C.c_int32),
                ("run_engaged", C.c_int32), ("run_jc", C.c_int32), ("last_dtmax", C.c_double),
                ("rel_err", C.c_double)]


def czc_state() -> CzcState:
    """CZ_CONSTRUCT's SAVE state at the start of a run."""
    st = CzcState()
    st.last_dtmax = 1.0e30
    return st


CZC_ARRAYS = ("flxcnv", "flxrad", "dltdlp", "grdadb", "flxerr", "t", "rhox", "ptotal", "rho",
              "heatcp", "dlrdlt", "p", "vturb", "abross", "hscale")


class CzcIn(C.Structure):
    _fields_ = [(k, C.c_int32) for k in ("nrhox", "ifconv", "iter", "use_constructor", "use4pt",
                                         "irosstab")] + \
               [(k, C.c_double) for k in ("flux", "mixlth", "grav")] + \
               [(k, P_D) for k in CZC_ARRAYS]


def cz_construct(ctx: Context, z, t1, state: CzcState, it: int, ifconv: int = 1,
                 use4pt: int = 1, mode: int = 1, use_constructor: int = 1):
    """CZ_CONSTRUCT(T1) on the device: z maps CZC_ARRAYS and flux, mixlth,
    grav; state (czc_state()) is updated.  Returns (t1 out, engaged)."""
    keep = {k: _f64(z[k]) for k in CZC_ARRAYS}
    n = len(keep["t"])
    for k, a in keep.items():
        if a.shape != (n,):
            raise ValueError(f"{k} must have shape ({n},)")
    s = CzcIn()
    s.nrhox, s.ifconv, s.iter, s.use_constructor, s.use4pt, s.irosstab = (
        n, int(ifconv), int(it), int(use_constructor), int(use4pt), int(mode))
    s.flux, s.mixlth, s.grav = float(z["flux"]), float(z["mixlth"]), float(z["grav"])
    for k, a in keep.items():
        setattr(s, k, _ptr(a))
    out = _f64(np.array(t1, np.float64).copy())
    if out.shape != (n,):
        raise ValueError(f"t1 must have shape ({n},)")
    e = C.c_int32(0)
    ctx._check(ctx.lib.a12_cz_construct(ctx.h, C.byref(s), C.byref(state), _ptr(out), C.byref(e)),
               "a12_cz_construct")
    return out, e.value


FIN_IN = ("flxcnv", "flxrad", "dltdlp", "grdadb", "flxerr", "ptotal", "rho", "heatcp",
          "dlrdlt", "hscale")
FIN_IO = ("tauros", "rhox", "t", "p", "xne", "abross", "prad", "vturb", "bmin", "pturb")
FIN_OUT = ("tk", "hkt", "hckt", "tkev", "tlog", "pradk")


class TcorrFinIn(C.Structure):
    _fields_ = [(k, C.c_int32) for k in ("nrhox", "ifconv", "iter", "use_constructor", "use4pt",
                                         "irosstab", "j1smooth", "j2smooth")] + \
               [(k, C.c_double) for k in ("flux", "mixlth", "grav", "tau1lg", "steplg", "pradk0",
                                          "tfloor", "wtjm1", "wtj", "wtjp1")] + \
               [(k, P_D) for k in FIN_IN]


class TcorrFinIo(C.Structure):
    _fields_ = [(k, P_D) for k in FIN_IO] + [("accrad", P_D), ("bhyd", P_D)] + \
               [(k, P_D) for k in FIN_OUT]


def atlas_tcorr3_finish(ctx: Context, sc, arr, t1, state: CzcState, mode: int = 1):
    """TCORR(3) sections (G2)-(K) on the device.  sc: the scalars of
    TcorrFinIn (use_constructor and use4pt default to the reference's 1);
    arr: FIN_IN, FIN_IO, accrad (n) and bhyd (6, n).  Returns (out dict of
    FIN_IO + FIN_OUT + bhyd, t1 after (G2), drhox, TTAUP depths not
    converged)."""
    n = len(arr["t"])
    s = TcorrFinIn()
    s.nrhox = n
    for k in ("ifconv", "iter", "j1smooth", "j2smooth"):
        setattr(s, k, int(sc.get(k, 0)))
    s.use_constructor = int(sc.get("use_constructor", 1))
    s.use4pt = int(sc.get("use4pt", 1))
    s.irosstab = int(mode)
    for k in ("flux", "mixlth", "grav", "tau1lg", "steplg", "pradk0", "tfloor", "wtjm1", "wtj",
              "wtjp1"):
        setattr(s, k, float(sc[k]))
    keep = {k: _f64(arr[k]) for k in FIN_IN}
    out = {k: _f64(np.array(arr[k], np.float64).copy()) for k in FIN_IO}
    out.update({k: np.zeros(n) for k in FIN_OUT})
    out["bhyd"] = _f64(np.array(arr["bhyd"], np.float64).copy())
    acc = _f64(arr["accrad"])
    for k, a in list(keep.items()) + [(k, out[k]) for k in FIN_IO] + [("accrad", acc)]:
        if a.shape != (n,):
            raise ValueError(f"{k} must have shape ({n},)")
    if out["bhyd"].shape != (6, n):
        raise ValueError(f"bhyd must have shape (6, {n})")
    for k in FIN_IN:
        setattr(s, k, _ptr(keep[k]))
    io = TcorrFinIo()
    for k in FIN_IO + FIN_OUT:
        setattr(io, k, _ptr(out[k]))
    io.accrad, io.bhyd = _ptr(acc), _ptr(out["bhyd"])
    t1o = _f64(np.array(t1, np.float64).copy())
    if t1o.shape != (n,):
        raise ValueError(f"t1 must have shape ({n},)")
    drhox = np.zeros(n)
    nf = C.c_int32(0)
    ctx._check(ctx.lib.a12_atlas_tcorr3_finish(ctx.h, C.byref(s), C.byref(io), C.byref(state),
                                               _ptr(t1o), _ptr(drhox), C.byref(nf)),
               "a12_atlas_tcorr3_finish")
    return out, t1o, drhox, nf.value


CNV_IN = ("ptotal", "rho", "pradk", "tauros", "rhox", "abross", "vturb", "height", "flxrad")
CNV_STATE = ("t", "tk", "hkt", "hckt", "tkev", "tlog", "p")
CNV_OUT = ("dltdlp", "heatcp", "velsnd", "dlrdlt", "grdadb", "hscale", "vconv", "flxcnv",
           "flxcnv0", "flxcnv1", "abconv", "deltat")


class ConvecIn(C.Structure):
    _fields_ = [(k, C.c_int32) for k in ("nrhox", "ifconv", "ifcorr", "nconv", "use4pt", "smooth",
                                         "irosstab")] + \
               [(k, C.c_double) for k in ("grav", "mixlth", "flux", "overwt")] + \
               [(k, P_D) for k in CNV_IN + ("edens", "rho4")]


class ConvecIo(C.Structure):
    _fields_ = [(k, P_D) for k in CNV_STATE + CNV_OUT]


def convec(ctx: Context, sc, arr, edens, rho4, mode: int = 1, mlt_only: bool = False):
    """CONVEC on the device.  sc: ifconv, ifcorr, nconv, use4pt, smooth,
    grav, mixlth, flux, overwt; arr: CNV_IN, CNV_STATE and flxcnv1 (n);
    edens, rho4 (4, n): the EOS's EDENS and RHO at T x 1.001, T x 0.999,
    P x 1.001, P x 0.999.  Returns a dict of CNV_STATE (after CONVEC's
    perturb-and-restore) and CNV_OUT."""
    n = len(arr["t"])
    s = ConvecIn()
    s.nrhox = n
    for k in ("ifconv", "ifcorr", "nconv", "use4pt", "smooth"):
        setattr(s, k, int(sc[k]))
    s.irosstab = int(mode)
    for k in ("grav", "mixlth", "flux", "overwt"):
        setattr(s, k, float(sc[k]))
    keep = {k: _f64(arr[k]) for k in CNV_IN}
    e = _f64(np.asarray(edens, np.float64))
    r = _f64(np.asarray(rho4, np.float64))
    for k, a in keep.items():
        if a.shape != (n,):
            raise ValueError(f"{k} must have shape ({n},)")
    if e.shape != (4, n) or r.shape != (4, n):
        raise ValueError(f"edens and rho4 must have shape (4, {n})")
    for k in CNV_IN:
        setattr(s, k, _ptr(keep[k]))
    s.edens, s.rho4 = _ptr(e), _ptr(r)
    out = {k: _f64(np.array(arr[k], np.float64).copy()) for k in CNV_STATE}
    out.update({k: np.zeros(n) for k in CNV_OUT})
    out["flxcnv1"] = _f64(np.array(arr["flxcnv1"], np.float64).copy())
    for k in CNV_STATE + ("flxcnv1",):
        if out[k].shape != (n,):
            raise ValueError(f"{k} must have shape ({n},)")
    io = ConvecIo()
    for k in CNV_STATE + CNV_OUT:
        setattr(io, k, _ptr(out[k]))
    ctx._check(ctx.lib.a12_convec(ctx.h, C.byref(s), C.byref(io), int(bool(mlt_only))),
               "a12_convec")
    return out


class MolTab(C.Structure):
    _fields_ = [("nequa", C.c_int32), ("nummol", C.c_int32), ("locj", C.POINTER(C.c_int32)),
                ("kcomps", C.POINTER(C.c_int32)), ("idequa", C.POINTER(C.c_int32))]


def molec_set_table(ctx: Context, locj, kcomps, idequa):
    """READMOL's molecule table for NMOLEC's Newton solve (1-based arrays
    as the reference holds them)."""
    a = [np.ascontiguousarray(x, np.int32) for x in (locj, kcomps, idequa)]
    t = MolTab(len(a[2]), len(a[0]) - 1, *[x.ctypes.data_as(C.POINTER(C.c_int32)) for x in a])
    ctx._check(ctx.lib.a12_molec_set_table(ctx.h, C.byref(t)), "a12_molec_set_table")
    ctx._moltab = (len(a[2]), len(a[0]) - 1)


def molec_newton(ctx: Context, xntot, xab, equilj, xn0, chain_off=None, pratio=None,
                 xnmol: bool = True):
    """NMOLEC's gas-phase Newton solve (src/atlas12_modules.f90:9078-9217)
    for a batch of systems on the device.  xntot (nsys), xab (nsys, nequa),
    equilj (nsys, nummol), xn0 (nsys, nequa): the start of each chain head.
    chain_off (nchain + 1) groups consecutive systems into NMOLEC depth
    loops that warm-start with pratio[s] = P(J)/P(J-1); None: every system
    on its own.  Returns (xn, xnmol or None, iterations, status)."""
    ne, nm = ctx._moltab
    xt = _f64(xntot)
    n = len(xt)
    xa, eq = _f64(xab).reshape(n, ne), _f64(equilj).reshape(n, nm)
    xn = _f64(np.array(xn0, np.float64).reshape(n, ne).copy())
    co = np.ascontiguousarray(np.arange(n + 1) if chain_off is None else chain_off, np.int32)
    pr = None if pratio is None else _f64(pratio)
    xm = np.zeros((n, nm)) if xnmol else None
    it = np.zeros(n, np.int32)
    st = np.zeros(n, np.int32)
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))  # noqa: E731
    ctx._check(ctx.lib.a12_molec_newton_batched(
        ctx.h, C.c_int32(n), C.c_int32(len(co) - 1), ip(co), _ptr(xt),
        None if pr is None else _ptr(pr), _ptr(xa), _ptr(eq), _ptr(xn),
        None if xm is None else _ptr(xm), ip(it), ip(st)), "a12_molec_newton_batched")
    return xn, xm, it, st


def atlas_stage_ms(ctx: Context):
    """LINOP1 set-up, reach, gather and RT-pass times [ms] of the last runs
    (with profiling on); -1 where not measured."""
    ms = (C.c_float * 4)()
    ctx._check(ctx.lib.a12_atlas_stage_ms(ctx.h, ms), "a12_atlas_stage_ms")
    return np.array(ms[:], np.float64)


def atlas_select_lines(ctx: Context, z, pred, mol2, mion: int = 1006):
    """SELECTLINES' filter (src/atlas12_modules.f90:19171-19781) on the GPU:
    predicted records (n, 4) int32 then MOL2 TiO records.  z holds tabcont,
    iwavetab, wavetab, hckt, teff and the xnfdop rows ion_idx.  Returns the
    kept LINEDATA (m, 4) int32."""
    n = len(z["hckt"])
    xnfdop = np.zeros((mion, n))
    xnfdop[z["ion_idx"]] = z["xnfdop"]
    keep = {"tabcont": _f64(z["tabcont"]), "xnfdop": xnfdop, "hckt": _f64(z["hckt"]),
            "wavetab": _f64(z["wavetab"])}
    iw = np.ascontiguousarray(z["iwavetab"], np.int32)
    a = AtlasAtm()
    a.nrhox, a.numnu, a.mion, a.ldion, a.ldtab, a.nwave1 = n, 0, mion, n, n, len(iw)
    a.tabcont, a.xnfdop, a.hckt = _ptr(keep["tabcont"]), _ptr(keep["xnfdop"]), _ptr(keep["hckt"])
    a.iwavetab = iw.ctypes.data_as(C.c_void_p)
    a.teff = float(z["teff"])
    pr = np.ascontiguousarray(pred, np.int32)
    mo = np.ascontiguousarray(mol2)
    if mo.dtype.itemsize != 32:
        raise ValueError("MOL2 records must be 32 bytes")
    out = np.zeros((len(pr) + len(mo), 4), np.int32)
    nk = C.c_int64(0)
    ctx._check(ctx.lib.a12_atlas_select_lines(
        ctx.h, C.byref(a), _ptr(keep["wavetab"]), pr.ctypes.data_as(C.c_void_p),
        C.c_int64(len(pr)), mo.ctypes.data_as(C.c_void_p), C.c_int64(len(mo)),
        out.ctypes.data_as(C.c_void_p), C.c_int64(len(out)), C.byref(nk)),
        "a12_atlas_select_lines")
    return out[:nk.value]


class SelLists(C.Structure):
    """a12_sel_lists"""
    _fields_ = [("pred", C.c_void_p), ("npred", C.c_int64), ("lowobs", C.c_void_p),
                ("nlowobs", C.c_int64), ("hilines", C.c_void_p), ("nhilines", C.c_int64),
                ("dia", C.c_void_p), ("ndia", C.c_int64), ("mol2", C.c_void_p),
                ("nmol2", C.c_int64), ("h2o", C.c_void_p), ("nh2o", C.c_int64),
                ("h3p", C.c_void_p), ("nh3p", C.c_int64)]


SEL_LISTS = {"pred": 16, "lowobs": 16, "hilines": 16, "dia": 48, "mol2": 32, "h2o": 8, "h3p": 16}


def atlas_select_all(ctx: Context, z, lists: dict, mion: int = 1006):
    """SELECTLINES over all seven lists (src/atlas12_modules.f90:19171-19781)
    on the GPU.  lists maps SEL_LISTS names to records: the 16-byte lists as
    (n, 4) int32, dia as records.DIATOMIC_DTYPE (reader order), mol2 as
    workloads.MOL2_DTYPE, h2o as workloads.H2O_DTYPE; absent ones are empty.
    z holds tabcont, iwavetab, wavetab, hckt, teff and xnfdop, either
    (mion, n) or the rows ion_idx.  Returns the kept LINEDATA (m, 4) int32."""
    n = len(z["hckt"])
    if "ion_idx" in z:
        xnfdop = np.zeros((mion, n))
        xnfdop[z["ion_idx"]] = z["xnfdop"]
    else:
        xnfdop = _f64(z["xnfdop"])
        mion = xnfdop.shape[0]
    keep = {"tabcont": _f64(z["tabcont"]), "xnfdop": xnfdop, "hckt": _f64(z["hckt"]),
            "wavetab": _f64(z["wavetab"])}
    iw = np.ascontiguousarray(z["iwavetab"], np.int32)
    a = AtlasAtm()
    a.nrhox, a.numnu, a.mion, a.ldion, a.ldtab, a.nwave1 = n, 0, mion, n, n, len(iw)
    a.tabcont, a.xnfdop, a.hckt = _ptr(keep["tabcont"]), _ptr(keep["xnfdop"]), _ptr(keep["hckt"])
    a.iwavetab = iw.ctypes.data_as(C.c_void_p)
    a.teff = float(z["teff"])
    sl = SelLists()
    held, total = [], 0
    for name, size in SEL_LISTS.items():
        r = lists.get(name)
        if r is None or len(r) == 0:
            continue
        r = np.ascontiguousarray(r)
        if r.nbytes != size * len(r):
            raise ValueError(f"{name} records must be {size} bytes")
        held.append(r)
        setattr(sl, name, r.ctypes.data_as(C.c_void_p))
        setattr(sl, "n" + name, len(r))
        total += len(r)
    out = np.zeros((max(total, 1), 4), np.int32)
    nk = C.c_int64(0)
    ctx._check(ctx.lib.a12_atlas_select_all(
        ctx.h, C.byref(a), _ptr(keep["wavetab"]), C.byref(sl), out.ctypes.data_as(C.c_void_p),
        C.c_int64(len(out)), C.byref(nk)), "a12_atlas_select_all")
    return out[:nk.value]


def atlas_shapes(ctx: Context):
    """LINOP1 gather tile shapes: list of (points per tile, lanes per
    block, lines per LDS batch, flat)."""
    out = (C.c_int32 * 4)()
    n = ctx.lib.a12_atlas_shape_info(C.c_int32(-1), None)
    shapes = []
    for k in range(n):
        ctx.lib.a12_atlas_shape_info(C.c_int32(k), out)
        shapes.append(tuple(out[:]))
    return shapes


def atlas_set_shape(ctx: Context, shape: int):
    ctx._check(ctx.lib.a12_atlas_set_shape(ctx.h, C.c_int32(shape)), "a12_atlas_set_shape")


def atlas_autotune(ctx: Context, prec: int = PREC_MIXED):
    """Times the LINOP1 gather in every tile shape on the loaded store and
    keeps the fastest.  Returns (best shape index, ms per shape)."""
    n = len(atlas_shapes(ctx))
    ms = (C.c_float * n)()
    best = C.c_int32(-1)
    ctx._check(ctx.lib.a12_atlas_autotune(ctx.h, C.c_int32(prec), ms, C.byref(best)),
               "a12_atlas_autotune")
    return best.value, np.array(ms[:], np.float64)


def scale_model(ctx: Context, sc, fields, kw: int = 80, tau1lg: float = -6.875,
                steplg: float = 0.125):
    """SCALE_MODEL (src/atlas12_modules.f90:6967-7060) for a batch of models.
    sc (nmodel, 7+): TEFF, GRAV, PRADK0, PCON, PTURB0, TEFF_NEW, LOGG_NEW;
    fields (nmodel, 7, nrhox): RHOX, T, P, XNE, ABROSS, PRAD, VTURB.
    Returns (out (nmodel, 9, kw): the 7 fields, PRADK, PTOTAL; so (nmodel,
    6): TEFF, GRAV, GLOG, FLUX, PRADK0, PZERO; TAUSTD (kw))."""
    f = np.ascontiguousarray(fields, np.float64)
    nm, nf, n = f.shape
    if nf != 7:
        raise ValueError("fields must be (nmodel, 7, nrhox)")
    s8 = np.zeros((nm, 8))
    s8[:, :7] = np.asarray(sc, np.float64)[:, :7]
    out = np.zeros((nm, 9, kw))
    so = np.zeros((nm, 6))
    tau = np.zeros(kw)
    ctx._check(ctx.lib.a12_scale_model(ctx.h, C.c_int32(nm), C.c_int32(n), C.c_int32(kw),
                                       C.c_double(tau1lg), C.c_double(steplg), _ptr(s8), _ptr(f),
                                       _ptr(out), _ptr(so), _ptr(tau)), "a12_scale_model")
    return out, so, tau


# ------------------------------------------------- starting-structure network
MLP_PARAMS = 235_664


def mlp_init(ctx: Context, batch: int, seed: int = 0):
    ctx._check(ctx.lib.a12_mlp_init(ctx.h, C.c_int32(batch), C.c_uint64(seed)), "a12_mlp_init")


def mlp_get_params(ctx: Context) -> np.ndarray:
    p = np.zeros(MLP_PARAMS, np.float32)
    ctx._check(ctx.lib.a12_mlp_get_params(ctx.h, p.ctypes.data_as(C.c_void_p)),
               "a12_mlp_get_params")
    return p


def mlp_set_params(ctx: Context, params):
    p = np.ascontiguousarray(params, np.float32)
    if p.size != MLP_PARAMS:
        raise ValueError(f"expected {MLP_PARAMS} parameters")
    ctx._check(ctx.lib.a12_mlp_set_params(ctx.h, p.ctypes.data_as(C.c_void_p)),
               "a12_mlp_set_params")


def mlp_set_data(ctx: Context, x, t):
    x = np.ascontiguousarray(x, np.float32)
    t = np.ascontiguousarray(t, np.float32)
    if x.shape[1:] != (4,) or t.shape != (len(x), 144):
        raise ValueError("x (n, 4), t (n, 144)")
    ctx._keep_mlp = (x, t)
    ctx._check(ctx.lib.a12_mlp_set_data(ctx.h, x.ctypes.data_as(C.c_void_p),
                                        t.ctypes.data_as(C.c_void_p), C.c_int32(len(x))),
               "a12_mlp_set_data")


def mlp_grads(ctx: Context):
    g = np.zeros(MLP_PARAMS, np.float32)
    loss = C.c_float(0.0)
    ctx._check(ctx.lib.a12_mlp_grads(ctx.h, C.byref(loss), g.ctypes.data_as(C.c_void_p)),
               "a12_mlp_grads")
    return loss.value, g


def mlp_train(ctx: Context, steps: int, lr: float) -> float:
    loss = C.c_float(0.0)
    ctx._check(ctx.lib.a12_mlp_train(ctx.h, C.c_int32(steps), C.c_float(lr), C.byref(loss)),
               "a12_mlp_train")
    return loss.value


def mlp_infer(ctx: Context, x) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    y = np.zeros((len(x), 144), np.float32)
    ctx._check(ctx.lib.a12_mlp_infer(ctx.h, x.ctypes.data_as(C.c_void_p), C.c_int32(len(x)),
                                     y.ctypes.data_as(C.c_void_p)), "a12_mlp_infer")
    return y


def mlp_gemm(ctx: Context, a_bits, b_bits, split: int = 1) -> np.ndarray:
    """C = A B on the MFMA tile kernel (k split `split` ways when > 1);
    A (m x k), B (k x n) as BF16 bit patterns (uint16), C FP32."""
    a = np.ascontiguousarray(a_bits, np.uint16)
    b = np.ascontiguousarray(b_bits, np.uint16)
    m, k = a.shape
    k2, n = b.shape
    assert k == k2
    c = np.zeros((m, n), np.float32)
    ctx._check(ctx.lib.a12_mlp_gemm(ctx.h, C.c_int32(m), C.c_int32(n), C.c_int32(k),
                                    a.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p),
                                    c.ctypes.data_as(C.c_void_p), C.c_int32(split)),
               "a12_mlp_gemm")
    return c


def atlas_rt_pass_noalloc(ctx: Context, n: int, nf: int, teff: float, d_waveset: int,
                          d_rcoset: int, d_depth3: int, d_kapp: int, d_xlines: int,
                          d_hnu1: int, d_acc: int):
    """a12_atlas_rt_pass_noalloc on device addresses (ints): the RT pass
    with no allocation, launched on the context's stream (not synchronised).
    Sizes are checked by the library against its fixed scratch array."""
    vp = C.c_void_p
    ctx._check(ctx.lib.a12_atlas_rt_pass_noalloc(
        ctx.h, C.c_int32(n), C.c_int32(nf), C.c_double(teff), vp(d_waveset), vp(d_rcoset),
        vp(d_depth3), vp(d_kapp), vp(d_xlines), vp(d_hnu1), vp(d_acc)),
        "a12_atlas_rt_pass_noalloc")


class ModelPar(C.Structure):
    _fields_ = [("teff", C.c_double), ("logg", C.c_double), ("mh", C.c_double),
                ("vturb", C.c_double)]


MODEL_NRHOX = 72
MODEL_NFREQ = 1200
MODEL_TAUSTD = 10.0 ** (-6.875 + 0.125 * np.arange(MODEL_NRHOX))


def model_converge(ctx: Context, par, t_start, itmax: int = 60, history: bool = False):
    """a12_model_converge: the ATLAS12 iteration (reduced-physics KAPP and
    EOS, device RT pass and TCORR(3)) for many models.  par (nmodel, 4):
    Teff, log g, [M/H], vturb [km/s]; t_start (nmodel, 72) on MODEL_TAUSTD.
    Returns (iters (nmodel,), T (nmodel, 72), structure dict of (nmodel, 72)
    arrays pg, ptot, rhox, abross, xne[, hist (nmodel, itmax, 3)])."""
    pa = np.ascontiguousarray(par, np.float64)
    if pa.ndim != 2 or pa.shape[1] != 4:
        raise ValueError("par must have shape (nmodel, 4)")
    nm = len(pa)
    ts = np.ascontiguousarray(t_start, np.float64)
    if ts.shape != (nm, MODEL_NRHOX):
        raise ValueError(f"t_start must have shape ({nm}, {MODEL_NRHOX})")
    ps = (ModelPar * nm)(*[ModelPar(*row) for row in pa])
    t_out = np.zeros((nm, MODEL_NRHOX))
    s_out = np.zeros((nm, 5, MODEL_NRHOX))
    iters = np.zeros(nm, np.int32)
    hist = np.zeros((nm, itmax, 3)) if history else None
    ctx._check(ctx.lib.a12_model_converge(
        ctx.h, C.c_int32(nm), ps, _ptr(ts), C.c_int32(itmax), _ptr(t_out), _ptr(s_out),
        iters.ctypes.data_as(C.c_void_p), None if hist is None else _ptr(hist)),
        "a12_model_converge")
    st = {k: s_out[:, i] for i, k in enumerate(("pg", "ptot", "rhox", "abross", "xne"))}
    if history:
        return iters, t_out, st, hist
    return iters, t_out, st


# ---- KAPP: continuum opacity ---------------------------------------------------
KAPP_SRC = ("ahyd", "shyd", "ah2p", "ahmin", "shmin", "sigh", "ahe1", "ahe2", "ahemin", "sighe",
            "ah2coll", "ah2min", "acont_metal", "sigel", "sigh2", "ahline", "shline",
            "acont", "sigmac", "scont", "sline", "aline", "sigmal")
KAPP_CONST = ("hck", "hoverk", "kbol_ev", "theta_coeff", "bnu_prefac", "saha_prefac", "freq_rydh",
              "ryd_ang", "ln10", "coeff_ff", "sigma_thomson", "clight_angs")

_vp = C.c_void_p


class KappStark(C.Structure):
    _fields_ = [("loaded", C.c_int32), ("n_upper_min", C.c_int32), ("n_upper_max", C.c_int32),
                ("n_dens", C.c_int32), ("dens", _vp), ("temp", _vp), ("log_dalpha", _vp),
                ("max_dens_idx", _vp), ("k_alpha", _vp), ("profiles", _vp)]


class KappTables(C.Structure):
    _fields_ = [("k_freqn", _vp), ("k_xn", _vp), ("k_xl", _vp), ("k_ekarzas", _vp),
                ("gff_logtab", _vp),
                ("cia_npair", C.c_int32), ("cia_ntmax", C.c_int32), ("cia_nnu", C.c_int32),
                ("cia_nu0", C.c_double), ("cia_dnu", C.c_double), ("cia_nt", _vp),
                ("cia_temp", _vp), ("cia_tab", _vp), ("cia_use", _vp),
                ("cross_ch", _vp), ("cross_oh", _vp),
                ("hi_nedge", C.c_int32), ("hi_edge", _vp),
                ("felo_nlev", C.c_int32 * 2), ("felo_xnfp_idx", C.c_int32 * 2),
                ("felo_nuth", _vp), ("felo_g", _vp), ("felo_e", _vp), ("felo_ifancy", _vp),
                ("felo_npts", _vp), ("felo_x", _vp), ("felo_logsig", _vp), ("felo_seaton", _vp),
                ("tb_nspecies", C.c_int32), ("tb_nlevels", C.c_int32), ("tb_npoints", C.c_int32),
                ("tb_xnfp_idx", _vp), ("tb_nlev", _vp), ("tb_lowest_thr", _vp), ("tb_g", _vp),
                ("tb_np", _vp), ("tb_ebind", _vp), ("tb_eexc", _vp), ("tb_logemin", _vp),
                ("tb_logestride", _vp), ("tb_ephot", _vp), ("tb_sigma", _vp),
                ("stark", KappStark * 4)]


class KappState(C.Structure):
    _fields_ = [("nrhox", C.c_int32), ("ld", C.c_int32), ("mion", C.c_int32),
                ("t", _vp), ("tkev", _vp), ("tlog", _vp), ("hkt", _vp), ("hckt", _vp),
                ("xne", _vp), ("rho", _vp), ("bmin", _vp), ("bhyd", _vp), ("equilh2", _vp),
                ("xnf", _vp), ("xnfp", _vp)]


class _Keep:
    """A ctypes struct together with the numpy arrays its pointers point into."""

    def __init__(self, struct, arrays):
        self.struct = struct
        self.arrays = arrays


def _arr(keep, a, dtype):
    a = np.ascontiguousarray(a, dtype)
    keep.append(a)
    return a.ctypes.data_as(_vp)


def kapp_tables(t) -> _Keep:
    """a12_kapp_tables from a mapping of arrays in the layout the C-ABI
    documents (the keys of tests/golden/kapp_tables.npz)."""
    keep = []
    T = KappTables()
    T.k_freqn = _arr(keep, t["k_freqn"], np.float32)
    T.k_xn = _arr(keep, t["k_xn"], np.float32)
    T.k_xl = _arr(keep, t["k_xl"], np.float32)
    T.k_ekarzas = _arr(keep, t["k_ekarzas"], np.float32)
    T.gff_logtab = _arr(keep, t["gff_logtab"], np.float64)
    nt = np.asarray(t["cia_nt"] if "cia_nt" in t else t["cia_nciat"], np.int32)
    temp = np.asarray(t["cia_temp"], np.float64)
    tab = np.asarray(t["cia_tab"], np.float64)
    T.cia_npair, T.cia_ntmax, T.cia_nnu = len(nt), temp.shape[1], tab.shape[2]
    T.cia_nu0, T.cia_dnu = (float(x) for x in t["cia_nu"])
    T.cia_nt = _arr(keep, nt, np.int32)
    T.cia_temp = _arr(keep, temp, np.float64)
    T.cia_tab = _arr(keep, tab, np.float64)
    T.cia_use = None
    T.cross_ch = _arr(keep, t["cross_ch"], np.float64)
    T.cross_oh = _arr(keep, t["cross_oh"], np.float64)
    hi = np.asarray(t["hi_edge"], np.float64)
    T.hi_nedge = len(hi)
    T.hi_edge = _arr(keep, hi, np.float64)
    for k in range(2):
        T.felo_nlev[k] = int(t["felo_nlev"][k])
        T.felo_xnfp_idx[k] = int(t["felo_xnfp_idx"][k])
    for k in ("nuth", "g", "e", "x", "logsig", "seaton"):
        setattr(T, "felo_" + k, _arr(keep, t["felo_" + k], np.float64))
    T.felo_ifancy = _arr(keep, t["felo_ifancy"], np.int32)
    T.felo_npts = _arr(keep, t["felo_npts"], np.int32)
    T.tb_nspecies = len(t["tb_nlev"])
    T.tb_nlevels = len(t["tb_g"])
    T.tb_npoints = len(t["tb_sigma"])
    T.tb_xnfp_idx = _arr(keep, t["tb_xnfp_idx"], np.int32)
    T.tb_nlev = _arr(keep, t["tb_nlev"], np.int32)
    T.tb_lowest_thr = _arr(keep, t["tb_lowest_thr"], np.float64)
    T.tb_g = _arr(keep, t["tb_g"], np.int32)
    T.tb_np = _arr(keep, t["tb_np"], np.int32)
    for k in ("ebind", "eexc", "logemin", "logestride", "ephot", "sigma"):
        setattr(T, "tb_" + k, _arr(keep, t["tb_" + k], np.float64))
    for s in range(4):
        hdr = np.asarray(t[f"st{s}_hdr"])
        S = T.stark[s]
        S.loaded, S.n_upper_min, S.n_upper_max, S.n_dens = (int(x) for x in hdr[[0, 2, 3, 4]])
        if S.loaded:
            S.dens = _arr(keep, t[f"st{s}_dens"], np.float64)
            S.temp = _arr(keep, t[f"st{s}_temp"], np.float64)
            S.log_dalpha = _arr(keep, t[f"st{s}_ldalpha"], np.float64)
            S.max_dens_idx = _arr(keep, t[f"st{s}_maxdens"], np.int32)
            S.k_alpha = _arr(keep, t[f"st{s}_kalpha"], np.float64)
            S.profiles = _arr(keep, t[f"st{s}_prof"], np.float64)
    return _Keep(T, keep)


def kapp_state(st, mion: int = 1006) -> _Keep:
    """a12_kapp_state from depth arrays t, tkev, tlog, hkt, hckt, xne, rho,
    bmin, equilh2 (nrhox), bhyd (6, nrhox) and xnf, xnfp (mion, nrhox)."""
    keep = []
    n = len(st["t"])
    S = KappState()
    S.nrhox, S.ld, S.mion = n, n, mion
    for k in ("t", "tkev", "tlog", "hkt", "hckt", "xne", "rho", "bmin", "equilh2"):
        setattr(S, k, _arr(keep, st[k], np.float64))
    S.bhyd = _arr(keep, st["bhyd"], np.float64)
    S.xnf = _arr(keep, st["xnf"], np.float64)
    S.xnfp = _arr(keep, st["xnfp"], np.float64)
    return _Keep(S, keep)


def kapp_constants() -> dict:
    """a12_kapp_constants (no GPU needed)."""
    out = np.zeros(12)
    L = load_library()
    if L.a12_kapp_constants(out.ctypes.data_as(_vp)) != 0:
        raise A12Error("a12_kapp_constants failed")
    return dict(zip(KAPP_CONST, out))


def kapp_set_tables(ctx: Context, tables: _Keep):
    ctx._check(ctx.lib.a12_kapp_set_tables(ctx.h, C.byref(tables.struct)), "a12_kapp_set_tables")


def kapp_set_state(ctx: Context, state: _Keep):
    ctx._check(ctx.lib.a12_kapp_set_state(ctx.h, C.byref(state.struct)), "a12_kapp_set_state")


def kapp_eval(ctx: Context, waves_nm, nrhox: int) -> dict:
    """a12_kapp_eval: every per-source array of KAPP at the given
    wavelengths; returns {name: (nwave, nrhox)} in KAPP_SRC order."""
    w = np.ascontiguousarray(waves_nm, np.float64)
    out = np.zeros((len(w), len(KAPP_SRC), nrhox))
    ctx._check(ctx.lib.a12_kapp_eval(ctx.h, C.c_int32(len(w)), _ptr(w), _ptr(out)), "a12_kapp_eval")
    return {k: np.ascontiguousarray(out[:, i]) for i, k in enumerate(KAPP_SRC)}


def atlas_kapp_run(ctx: Context):
    ctx._check(ctx.lib.a12_atlas_kapp_run(ctx.h), "a12_atlas_kapp_run")


def atlas_kapp_fetch(ctx: Context) -> dict:
    """The six KAPP tables on the device: {acont, sigmac, scont, sline, aline,
    sigmal} each (numnu, nrhox)."""
    n, nf = _atl_dims(ctx)
    out = np.zeros((6, nf, n))
    ctx._check(ctx.lib.a12_atlas_kapp_fetch(ctx.h, _ptr(out)), "a12_atlas_kapp_fetch")
    return dict(zip(("acont", "sigmac", "scont", "sline", "aline", "sigmal"), out))


def kapp_ms(ctx: Context) -> float:
    ms = C.c_float(-1.0)
    ctx._check(ctx.lib.a12_kapp_ms(ctx.h, C.byref(ms)), "a12_kapp_ms")
    return ms.value


# ---- SYNTHE complex-profile lines (a12_cplx_*) ------------------------------
# a12_cplx_line (nlte_line_t, src/mod_mklinelist.f90:43-62) with C's
# alignment: 60 bytes of fields, 64 per record.
CPLX_DTYPE = np.dtype([("wlvac", "<f8"), ("elo", "<f4"), ("gf", "<f4"), ("nblo", "<i4"),
                       ("nbup", "<i4"), ("nelion", "<i4"), ("itype", "<i4"), ("ncon", "<i4"),
                       ("nelionx", "<i4"), ("gammar", "<f4"), ("gammas", "<f4"), ("gammaw", "<f4"),
                       ("nbuff", "<i4"), ("lim", "<i4")], align=True)


class CplxLine(C.Structure):
    _fields_ = [("wlvac", C.c_double), ("elo", C.c_float), ("gf", C.c_float), ("nblo", C.c_int32),
                ("nbup", C.c_int32), ("nelion", C.c_int32), ("itype", C.c_int32), ("ncon", C.c_int32),
                ("nelionx", C.c_int32), ("gammar", C.c_float), ("gammas", C.c_float),
                ("gammaw", C.c_float), ("nbuff", C.c_int32), ("lim", C.c_int32)]


def cplx_lines(recs) -> np.ndarray:
    """Line records (any structured array with the nlte_line_t field
    names) in the C-ABI layout."""
    out = np.zeros(len(recs), CPLX_DTYPE)
    for k in CPLX_DTYPE.names:
        out[k] = recs[k]
    return out


def kapp_stark(t) -> _Keep:
    """The four Stehle series (a12_kapp_stark[4]) from a mapping of arrays in
    the layout of tests/golden/kapp_tables.npz (st{s}_* keys)."""
    keep = []
    S = (KappStark * 4)()
    for s in range(4):
        hdr = np.asarray(t[f"st{s}_hdr"])
        x = S[s]
        x.loaded, x.n_upper_min, x.n_upper_max, x.n_dens = (int(v) for v in hdr[[0, 2, 3, 4]])
        if x.loaded:
            x.dens = _arr(keep, t[f"st{s}_dens"], np.float64)
            x.temp = _arr(keep, t[f"st{s}_temp"], np.float64)
            x.log_dalpha = _arr(keep, t[f"st{s}_ldalpha"], np.float64)
            x.max_dens_idx = _arr(keep, t[f"st{s}_maxdens"], np.int32)
            x.k_alpha = _arr(keep, t[f"st{s}_kalpha"], np.float64)
            x.profiles = _arr(keep, t[f"st{s}_prof"], np.float64)
    return _Keep(S, keep)


def cplx_load(ctx: Context, st: SyntheState, lines: np.ndarray, stark: _Keep, wlend: float,
              cutoff: float, ifvac: int = 1, velshift: float = 0.0):
    """a12_cplx_load: the complex lines of the SYNTHE state loaded on ctx
    (a12_synthe_load must have been called with the same state)."""
    lines = np.ascontiguousarray(lines, dtype=CPLX_DTYPE)
    a = ctx._atm(st)
    ctx._check(ctx.lib.a12_cplx_load(ctx.h, C.byref(a), lines.ctypes.data_as(_vp), C.c_int32(len(lines)),
                                     C.cast(stark.struct, _vp), C.c_double(wlend), C.c_double(cutoff),
                                     C.c_int32(ifvac), C.c_float(velshift)), "a12_cplx_load")
    ctx._cplx = (st.nrhox, st.length, len(lines))


def cplx_run(ctx: Context):
    ctx._check(ctx.lib.a12_cplx_run(ctx.h), "a12_cplx_run")


def cplx_fetch(ctx: Context, extents: bool = False):
    """The last run's buffer (nrhox, length) float32 and, with extents, the
    (nrhox, nlines, 4) first/last red and blue points."""
    n, L, nl = ctx._cplx
    buf = np.zeros((n, L), np.float32)
    ext = np.zeros((n, max(nl, 1), 4), np.int32) if extents else None
    ctx._check(ctx.lib.a12_cplx_fetch(ctx.h, buf.ctypes.data_as(_vp),
                                      None if ext is None else ext.ctypes.data_as(_vp)), "a12_cplx_fetch")
    return (buf, ext[:, :nl]) if extents else buf


def cplx_clear(ctx: Context):
    ctx._check(ctx.lib.a12_cplx_clear(ctx.h), "a12_cplx_clear")


def cplx_ms(ctx: Context):
    """[plans, convolutions, wing extents, accumulation] of the last run (ms)."""
    ms = (C.c_float * 4)()
    ctx._check(ctx.lib.a12_cplx_ms(ctx.h, ms), "a12_cplx_ms")
    return [float(v) for v in ms]
