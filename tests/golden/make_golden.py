#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ from the reference Fortran.

Runs in the development container only (it needs /root/reference and the
reference build from ``make -C oracle/ref``).  It

1. builds a synthetic data root: symlinks to the reference's data tables
   plus a generated ``lines.list`` naming a seeded synthetic gfall file;
2. runs the unmodified reference ``synthe.exe`` on a seeded synthetic deck
   (golden emergent spectrum, the ``.spec`` file);
3. runs ``ref_synthe_dump`` (the reference's own set-up stages) to capture
   the state SYNTHE's hot loops read, trimmed to the continuum-edge
   intervals the window touches;
4. runs ``ref_numerics`` on seeded vectors for the reference's Voigt,
   SOLVIT, MAP1, INTEG, PARCOE, DERIV, EXPI and JOSH routines.

Fixtures are data only (inputs and the reference's outputs).  Usage:
    python tests/golden/make_golden.py [--only NAME ...]
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import shutil
import math
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
from atlas12_amd import workloads as wl  # noqa: E402
from atlas12_amd.records import DIATOMIC_DTYPE  # noqa: E402

REF = "/root/reference"
REFBIN = os.path.join(REPO, "oracle", "_ref", "bin")
MW = 139


def data_root(tmp: str, gfall_path: str) -> str:
    root = os.path.join(tmp, "a12root")
    d = os.path.join(root, "data")
    os.makedirs(d, exist_ok=True)
    for name in os.listdir(os.path.join(REF, "data")):
        if name == "lines.list":
            continue
        os.symlink(os.path.join(REF, "data", name), os.path.join(d, name))
    shutil.copy(gfall_path, os.path.join(d, "gfall_synth.dat"))
    with open(os.path.join(d, "lines.list"), "w") as fh:
        fh.write("gfall\tgfall_synth.dat\n")
    return root


def read_dump(path: str) -> dict:
    with open(path, "rb") as fh:
        buf = fh.read()
    off = 0

    def take(dt, n):
        nonlocal off
        a = np.frombuffer(buf, dtype=dt, count=n, offset=off)
        off += a.nbytes
        return a.copy()
    n, length, ixb, nedge, numnu, mw, nlte, nnlte = take("<i4", 8)
    wbeg, rat, res = take("<f8", 3)
    d = dict(nrhox=int(n), length=int(length), ixwlbeg=int(ixb), nedge=int(nedge),
             numnu=int(numnu), mw=int(mw), nlines=int(nlte), nlines_nlte=int(nnlte),
             wbegin=float(wbeg), ratio=float(rat), resolu=float(res))
    d["wledge"] = np.abs(take("<f8", nedge))
    for k in ("continall", "contabs", "contscat"):
        d[k] = take("<f8", numnu * n).reshape(n, numnu)
    d["xnfpel"] = take("<f8", 6 * mw * n).reshape(n, 6 * mw)
    d["dopple"] = take("<f8", 6 * mw * n).reshape(n, 6 * mw)
    for k in ("t", "xne", "rho", "xnfh", "xnfhe1", "xnfh2", "hckt", "hkt_syn", "hkt_atl", "rhox"):
        d[k] = take("<f8", n)
    d["coefj"] = take("<f8", 51 * 51)
    d["coefh"] = take("<f8", 51 * 51)
    d["lines"] = np.frombuffer(buf, dtype=wl.LTE_DTYPE, count=nlte, offset=off).copy()
    return d


def trim_edges(d: dict) -> dict:
    """Keep only the continuum-edge intervals the grid touches (the
    interpolation reads no others), renumbered from 1."""
    w = d["wledge"]
    L, wb, r = d["length"], d["wbegin"], d["ratio"]
    wave = wb * r ** np.arange(L, dtype=np.float64)
    ne = d["nedge"]
    # interval index per point exactly as the depth loop advances it
    ie = 0
    used = set()
    for x in (wave[0], wave[-1]):
        pass
    for x in wave:
        while x >= w[ie + 1] and ie < ne - 2:
            ie += 1
        used.add(ie)
    lo, hi = min(used), max(used)
    out = dict(d)
    out["wledge"] = w[lo:hi + 2].copy()
    out["nedge"] = hi - lo + 2
    for k in ("continall", "contabs", "contscat"):
        out[k] = d[k][:, 3 * lo:3 * (hi + 1)].copy()
    out["numnu"] = 3 * (hi - lo + 1)
    return out


def save_state(path: str, d: dict, species_nelion=None) -> None:
    """Atmosphere state for the hot path; xnfpel/dopple kept only for the
    ion slots the line list uses (others are zero when rebuilt)."""
    if species_nelion is None:
        species_nelion = np.unique(d["lines"]["nelion"])
    idx = np.asarray(sorted(set(int(x) - 1 for x in species_nelion)), dtype=np.int32)
    keep = {k: d[k] for k in ("t", "xne", "rho", "xnfh", "xnfhe1", "xnfh2", "hckt",
                              "hkt_syn", "hkt_atl", "rhox", "wledge", "continall",
                              "contabs", "contscat")}
    np.savez_compressed(path, nrhox=d["nrhox"], length=d["length"], ixwlbeg=d["ixwlbeg"],
                        wbegin=d["wbegin"], ratio=d["ratio"], resolu=d["resolu"],
                        mw6=6 * MW, ion_idx=idx, xnfpel_sub=d["xnfpel"][:, idx],
                        dopple_sub=d["dopple"][:, idx], **keep)


def read_spec(path: str) -> np.ndarray:
    return np.loadtxt(path)


def run(cmd, cwd, env=None):
    r = subprocess.run(cmd, cwd=cwd, env=env, capture_output=True, text=True)
    if r.returncode != 0:
        print(r.stdout[-3000:], r.stderr[-3000:])
        raise SystemExit(f"command failed: {cmd}")
    return r


def read_linform(path: str, nrhox: int):
    """synthe.exe's .linform (more_output=yes, src/synthe.f90:888-890): per
    wavelength point a record (F12.4, 1P2E12.4) wave, HNU(1), surface flux,
    then TAUNU(1..nrhox) ten per line (10E12.4).  Returns (wave, hnu1,
    surf, taunu (points, nrhox))."""
    per = 1 + (nrhox + 9) // 10
    with open(path) as fh:
        lines = fh.read().splitlines()
    assert len(lines) % per == 0, (len(lines), per)
    npt = len(lines) // per
    wave, h1, surf = np.zeros(npt), np.zeros(npt), np.zeros(npt)
    tau = np.zeros((npt, nrhox))
    for i in range(npt):
        head = lines[i * per]
        wave[i], h1[i], surf[i] = float(head[0:12]), float(head[12:24]), float(head[24:36])
        vals = []
        for ln in lines[i * per + 1:(i + 1) * per]:
            vals += [float(ln[k:k + 12]) for k in range(0, len(ln), 12) if ln[k:k + 12].strip()]
        tau[i] = vals[:nrhox]
    return wave, h1, surf, tau


def synthe_case(name: str, atm: wl.Atmosphere, wlbeg: float, wlend: float, resolu: float,
                nlines: int, seed: int, line_span=None, keep_spec=True, linform=False,
                **gkw) -> dict:
    tmp = tempfile.mkdtemp(prefix="a12g_")
    span = line_span or (wlbeg - 0.5, wlend + 0.5)
    rec = wl.synthetic_gfall(nlines, span[0], span[1], seed, **gkw)
    gpath = os.path.join(tmp, "gfall.dat")
    wl.write_gfall(rec, gpath)
    root = data_root(tmp, gpath)
    deck = os.path.join(tmp, f"{name}.dat")
    with open(deck, "w") as fh:
        fh.write(atm.deck())
    env = dict(os.environ, ATLAS12=root)
    dump = os.path.join(tmp, "dump.bin")
    run([os.path.join(REFBIN, "ref_synthe_dump"), deck, repr(wlbeg), repr(wlend),
         repr(resolu), dump], tmp, env)
    d = read_dump(dump)
    out = {"dump": d, "gfall": rec}
    if keep_spec:
        run([os.path.join(REFBIN, "synthe.exe"), deck, f"wlbeg={wlbeg}", f"wlend={wlend}",
             f"resolu={resolu}"], tmp, env)
        out["spec"] = read_spec(os.path.join(tmp, f"{name}.spec"))
    if linform:
        # the same synthesis with the per-depth optical-depth table
        os.rename(os.path.join(tmp, f"{name}.spec"), os.path.join(tmp, f"{name}.spec0"))
        run([os.path.join(REFBIN, "synthe.exe"), deck, f"wlbeg={wlbeg}", f"wlend={wlend}",
             f"resolu={resolu}", "more_output=yes"], tmp, env)
        spec2 = read_spec(os.path.join(tmp, f"{name}.spec"))
        assert np.array_equal(spec2, out["spec"]), "more_output changed the spectrum"
        out["linform"] = read_linform(os.path.join(tmp, f"{name}.linform"), d["nrhox"])
    shutil.rmtree(tmp)
    return out


def golden_synthe_small():
    """5 nm window at R = 500,000 on the plain-SYNTHE atmosphere; line
    density of the 200k-line workload, plus lines just outside the window
    (off-grid centres) and a few strong resonance-like lines."""
    atm = wl.scurve_atmosphere()
    c = synthe_case("synsmall", atm, 500.0, 505.0, 500000.0, 4000, seed=11,
                    line_span=(499.5, 505.5), linform=True)
    d = trim_edges(c["dump"])
    save_state(os.path.join(HERE, "synthe_small_state.npz"), d)
    grid = wl.Grid.make(500.0, 505.0, 500000.0)
    np.savez_compressed(os.path.join(HERE, "synthe_small_ref.npz"),
                        lines=d["lines"].view(np.uint8), spec=c["spec"],
                        gfall=c["gfall"], grid_length=grid.length,
                        linform_wave=c["linform"][0], linform_hnu1=c["linform"][1],
                        linform_surf=c["linform"][2], linform_taunu=c["linform"][3])
    np.savez_compressed(os.path.join(HERE, "josh_coef.npz"), coefj=d["coefj"],
                        coefh=d["coefh"])
    return {"synthe_small": {"nlines": int(d["nlines"]), "length": d["length"],
                             "nedge": d["nedge"]}}


def golden_state_plain():
    """Atmosphere state of the plain-SYNTHE workload (400-700 nm, R=5e5) for
    full-size runs; lines are regenerated from their seed on any machine."""
    atm = wl.scurve_atmosphere()
    c = synthe_case("plain", atm, 400.0, 700.0, 500000.0, 2000, seed=1, keep_spec=False)
    d = trim_edges(c["dump"])
    nel = sorted({int(round(x)) for x in (26 * 6 - 6 + 1, 26 * 6 - 6 + 2, 20 * 6 - 6 + 1,
                                          20 * 6 - 6 + 2, 22 * 6 - 6 + 1, 11 * 6 - 6 + 1,
                                          12 * 6 - 6 + 1)})
    save_state(os.path.join(HERE, "synthe_plain_state.npz"), d, nel)
    return {"synthe_plain_state": {"length": d["length"], "nedge": d["nedge"]}}


def _ref_numerics(op, payload: bytes, outbytes: int, extra=()):
    tmp = tempfile.mkdtemp(prefix="a12n_")
    fin, fout = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
    with open(fin, "wb") as fh:
        fh.write(payload)
    run([os.path.join(REFBIN, "ref_numerics"), op, fin, fout, *extra], tmp)
    with open(fout, "rb") as fh:
        b = fh.read()
    shutil.rmtree(tmp)
    assert len(b) == outbytes, (op, len(b), outbytes)
    return b


def golden_numerics():
    rng = np.random.default_rng(7)
    out = {}
    # Voigt: a grid over the regimes plus random points
    v = np.concatenate([np.linspace(0, 20, 201), rng.uniform(0, 30, 3000),
                        rng.uniform(0, 1, 1000), np.zeros(50)]).astype(np.float32)
    a = np.concatenate([np.full(201, 0.01), 10 ** rng.uniform(-6, 1.5, 3000),
                        10 ** rng.uniform(-4, 0, 1000), 10 ** np.linspace(-6, 1, 50)]).astype(np.float32)
    n = len(v)
    h = np.frombuffer(_ref_numerics("voigt", np.int32(n).tobytes() + v.tobytes() + a.tobytes(),
                                    4 * n), dtype="<f4")
    vec = {"voigt_v": v, "voigt_a": a, "voigt_h": h}
    # SOLVIT on I - diag(alpha) C style systems and random dense systems
    coef = np.load(os.path.join(HERE, "josh_coef.npz"))
    cj = coef["coefj"].reshape(51, 51, order="F")
    mats, rhs = [], []
    for k in range(24):
        if k < 16:
            al = rng.uniform(0, 0.999 if k % 2 else 0.3, 51)
            A = np.eye(51) - al[:, None] * cj
        else:
            A = rng.normal(size=(51, 51)) + 3 * np.eye(51)
        mats.append(A)
        rhs.append(rng.uniform(0.1, 2.0, 51))
    pay = np.int32([51, len(mats)]).tobytes()
    for A, b in zip(mats, rhs):
        pay += np.asfortranarray(A).tobytes(order="F") + b.tobytes()
    x = np.frombuffer(_ref_numerics("solvit", pay, 8 * 51 * len(mats)), "<f8").reshape(-1, 51)
    vec.update(solvit_a=np.stack(mats), solvit_b=np.stack(rhs), solvit_x=x)
    # MAP1 / INTEG / PARCOE / DERIV on log-spaced optical-depth-like grids
    nold, nnew, cnt = 72, 51, 20
    xo = np.sort(10 ** rng.uniform(-6, 3, (cnt, nold)), axis=1)
    fo = np.exp(-xo / rng.uniform(0.5, 50, (cnt, 1))) + rng.uniform(0, 1, (cnt, 1)) * np.log1p(xo)
    xn = np.tile(np.array([0.0, 3.2e-6, 5.6e-6, 1e-5, 1.8e-5, 3.2e-5, 5.6e-5, 1e-4, 1.8e-4, 3.2e-4,
                           5.6e-4, 1e-3, 1.8e-3, 3.2e-3, 5.6e-3, 0.01, 0.016, 0.025, 0.042, 0.065,
                           0.096, 0.139, 0.196, 0.273, 0.375, 0.5, 0.63, 0.78, 0.95, 1.15, 1.35, 1.6,
                           1.85, 2.15, 2.45, 2.75, 3.15, 3.65, 4.25, 5.0, 6.0, 7.0, 8.0, 9.0, 10.0,
                           11.5, 13.0, 14.5, 16.0, 18.0, 20.0]), (cnt, 1))
    xn[cnt // 2:] *= rng.uniform(0.5, 5e3, (cnt - cnt // 2, 1))
    pay = np.int32([nold, nnew, cnt]).tobytes()
    for k in range(cnt):
        pay += xo[k].tobytes() + fo[k].tobytes() + xn[k].tobytes()
    b = _ref_numerics("map1", pay, cnt * (8 * nnew + 4))
    fn = np.zeros((cnt, nnew))
    ret = np.zeros(cnt, np.int32)
    rec = np.dtype([("f", "<f8", nnew), ("r", "<i4")])
    arr = np.frombuffer(b, rec)
    fn[:], ret[:] = arr["f"], arr["r"]
    vec.update(map1_xold=xo, map1_fold=fo, map1_xnew=xn, map1_fnew=fn, map1_ret=ret)
    st = rng.uniform(0, 1e-3, cnt)
    pay = np.int32([nold, cnt]).tobytes()
    for k in range(cnt):
        pay += xo[k].tobytes() + fo[k].tobytes() + np.float64(st[k]).tobytes()
    fi = np.frombuffer(_ref_numerics("integ", pay, 8 * nold * cnt), "<f8").reshape(cnt, nold)
    vec.update(integ_x=xo, integ_f=fo, integ_start=st, integ_out=fi)
    pay = np.int32([nold, cnt]).tobytes()
    for k in range(cnt):
        pay += fo[k].tobytes() + xo[k].tobytes()
    pc = np.frombuffer(_ref_numerics("parcoe", pay, 24 * nold * cnt), "<f8").reshape(cnt, 3, nold)
    vec.update(parcoe_out=pc)
    pay = np.int32([nold, cnt]).tobytes()
    for k in range(cnt):
        pay += xo[k].tobytes() + fo[k].tobytes()
    dv = np.frombuffer(_ref_numerics("deriv", pay, 8 * nold * cnt), "<f8").reshape(cnt, nold)
    vec.update(deriv_out=dv)
    # EXPI over its three ranges and the recurrence
    ex = np.concatenate([10 ** rng.uniform(-6, 0, 200), rng.uniform(1, 4, 200),
                         rng.uniform(4, 80, 200), [0.5, 1.0, 4.0, 1e-8]])
    nn = rng.integers(1, 5, len(ex)).astype(np.int32)
    pay = np.int32(len(ex)).tobytes() + nn.tobytes() + ex.tobytes()
    e = np.frombuffer(_ref_numerics("expi", pay, 8 * len(ex)), "<f8")
    vec.update(expi_n=nn, expi_x=ex, expi_out=e)
    np.savez_compressed(os.path.join(HERE, "numerics_ref.npz"), **vec)
    out["numerics"] = {k: list(np.shape(v)) for k, v in vec.items()}
    return out


def golden_josh():
    """JOSH on the small-case atmosphere at several wavelengths, both output
    modes, with scattering and line opacity of several strengths, plus the
    deep-only branch (TAUNU(1) > 20)."""
    st = np.load(os.path.join(HERE, "synthe_small_state.npz"))
    rng = np.random.default_rng(5)
    n = int(st["nrhox"])
    rhox = st["rhox"]
    hkt = st["hkt_atl"]
    cases = []
    for k in range(40):
        wave = rng.uniform(300, 2000)
        freq = 2.99792458e17 / wave
        ehv = np.exp(-freq * hkt)
        stim = 1 - ehv
        bnu = 2 * 6.62607015e-27 / 2.99792458e10 ** 2 * 1e45 * (freq / 1e15) ** 3 * ehv / stim
        acont = 10 ** rng.uniform(-2, 0, 1) * (rhox / rhox[0]) ** 0.3 * np.exp(rng.normal(0, 0.1, n))
        sigmac = acont * 10 ** rng.uniform(-4, 0.5)
        aline = acont * 10 ** rng.uniform(-3, 4) * np.exp(-np.linspace(0, 3, n) * rng.uniform(0, 2))
        if k % 5 == 0:
            aline[:] = 0
        sigmal = np.zeros(n) if k % 3 else aline * rng.uniform(0, 0.5)
        if k == 39:
            acont = acont * 1e6  # TAUNU(1) > 20
        ifscat = 1 if k % 7 else 0
        ifsurf = k % 2
        cases.append((ifscat, ifsurf, acont, aline, sigmac, sigmal, bnu.copy(), bnu.copy(), rhox, bnu))
    pay = np.int32([n, len(cases)]).tobytes()
    for c in cases:
        pay += np.int32([c[0], c[1]]).tobytes()
        for arr in c[2:]:
            pay += np.asarray(arr, np.float64).tobytes()
    data = _ref_numerics("josh", pay, len(cases) * 8 * (2 + 5 * n),
                         extra=(os.path.join(REF, "data") + "/",))
    o = np.frombuffer(data, "<f8").reshape(len(cases), 2 + 5 * n)
    np.savez_compressed(os.path.join(HERE, "josh_ref.npz"),
                        ifscat=np.int32([c[0] for c in cases]),
                        ifsurf=np.int32([c[1] for c in cases]),
                        inputs=np.stack([np.stack(c[2:10]) for c in cases]),
                        out=o)
    return {"josh": {"cases": len(cases)}}


def read_atlas_dump(path: str) -> dict:
    with open(path, "rb") as fh:
        buf = fh.read()
    off = 0

    def take(dt, n):
        nonlocal off
        a = np.frombuffer(buf, dtype=dt, count=n, offset=off)
        off += a.nbytes
        return a.copy()
    n, numnu, nl, mion, nwave = (int(x) for x in take("<i4", 5))
    d = dict(nrhox=n, numnu=numnu, nlines=nl, mion=mion, nwave1=nwave + 1)
    d["teff"], d["flux"] = (float(x) for x in take("<f8", 2))
    d["waveset"] = take("<f8", numnu)
    d["rcoset"] = take("<f8", numnu)
    for k in ("t", "hkt", "hckt", "xne", "rhox", "xnf1", "xnf3", "xnf841"):
        d[k] = take("<f8", n)
    d["xnfdop"] = take("<f8", n * mion).reshape(mion, n)
    d["dopple"] = take("<f8", n * mion).reshape(mion, n)
    d["tabcont"] = take("<f8", n * (nwave + 1)).reshape(nwave + 1, n)
    d["iwavetab"] = take("<i4", nwave + 1)
    for k in ("h0", "h1", "h2"):
        d[k] = take("<f8", 2001)
    d["coefj"] = take("<f8", 2601)
    d["coefh"] = take("<f8", 2601)
    d["xlines"] = take("<f8", n * numnu).reshape(numnu, n)
    kapp = take("<f8", 6 * n * numnu).reshape(numnu, 6, n)
    for i, k in enumerate(("acont", "sigmac", "scont", "sline", "aline", "sigmal")):
        d[k] = np.ascontiguousarray(kapp[:, i, :])
    acc = take("<f8", 5 * n + 1 + 3 * n)
    for i, k in enumerate(("rjmins", "flxrad", "raden", "accrad", "abross")):
        d["acc_" + k] = acc[i * n:(i + 1) * n]
    d["acc_pradk0"] = acc[5 * n]
    d["last_rdiagjnu"] = acc[5 * n + 1:6 * n + 1]
    d["last_hnu"] = acc[6 * n + 1:7 * n + 1]
    d["last_taunu"] = acc[7 * n + 1:8 * n + 1]
    return d


KAPP_SRC = ("ahyd", "shyd", "ah2p", "ahmin", "shmin", "sigh", "ahe1", "ahe2", "ahemin", "sighe",
            "ah2coll", "ah2min", "acont_metal", "sigel", "sigh2", "ahline", "shline",
            "acont", "sigmac", "scont", "sline", "aline", "sigmal")
KAPP_CONST = ("hck", "hoverk", "kbol_ev", "theta_coeff", "bnu_prefac", "saha_prefac", "freq_rydh",
              "ryd_ang", "ln10", "coeff_ff", "sigma_thomson", "clight_angs")


def read_kapp_dump(path):
    """ref_atlas_dump's A12_KAPP_DUMP file: KAPP's depth inputs, the tables
    its sources hold in module memory after first use, and the per-source
    arrays at every requested wavelength."""
    b = open(path, "rb").read()
    off = 0

    def take(dt, n):
        nonlocal off
        a = np.frombuffer(b, dtype=dt, count=n, offset=off)
        off += a.nbytes
        return a.copy()
    n, mion, kw, nw = (int(x) for x in take("<i4", 4))
    d = dict(nrhox=n, mion=mion)
    d["const"] = dict(zip(KAPP_CONST, take("<f8", len(KAPP_CONST))))
    st = {}
    for k in ("t", "tkev", "tlog", "hkt", "hckt", "xne", "rho", "bmin"):
        st[k] = take("<f8", n)
    st["bhyd"] = take("<f8", 6 * n).reshape(6, n)
    st["equilh2"] = take("<f8", n)
    st["xnf"] = take("<f8", n * mion).reshape(mion, n)
    st["xnfp"] = take("<f8", n * mion).reshape(mion, n)
    d["state"] = st
    tb = {}
    tb["q_holtsmark"] = take("<f8", 150)
    tb["log_beta_min"], tb["log_beta_step"] = take("<f8", 2)
    nsp = int(take("<i4", 1)[0])
    sp_idx, sp_nlev, sp_thr = [], [], []
    lv = {k: [] for k in ("g", "np", "ebind", "eexc", "logemin", "logestride")}
    ephot, sigma = [], []
    for _ in range(nsp):
        sp_idx.append(int(take("<i4", 1)[0]))
        nl = int(take("<i4", 1)[0])
        sp_nlev.append(nl)
        sp_thr.append(float(take("<f8", 1)[0]))
        for _ in range(nl):
            g, npt = (int(x) for x in take("<i4", 2))
            eb, ee, lmin, lstr = take("<f8", 4)
            lv["g"].append(g)
            lv["np"].append(npt)
            lv["ebind"].append(eb)
            lv["eexc"].append(ee)
            lv["logemin"].append(lmin)
            lv["logestride"].append(lstr)
            ephot.append(take("<f8", npt))
            sigma.append(take("<f8", npt))
    tb["tb_xnfp_idx"] = np.array(sp_idx, np.int32)
    tb["tb_nlev"] = np.array(sp_nlev, np.int32)
    tb["tb_lowest_thr"] = np.array(sp_thr)
    tb["tb_g"] = np.array(lv["g"], np.int32)
    tb["tb_np"] = np.array(lv["np"], np.int32)
    for k in ("ebind", "eexc", "logemin", "logestride"):
        tb["tb_" + k] = np.array(lv[k])
    tb["tb_ephot"] = np.concatenate(ephot)
    tb["tb_sigma"] = np.concatenate(sigma)
    nhi = int(take("<i4", 1)[0])
    tb["hi_edge"] = take("<f8", 7 * nhi).reshape(nhi, 7)
    tb["felo_nlev"] = take("<i4", 2)
    tb["felo_xnfp_idx"] = take("<i4", 2)
    for k in ("nuth", "g", "e"):
        tb["felo_" + k] = take("<f8", 120).reshape(2, 60)
    tb["felo_ifancy"] = take("<i4", 120).reshape(2, 60)
    tb["felo_npts"] = take("<i4", 120).reshape(2, 60)
    tb["felo_x"] = take("<f8", 200 * 120).reshape(2, 60, 200)
    tb["felo_logsig"] = take("<f8", 200 * 120).reshape(2, 60, 200)
    tb["felo_seaton"] = take("<f8", 4 * 120).reshape(2, 60, 4)
    for s in range(4):
        loaded, nlo, numin, numax, ndens = (int(x) for x in take("<i4", 5))
        tb[f"st{s}_hdr"] = np.array([loaded, nlo, numin, numax, ndens], np.int32)
        tb[f"st{s}_dens"] = take("<f8", 20)
        tb[f"st{s}_temp"] = take("<f8", 10)
        tb[f"st{s}_ldalpha"] = take("<f8", 60)
        if loaded:
            nt = numax - numin + 1
            tb[f"st{s}_maxdens"] = take("<i4", nt)
            tb[f"st{s}_kalpha"] = take("<f8", nt)
            tb[f"st{s}_prof"] = take("<f8", 60 * 10 * ndens * nt).reshape(nt, ndens, 10, 60)
    d["tables"] = tb
    src = take("<f8", nw * len(KAPP_SRC) * n).reshape(nw, len(KAPP_SRC), n)
    d["src"] = {k: np.ascontiguousarray(src[:, i, :]) for i, k in enumerate(KAPP_SRC)}
    assert off == len(b), (off, len(b))
    return d


def atlas_case(atm, nustart, nuhi, linedata, raw=None, tcorr=False, kapp_waves=None,
               sel_files=None, sel_mols=()):
    """Runs ref_atlas_dump.  raw = (predicted records, TiO MOL2 records):
    the store is then SELECTLINES' selection from these lists, returned
    as d["selected"] with d["wavetab"].  sel_files {name: bytes} are
    further files of the selection directory (lowobsat12.bin, hilines.bin,
    h2opokazatel.bin, h3plus.dat, ASCII molecular lists), sel_mols the
    ASCII lists lines.list names; read_diatomics_for_atlas's records come
    back as d["diatomic"]."""
    tmp = tempfile.mkdtemp(prefix="a12a_")
    g = os.path.join(tmp, "gfall.dat")
    open(g, "w").close()
    root = data_root(tmp, g)
    deck = os.path.join(tmp, "model.dat")
    with open(deck, "w") as fh:
        fh.write(atm.deck())
    lp = os.path.join(tmp, "lines.bin")
    if linedata is not None:
        with open(lp, "wb") as fh:
            fh.write(np.int32(len(linedata)).tobytes() + np.ascontiguousarray(linedata).tobytes())
    out = os.path.join(tmp, "dump.bin")
    extra = []
    if raw is not None:
        sel = os.path.join(tmp, "sel")
        os.makedirs(sel)
        with open(os.path.join(sel, "gfpred29dec2014.bin"), "wb") as fh:
            fh.write(np.ascontiguousarray(raw[0], "<i4").tobytes())
        wl.write_mol2(raw[1], os.path.join(sel, "tio.bin"))
        for name, data in (sel_files or {}).items():
            with open(os.path.join(sel, name), "wb") as fh:
                fh.write(data)
        with open(os.path.join(sel, "lines.list"), "w") as fh:
            fh.write("mol tio.bin\n")
            for name in sel_mols:
                fh.write(f"mol {name}\n")
        extra = [sel]
    env = dict(os.environ, ATLAS12=root)
    tco = os.path.join(tmp, "tcorr.bin")
    if tcorr:
        env["A12_TCORR_OUT"] = tco
    if kapp_waves is not None:
        kw_path = os.path.join(tmp, "kapp_waves.bin")
        with open(kw_path, "wb") as fh:
            fh.write(np.int32(len(kapp_waves)).tobytes() + np.asarray(kapp_waves, "<f8").tobytes())
        env["A12_KAPP_WAVES"] = kw_path
        env["A12_KAPP_DUMP"] = os.path.join(tmp, "kapp.bin")
    run([os.path.join(REFBIN, "ref_atlas_dump"), deck, str(nustart), str(nuhi), lp, out, *extra],
        tmp, env)
    d = read_atlas_dump(out)
    if kapp_waves is not None:
        d["kapp"] = read_kapp_dump(env["A12_KAPP_DUMP"])
    if tcorr:
        d["tcorr_in"] = read_tcorr_in(tco)
        d["tcorr_iter"] = read_iter_table(tco + ".iter", d["tcorr_in"]["n"])
        d["tcorr_fin"] = read_tcorr_fin(tco + ".fin", d["tcorr_in"]["n"])
        if os.path.exists(tco + ".cnv"):
            d["convec"] = read_convec(tco + ".cnv")
    if raw is not None:
        with open(out + ".sel", "rb") as fh:
            buf = fh.read()
        ns = int(np.frombuffer(buf, "<i4", 1)[0])
        d["selected"] = np.frombuffer(buf, "<i4", 4 * ns, 4).reshape(ns, 4).copy()
        d["wavetab"] = np.frombuffer(buf, "<f8", 361, 4 + 16 * ns).copy()
        with open(out + ".dia", "rb") as fh:
            buf = fh.read()
        nd = int(np.frombuffer(buf, "<i4", 1)[0])
        dia = np.zeros(nd, DIATOMIC_DTYPE)
        for k, name in enumerate(("wlvac_nm", "gflog_dex", "elo_cm", "gammar_log", "gammaw_log")):
            dia[name] = np.frombuffer(buf, "<f8", nd, 4 + 8 * nd * k)
        dia["molcode"] = np.frombuffer(buf, "<i4", nd, 4 + 40 * nd)
        assert len(buf) == 4 + 44 * nd
        d["diatomic"] = dia
    shutil.rmtree(tmp)
    return d


TCORR_FIELDS = ("tauros", "rhox", "t", "dltdlp", "abross", "flxcnv", "flxcnv0", "grdadb",
                "ptotal", "rho", "dlrdlt", "heatcp", "hscale", "flxrad", "rjmins")
ITER_COLS = ("j", "logtau", "t", "dtlamb", "dtsurf", "dtflux", "t1", "flxerr", "flxdrv",
             "r_newt", "dltdlp", "grdadb", "hratio", "p", "xne", "height", "accrad")
# FORMAT(I3, F8.3, F10.1, 4F8.1, 1X,ES11.2, 1X,ES11.2, 1X,ES10.2, 2F8.3, 1X,ES11.2,
#        1X,ES12.3, 1X,ES12.3, 1X,ES10.1, 1X,ES11.2), src/atlas12_modules.f90:2765-2774
ITER_WIDTHS = (3, 8, 10, 8, 8, 8, 8, 12, 12, 11, 8, 8, 12, 13, 13, 11, 12)


def read_tcorr_in(path):
    b = open(path, "rb").read()
    n, ifconv, it = (int(x) for x in np.frombuffer(b, "<i4", 3))
    flux, teff, mixlth = np.frombuffer(b, "<f8", 3, 12)
    arr = np.frombuffer(b, "<f8", len(TCORR_FIELDS) * n, 36).reshape(len(TCORR_FIELDS), n)
    d = dict(n=n, ifconv=ifconv, iter=it, flux=flux, teff=teff, mixlth=mixlth)
    d.update({k: arr[i].copy() for i, k in enumerate(TCORR_FIELDS)})
    return d


def read_iter_table(path, n):
    """The per-iteration table TCORR(3) prints: n rows after its 3 header lines."""
    rows = []
    for line in open(path).read().splitlines():
        if len(rows) == n:
            break
        if len(line) < 60 or not line[:3].strip().isdigit():
            continue
        vals, o = [], 0
        for w in ITER_WIDTHS:
            vals.append(float(line[o:o + w]))
            o += w
        rows.append(vals)
    a = np.array(rows)
    return {k: a[:, i] for i, k in enumerate(ITER_COLS)}


class _Stream:
    """Sequential reader of a Fortran stream-access file."""

    def __init__(self, path):
        self.b = open(path, "rb").read()
        self.o = 0

    def i4(self, cnt=1):
        a = np.frombuffer(self.b, "<i4", cnt, self.o).copy()
        self.o += 4 * cnt
        return a

    def f8(self, cnt=1):
        a = np.frombuffer(self.b, "<f8", cnt, self.o).copy()
        self.o += 8 * cnt
        return a


FIN_BEFORE = ("p", "xne", "prad", "pturb", "vturb", "bmin", "accrad")
FIN_AFTER = ("rhox", "t", "tk", "hkt", "hckt", "tkev", "tlog", "p", "xne", "abross", "prad",
             "pradk", "vturb", "bmin", "pturb", "tauros")
CZ_IN = ("flxcnv", "flxrad", "dltdlp", "grdadb", "t", "rhox", "ptotal", "rho", "heatcp",
         "dlrdlt", "p", "vturb", "abross", "hscale")


def read_tcorr_fin(path, n):
    """ref_atlas_dump's <A12_TCORR_OUT>.fin: the inputs of TCORR(3)'s sections
    G2-K, TTAUP on its own, the module state TCORR(3) leaves, ROSSTAB lookups
    in modes 1-3 and (decks that convect) three CZ_CONSTRUCT calls."""
    r = _Stream(path)
    d = {}
    nross, irosstab, j1s, j2s, numits, itemp = (int(x) for x in r.i4(6))
    d.update(nross=nross, irosstab=irosstab, j1smooth=j1s, j2smooth=j2s, numits=numits,
             itemp=itemp)
    (d["grav"], d["tau1lg"], d["steplg"], d["pradk0"], d["tfloor"], d["wtjm1"], d["wtj"],
     d["wtjp1"]) = r.f8(8)
    d["rt_zero"] = r.f8(4)  # ZEROT, SLOPET, ZEROP, SLOPEP
    d["rt_tab"] = r.f8(3 * nross).reshape(3, nross)  # TABT, TABP, ROSS
    for k in FIN_BEFORE:
        d["in_" + k] = r.f8(n)
    d["in_bhyd"] = r.f8(6 * n).reshape(6, n)
    for k in ("tt_t", "tt_prad", "tt_ab", "tt_ptot", "tt_p"):
        d[k] = r.f8(n)
    for k in FIN_AFTER:
        d["out_" + k] = r.f8(n)
    d["out_bhyd"] = r.f8(6 * n).reshape(6, n)
    d["out_flxerr"] = r.f8(n)
    d["out_itemp"], d["out_czc_engaged"], d["out_czc_jc"] = (int(x) for x in r.i4(3))
    nross2, nq = (int(x) for x in r.i4(2))
    d["lk_tab"] = r.f8(3 * nross2).reshape(3, nross2)
    d["lk_q"] = r.f8(2 * nq).reshape(2, nq)
    d["lk_r"] = r.f8(3 * nq).reshape(3, nq)
    d["cz"] = int(r.i4(1)[0])
    if d["cz"]:
        for k in CZ_IN:
            d["cz_" + k] = r.f8(n)
        d["cz_grav"], d["cz_flux"], d["cz_mixlth"] = r.f8(3)
        d["cz_4pt"] = int(r.i4(1)[0])
        errs, t1s, eng, jcs = [], [], [], []
        for _ in range(3):
            errs.append(r.f8(n))
            t1s.append(r.f8(n))
            e, jc = r.i4(2)
            eng.append(int(e))
            jcs.append(int(jc))
        d["cz_flxerr"], d["cz_t1"] = np.array(errs), np.array(t1s)
        d["cz_engaged"], d["cz_jc"] = np.array(eng, np.int32), np.array(jcs, np.int32)
    assert r.o == len(r.b), (r.o, len(r.b))
    return d


CNV_IN = ("t", "tk", "hkt", "hckt", "tkev", "tlog", "p", "ptotal", "rho", "pradk", "tauros",
          "rhox", "abross", "vturb", "height", "flxrad", "flxcnv1")
CNV_STATE = ("t", "tk", "hkt", "hckt", "tkev", "tlog", "p")
CNV_OUT = ("dltdlp", "heatcp", "velsnd", "dlrdlt", "grdadb", "hscale", "vconv", "flxcnv",
           "flxcnv0", "flxcnv1")


def read_convec(path):
    """ref_atlas_dump's <A12_TCORR_OUT>.cnv: CONVEC's inputs, the four EOS
    evaluations (EDENS, RHO) it makes, the ROSSTAB table, and what CONVEC
    leaves in the module arrays."""
    r = _Stream(path)
    n, ifcorr, nconv, use4pt, smooth = (int(x) for x in r.i4(5))
    d = dict(n=n, ifcorr=ifcorr, nconv=nconv, use4pt=use4pt, smooth=smooth)
    d["grav"], d["mixlth"], d["flux"], d["overwt"] = r.f8(4)
    for k in CNV_IN:
        d["in_" + k] = r.f8(n)
    d["edens"] = r.f8(4 * n).reshape(4, n)
    d["rho4"] = r.f8(4 * n).reshape(4, n)
    nross = int(r.i4(1)[0])
    d["rt_zero"] = r.f8(4)
    d["rt_tab"] = r.f8(3 * nross).reshape(3, nross)
    for k in CNV_STATE:
        d["st_" + k] = r.f8(n)
    d["st_itemp"] = int(r.i4(1)[0])
    for k in CNV_OUT:
        d["out_" + k] = r.f8(n)
    for k in ("flxcnv", "flxcnv0", "flxcnv1"):  # the same call with OVERWT = 0.5
        d["ovr_" + k] = r.f8(n)
    assert r.o == len(r.b), (r.o, len(r.b))
    return d


def golden_tcorr():
    """TCORR(3) after a full 30,000-frequency pass (100,000 raw lines through
    SELECTLINES) on the plain-SYNTHE atmosphere: the state it reads, RDABH
    and RDIAGJ from the oracle's RT pass on the reference's own KAPP tables
    and XLINES (TCORR keeps those two private), and the per-iteration table
    it prints (DTLAMB, DTSURF, DTFLUX, T1 to 0.1 K; ERROR, DERIV to 3
    digits; CONV/TOT)."""
    from oracle import oracle as orc
    atm = wl.scurve_atmosphere()
    st = np.load(os.path.join(HERE, "atlas_plain_state.npz"))
    mols = tuple(int(m) for m in st["mol_nelion"])
    pred, tio = wl.atlas_iteration_lists(100_000, 95.0, 90_000.0, seed=41, molecules=mols)
    d = atlas_case(atm, 9599, 30000, None, raw=(pred, tio), tcorr=True)
    orc.build()
    cj, ch = d["coefj"], d["coefh"]
    acc, _ = orc.rt_pass(d, d["xlines"], cj, ch)
    for k in ("rjmins", "flxrad", "raden", "accrad", "abross"):
        assert np.array_equal(acc[k], d["acc_" + k]), k  # the oracle's pass is the reference's
    ti, tab = d["tcorr_in"], d["tcorr_iter"]
    keep = {k: ti[k] for k in TCORR_FIELDS}
    keep.update({"iter_" + k: v for k, v in tab.items()})
    keep.update({"fin_" + k: v for k, v in d["tcorr_fin"].items()})
    keep.update({"cnv_" + k: v for k, v in d["convec"].items()})
    np.savez_compressed(os.path.join(HERE, "tcorr_ref.npz"), n=ti["n"], ifconv=ti["ifconv"],
                        iter=ti["iter"], flux=ti["flux"], teff=ti["teff"], mixlth=ti["mixlth"],
                        rdabh=acc["rdabh"], rdiagj=acc["rdiagj"], **keep)
    return {"tcorr": {"selected_lines": int(d["nlines"]), "ifconv": ti["ifconv"],
                      "rows": int(len(tab["j"]))}}


def golden_tcorr_conv():
    """TCORR(3) on a deck whose deep layers convect: the plain-SYNTHE RHOX
    grid with T rising from 4,000 to 16,000 K, after a full 30,000-frequency
    pass (20,000 raw lines through SELECTLINES).  The reference's CONVEC gives
    convective flux at 20 layers (CONV/TOT up to 0.59), so the convective
    branches of sections A-F are exercised: the DDEL factor, the convective
    terms of the gradient, the DTLAMB halving above convective layers and no
    damping where CONV/TOT > 0."""
    from oracle import oracle as orc
    atm = wl.scurve_atmosphere(t_top=4000.0, t_bot=16000.0)
    st = np.load(os.path.join(HERE, "atlas_plain_state.npz"))
    mols = tuple(int(m) for m in st["mol_nelion"])
    pred, tio = wl.atlas_iteration_lists(20_000, 95.0, 90_000.0, seed=41, molecules=mols)
    d = atlas_case(atm, 9599, 30000, None, raw=(pred, tio), tcorr=True)
    orc.build()
    acc, _ = orc.rt_pass(d, d["xlines"], d["coefj"], d["coefh"])
    for k in ("rjmins", "flxrad", "raden", "accrad", "abross"):
        assert np.array_equal(acc[k], d["acc_" + k]), k
    ti, tab = d["tcorr_in"], d["tcorr_iter"]
    assert ti["ifconv"] == 1 and np.sum(ti["flxcnv"] > 0) > 5, "no convective layers"
    keep = {k: ti[k] for k in TCORR_FIELDS}
    keep.update({"iter_" + k: v for k, v in tab.items()})
    keep.update({"fin_" + k: v for k, v in d["tcorr_fin"].items()})
    keep.update({"cnv_" + k: v for k, v in d["convec"].items()})
    np.savez_compressed(os.path.join(HERE, "tcorr_conv_ref.npz"), n=ti["n"],
                        ifconv=ti["ifconv"], iter=ti["iter"], flux=ti["flux"], teff=ti["teff"],
                        mixlth=ti["mixlth"], rdabh=acc["rdabh"], rdiagj=acc["rdiagj"], **keep)
    return {"tcorr_conv": {"selected_lines": int(d["nlines"]),
                           "convective_layers": int(np.sum(ti["flxcnv"] > 0)),
                           "max_conv_tot": float(tab["hratio"].max())}}


class _Records:
    """Fortran list-directed reading over the records (lines) of a text
    file: read(k) takes k values from as many records as needed and
    drops the rest of the last one, as READ(u, *) does."""

    def __init__(self, path, skip=0):
        self.lines = open(path).read().splitlines()[skip:]
        self.i = 0

    def line(self):
        s = self.lines[self.i]
        self.i += 1
        return s

    def read(self, k, conv=float):
        out = []
        while len(out) < k:
            out += [conv(t.replace("D", "E")) for t in self.line().split()]
        return out[:k]


def read_kapp_files(datadir):
    """The KAPP tables the reference keeps in SAVE variables of its source
    functions, read from the data files as those functions do: the Karzas &
    Latter tables (REAL(4)), the van Hoof free-free Gaunt table (stored as
    log10 g), the CIA tables and the CH/OH cross sections."""
    t = {}
    r = _Records(os.path.join(datadir, "karzas_ekarzas.dat"), 2)
    t["k_ekarzas"] = np.array(r.read(29), np.float32)
    r = _Records(os.path.join(datadir, "karzas_xn.dat"), 3)
    t["k_xn"] = np.array([r.read(1)[0] for _ in range(15 * 29)], np.float32).reshape(15, 29)
    r = _Records(os.path.join(datadir, "karzas_freqn.dat"), 2)
    t["k_freqn"] = np.array([r.read(1)[0] for _ in range(15 * 29)], np.float32).reshape(15, 29)
    r = _Records(os.path.join(datadir, "karzas_xl.dat"), 3)
    t["k_xl"] = np.array([r.read(1)[0] for _ in range(6 * 6 * 29)], np.float32).reshape(6, 6, 29)
    rows = []
    for line in open(os.path.join(datadir, "gauntff_vanhoof.dat")):
        line = line.split("#")[0].strip()
        if line:
            rows.append(line)
    assert int(rows[0]) == 20140210 and rows[1].split()[:2] == ["81", "146"]
    g = np.array([[float(v) for v in rows[5 + j].split()] for j in range(146)])
    t["gff_logtab"] = np.array([[math.log10(v) for v in row] for row in g])  # (146 u, 81 gam2)
    r = _Records(os.path.join(datadir, "h2collop.dat"))
    line = r.line()
    while line.startswith("#"):
        line = r.line()
    npair, ntmax = (int(v) for v in line.split()[:2])
    ncianu, nu0, dnu = r.read(3)
    ncianu = int(ncianu)
    nciat = np.zeros(npair, np.int32)
    temps = np.zeros((npair, ntmax))
    tab = np.full((npair, ntmax, ncianu), -99.0)
    for ip in range(npair):
        nciat[ip] = int(r.line().split()[1])
        temps[ip, :nciat[ip]] = r.read(nciat[ip])
        for it in range(nciat[ip]):
            tab[ip, it] = r.read(ncianu)
    t.update(cia_nciat=nciat, cia_temp=temps, cia_tab=tab, cia_nu=np.array([nu0, dnu]))
    for name, nb in (("ch", 105), ("oh", 130)):
        r = _Records(os.path.join(datadir, f"cross{name}.dat"), 5)
        t[f"cross_{name}"] = np.array([r.read(15) for _ in range(nb)])  # (energy bin, T)
    return t


def kapp_waves():
    """Wavelengths (nm) of the per-source KAPP fixture: 240 log-spaced
    points from 8 nm to 300 um, and clusters across the edges and lines the
    sources switch on: the H I series limits n = 1..15, the He I and He II
    edges, the H- threshold, the bands of the H Rayleigh fit, the H I lines
    HLINOP covers near Ly-a, H-a/b/g, Pa-a/b and Br-a."""
    w = list(np.geomspace(8.0, 300000.0, 240))
    limits = [109678.764, 27419.659, 12186.462, 6854.871, 4387.113, 3046.604, 2238.320,
              1713.713, 1354.044, 1096.776, 906.426, 761.650, 648.980, 559.579, 487.456]
    edges = [1e7 / x for x in limits] + [22.79, 50.43, 260.0, 312.0, 342.2, 368.0, 1643.9]
    nu_l = 109677.583403 * 2.99792458e10
    for f in (0.01, 0.74, 0.77, 0.885, 0.890, 0.936, 0.938, 0.959, 0.961, 1.0):
        edges.append(2.99792458e17 / (f * nu_l))
    lines = [121.567, 102.572, 656.28, 486.13, 434.05, 1875.1, 1281.8, 4051.2, 2625.1]
    for e in edges:
        w += list(e * (1.0 + np.array([-1e-5, 1e-5])))
    for c in lines:
        w += list(c * (1.0 + np.array([-2e-3, -2e-4, 0.0, 1e-4, 1e-3])))
    return np.array(sorted(w))


KAPP_DECKS = {"plain": dict(), "hot": dict(t_top=6000.0, t_bot=16000.0),
              "cool": dict(t_top=2600.0, t_bot=6000.0)}


KAPP_DEPTHS = np.r_[0:72:3, 71]
KAPP_XNFP = np.r_[1:6, 6, 7, 10, 11, 15, 16, 21, 22, 28, 29, 36, 37, 45, 46, 55, 56, 66, 67, 78, 79,
                  91, 92, 105, 106, 136, 137, 171, 172, 210, 211, 351, 352, 846, 848]
KAPP_XNF = np.r_[1, 3, 4, 22:27, 29:34, 37:42, 56:61, 79:84, 106:111, 137:142, 352:357]


def golden_kapp():
    """KAPP (src/atlas12_modules.f90:11202-11324) and every source it calls,
    at 360 wavelengths, on three decks: the plain-SYNTHE atmosphere
    (4,400-9,500 K), a hot one (6,000-16,000 K: He I/II, H I n = 1) and a
    cool one (2,600-6,000 K: H2, CIA, H2-, CH/OH).  Writes the tables the
    sources hold (module tables from the reference's own memory after first
    use, SAVE tables read from the data files as the sources read them) to
    kapp_tables.npz, and per deck the depth inputs and per-source outputs to
    kapp_ref.npz."""
    waves = kapp_waves()
    out = {"waves": waves, "depths": KAPP_DEPTHS.astype(np.int32)}
    tables = None
    for name, kwargs in KAPP_DECKS.items():
        atm = wl.scurve_atmosphere(**kwargs)
        d = atlas_case(atm, 9599, 20, np.zeros((0, 4), np.int32), kapp_waves=waves)
        k = d["kapp"]
        if tables is None:
            tables = dict(k["tables"])
            for c, v in k["const"].items():
                tables["const_" + c] = np.float64(v)
        hi_ids = np.unique(k["tables"]["hi_edge"][:, 6].astype(int))
        xnfp_ids = np.unique(np.concatenate([KAPP_XNFP, hi_ids]))
        out["xnfp_ids"], out["xnf_ids"] = xnfp_ids.astype(np.int32), KAPP_XNF.astype(np.int32)
        for sk, sv in k["state"].items():
            if sk == "xnfp":
                sv = sv[xnfp_ids - 1]
            elif sk == "xnf":
                sv = sv[KAPP_XNF - 1]
            out[f"{name}_{sk}"] = sv
        for sk, sv in k["src"].items():
            out[f"{name}_{sk}"] = sv[:, KAPP_DEPTHS]
    tables.update(read_kapp_files("/root/reference/data"))
    np.savez_compressed(os.path.join(HERE, "kapp_tables.npz"), **tables)
    np.savez_compressed(os.path.join(HERE, "kapp_ref.npz"), **out)
    return {"kapp": {"waves": len(waves), "decks": list(KAPP_DECKS),
                     "tb_levels": int(len(tables["tb_g"])), "hi_edges": int(len(tables["hi_edge"]))}}


def gfall_line(wl_nm, loggf, code, e, xj, ep, xjp, gr=0.0, gs=0.0, gw=0.0, nblo=0, nbup=0,
               iso1=0, iso2=0, labelp="SYNTH   ", auto="   "):
    """One gfall record with the fields the complex-line routing reads
    (src/mod_mklinelist.f90:861-1060): NBLO/NBUP, ISO1/ISO2 (NCON), the
    CONTINUU label and the AUT/PRD flag in the last OTHER field."""
    other2 = f"{'':6s}{0:1d}{auto:3s}"
    return (f"{wl_nm:11.4f}{loggf:7.3f}{code:6.2f}{e:12.3f}{xj:5.1f} {'SYNTH':8s}{'':2s}"
            f"{ep:12.3f}{xjp:5.1f} {labelp:8s}{'':2s}{gr:6.2f}{gs:6.2f}{gw:6.2f}"
            f"SYN {nblo:2d}{nbup:2d}{iso1:3d}{0.0:6.3f}{iso2:3d}{0.0:6.3f}{'':10s}{other2:10s}"
            f"{0:5d}{0:5d}{0:6d}\n")


def _hseries(n, ms, ncon_from, ncon):
    """H I lines n -> m (gfall code 1.00) with NCON (ISO2) from m >= ncon_from."""
    ryd, elim = 109677.583403, 109678.764
    el = elim - ryd / n ** 2
    out = []
    for m in ms:
        eu = elim - ryd / m ** 2
        w = 1.0e7 / (eu - el)
        # Kramers-like f_nm, large enough that the wings reach the cutoff
        f = 0.64 / m ** 3 * (1.0 / n ** 2 - 1.0 / m ** 2) ** -3 / n ** 5
        out.append((w, gfall_line(w, float(np.log10(2 * n * n * f)), 1.00, el, 0.5, eu, 0.5,
                                  nblo=n, nbup=m, iso2=(ncon if m >= ncon_from else 0))))
    return out


def cplx_gfall(window="vis"):
    """Complex-profile lines of every type compute_line_opacity handles
    (src/synthe_module.f90:3167-3702), sorted by wavelength as gfall is.
    vis (365-660 nm): H I Balmer n = 2 -> 3..30 (Halpha alone, Hbeta with
    the alpha limit, the rest merging into the Balmer continuum, NCON = 2),
    D I alpha just before Halpha (the same transition: the reference's
    per-transition cache serves Halpha with D's convolution), He I lines
    without an edge, with NCON 1 and inside the tapered stretch below the
    NCON 5 edge, NLTE-flagged metals (Mg I, Ca II, Na I), a PRD line, an
    autoionising Ca I line and CONTINUU records: Mg I (NLAST 6), H I (NLAST
    8) and one of type 2 (coronal).  ir (1450-2400 nm): Brackett n = 4 ->
    6..30 and Pfund n = 5 -> 13..30 -- the Stehle tables stop at Brackett
    m = 7, so most take the Peterson/Kurucz path -- and a Si I line.  uv
    (118-180 nm): Lyman n = 1 -> 2..40 (Lyalpha with its H2 and H2+
    satellites, Lybeta with the alpha limit, m > 30 past the Stehle tables)
    and a C I line."""
    out = []
    if window == "vis":
        out += _hseries(2, range(3, 31), 4, 2)
        # D I alpha: ISO1 = 2 (itype -2); 2 dex weaker than H alpha
        ryd, elim = 109677.583403, 109678.764
        e2, e3 = elim - ryd / 4.0, elim - ryd / 9.0
        f = 0.64 / 27 * (1.0 / 4.0 - 1.0 / 9.0) ** -3 / 32.0
        out.append((656.2905, gfall_line(656.2905, float(np.log10(8 * f)) - 2.0, 1.00, e2, 0.5, e3, 0.5,
                                         nblo=2, nbup=3, iso1=2)))
        # He I (code 2.00): 447.1 (no edge), 492.2 (NCON 1), 376.0 (NCON 5:
        # the edge at 27175.76 cm^-1 tapers it between wcon and wtail)
        out.append((447.1480, gfall_line(447.1480, 0.052, 2.00, 169086.766, 5.0, 191444.482, 7.0,
                                         nblo=4, nbup=9)))
        out.append((492.1931, gfall_line(492.1931, -0.435, 2.00, 171134.897, 1.0, 191446.456, 2.0,
                                         nblo=5, nbup=10, iso2=1)))
        out.append((376.0000, gfall_line(376.0000, 1.5, 2.00, 169086.766, 1.0, 195682.0, 2.0,
                                         nblo=4, nbup=12, iso2=5)))
        # NLTE-flagged metals: Mg I b (518.36), Ca II H (396.96), Na I D2 (589.16)
        out.append((518.3604, gfall_line(518.3604, -0.167, 12.00, 21911.178, 1.0, 41197.403, 1.0,
                                         nblo=3, nbup=6)))
        out.append((396.9591, gfall_line(396.9591, -0.166, 20.01, 0.0, 0.5, 25191.51, 0.5,
                                         nblo=1, nbup=2)))
        out.append((589.1583, gfall_line(589.1583, 0.108, 11.00, 0.0, 0.5, 16973.366, 1.5,
                                         nblo=1, nbup=3, iso2=1)))
        # a PRD line (Fe II), an autoionising Ca I line, CONTINUU records
        out.append((455.0, gfall_line(455.0, -1.2, 26.01, 20000.0, 2.5, 41978.0, 3.5, auto="PRD")))
        out.append((520.5, gfall_line(520.5, -2.0, 20.00, 20000.0, 1.0, 39212.4, 2.0, gr=13.0, gs=1.5,
                                      gw=2.0, auto="AUT")))
        out.append((430.0, gfall_line(430.0, -3.0, 12.00, 35000.0, 0.5, 58255.8, 6.0,
                                      labelp="CONTINUU")))
        out.append((500.0, gfall_line(500.0, -2.5, 1.00, 97492.302, 0.5, 117492.302, 8.0,
                                      labelp="CONTINUU")))
        out.append((480.0, gfall_line(480.0, -3.0, 12.00, 35000.0, 0.5, 55833.333, 2.0,
                                      labelp="CONTINUU")))
    elif window == "ir":
        out += _hseries(4, range(6, 31), 6, 4)
        out += _hseries(5, range(13, 31), 7, 5)
        out.append((1588.8, gfall_line(1588.8, 0.2, 14.00, 47351.0, 2.0, 53645.0, 3.0, nblo=2, nbup=5)))
    elif window == "uv":
        out += _hseries(1, range(2, 41), 3, 1)
        out.append((165.7, gfall_line(165.7, -0.3, 6.00, 16.4, 1.0, 60369.0, 1.0, nblo=1, nbup=4)))
    else:
        raise ValueError(window)
    out.sort()
    return "".join(r for _, r in out)


def read_cplx_dump(path):
    """ref_synthe_dump's A12_CPLX_DUMP file."""
    b = open(path, "rb").read()
    off = 0

    def take(dt, n):
        nonlocal off
        a = np.frombuffer(b, dtype=dt, count=n, offset=off)
        off += a.nbytes
        return a.copy()
    n, length, n19, mw6 = (int(x) for x in take("<i4", 4))
    d = dict(nrhox=n, length=length, nlines=n19, mw6=mw6, cutoff=float(take("<f8", 1)[0]))
    recs = []
    for _ in range(n19):
        wlv = float(take("<f8", 1)[0])
        elo, gf = (float(x) for x in take("<f4", 2))
        nblo, nbup, nelion, itype, ncon, nelionx = (int(x) for x in take("<i4", 6))
        gr, gs, gw = (float(x) for x in take("<f4", 3))
        nbuff, lim = (int(x) for x in take("<i4", 2))
        recs.append((wlv, elo, gf, nblo, nbup, nelion, itype, ncon, nelionx, gr, gs, gw, nbuff, lim))
    d["lines"] = np.array(recs, dtype=CPLX_DTYPE)
    for k in ("t", "xne", "rho", "hckt", "xnfhe1", "xnfh2"):
        d[k] = take("<f8", n)
    per = {k: [] for k in ("txnxn", "xnfph1", "xnfph2", "xnfpel", "dopple", "xnfdop", "continuum", "buffer")}
    for _ in range(n):
        a = take("<f4", 1)
        per["txnxn"].append(float(a[0]))
        x = take("<f8", 2)
        per["xnfph1"].append(x[0])
        per["xnfph2"].append(x[1])
        per["xnfpel"].append(take("<f8", mw6))
        per["dopple"].append(take("<f8", mw6))
        per["xnfdop"].append(take("<f8", mw6))
        per["continuum"].append(take("<f4", length))
        per["buffer"].append(take("<f4", length))
    for k, v in per.items():
        d[k] = np.array(v, dtype=np.float32 if k in ("txnxn", "continuum", "buffer") else np.float64)
    assert off == len(b), (off, len(b))
    return d


CPLX_DTYPE = np.dtype([("wlvac", "<f8"), ("elo", "<f4"), ("gf", "<f4"), ("nblo", "<i4"),
                       ("nbup", "<i4"), ("nelion", "<i4"), ("itype", "<i4"), ("ncon", "<i4"),
                       ("nelionx", "<i4"), ("gammar", "<f4"), ("gammas", "<f4"), ("gammaw", "<f4"),
                       ("nbuff", "<i4"), ("lim", "<i4")])


CPLX_DEPTHS = np.r_[0:72:6, 71]
# window: (wlbeg, wlend, resolving power, depths kept, file suffix)
CPLX_WINDOWS = {"vis": (365.0, 660.0, 30000.0, CPLX_DEPTHS, ""),
                "ir": (1450.0, 2400.0, 30000.0, np.r_[0:72:12, 71], "_ir"),
                "uv": (118.0, 180.0, 30000.0, np.r_[0:72:12, 71], "_uv")}


def golden_cplx():
    """compute_line_opacity (src/synthe_module.f90:3167-3702) on the
    plain-SYNTHE atmosphere in three windows at R = 30,000, for the complex
    lines of cplx_gfall(window): the line records as run_mklinelist builds
    them, the per-depth inputs the driver forms (src/synthe.f90:476-551) and
    the CONTINUUM and BUFFER arrays after the call at the kept depths; the
    SYNTHE state of each window for the device path."""
    atm = wl.scurve_atmosphere()
    info = {}
    for name, (wlb, wle, res, js, suf) in CPLX_WINDOWS.items():
        tmp = tempfile.mkdtemp(prefix="a12c_")
        gpath = os.path.join(tmp, "gfall.dat")
        with open(gpath, "w") as fh:
            fh.write(cplx_gfall(name))
        root = data_root(tmp, gpath)
        deck = os.path.join(tmp, "cplx.dat")
        with open(deck, "w") as fh:
            fh.write(atm.deck())
        env = dict(os.environ, ATLAS12=root, A12_CPLX_DUMP=os.path.join(tmp, "cplx.bin"))
        dump = os.path.join(tmp, "dump.bin")
        run([os.path.join(REFBIN, "ref_synthe_dump"), deck, repr(wlb), repr(wle), repr(res), dump], tmp, env)
        d = trim_edges(read_dump(dump))
        c = read_cplx_dump(env["A12_CPLX_DUMP"])
        shutil.rmtree(tmp)
        nel = sorted(set(int(x) for x in c["lines"]["nelion"]) | {1, 2, 7, 8, 9})
        save_state(os.path.join(HERE, f"cplx{suf}_state.npz"), d, nel)
        keep = {k: c[k] for k in ("t", "xne", "rho", "hckt", "xnfhe1", "xnfh2")}
        for k in ("txnxn", "xnfph1", "xnfph2", "continuum", "buffer"):
            keep[k] = c[k][js]
        idx = np.array([i - 1 for i in nel], np.int32)
        for k in ("xnfpel", "dopple", "xnfdop"):
            keep[k] = c[k][js][:, idx]
        np.savez_compressed(os.path.join(HERE, f"cplx{suf}_ref.npz"), lines=c["lines"],
                            depths=np.asarray(js, np.int32), ion_idx=idx, length=c["length"],
                            cutoff=c["cutoff"], wlbeg=wlb, wlend=wle, resolu=res, **keep)
        types = {int(t): int(np.sum(c["lines"]["itype"] == t)) for t in np.unique(c["lines"]["itype"])}
        info[name] = {"nlines": int(c["nlines"]), "length": int(c["length"]), "itypes": types,
                      "nonzero_points": int(np.count_nonzero(c["buffer"]))}
    return {"cplx": info}


def golden_atlas_small():
    """One ATLAS12 opacity-sampling pass on the plain-SYNTHE atmosphere:
    240 frequencies from 500 nm (1e-4 dex steps), 3000 lines with the
    distributions of the ATLAS12 workload.  Captures LINOP1's XLINES, the
    KAPP outputs per frequency and RUN_RT_PASS's accumulators."""
    atm = wl.scurve_atmosphere()
    nustart = int(round((np.log10(500.0) - 1.0) / 1e-4)) + 1
    nuhi = 240
    w0 = 10 ** (1 + 1e-4 * nustart)
    w1 = 10 ** (1 + 1e-4 * (nuhi + nustart - 1))
    lines = wl.atlas_linedata(3000, w0 - 0.05, w1 - 0.2, seed=21)
    d = atlas_case(atm, nustart, nuhi, lines)
    ielion = np.ascontiguousarray(lines).view(np.int16).reshape(len(lines), 8)[:, 2]
    used = np.unique(np.abs(ielion.astype(int)) // 10)
    keep = {k: v for k, v in d.items() if not isinstance(v, (int, float))}
    keep["xnfdop"] = d["xnfdop"][used - 1]
    keep["dopple"] = d["dopple"][used - 1]
    keep["ion_idx"] = (used - 1).astype(np.int32)
    keep["linedata"] = lines
    for k in ("coefj", "coefh"):
        keep.pop(k)
    np.savez_compressed(os.path.join(HERE, "atlas_small.npz"), nustart=nustart,
                        teff=d["teff"], flux=d["flux"], mion=d["mion"],
                        acc_pradk0=d["acc_pradk0"], **keep)
    return {"atlas_small": {"nlines": len(lines), "numnu": d["numnu"],
                            "nwave1": d["nwave1"]}}


def golden_state_atlas():
    """ATLAS12 state of the one-iteration workload on the plain-SYNTHE
    atmosphere: the driver's 30,000-point grid for Teff < 7250 K (NUSTART =
    9599, 91.2 nm - 91.2 um), TABCONT, populations and Doppler widths of the
    atomic and diatomic species the line store uses, and KAPP's outputs at
    every 100th frequency (a full KAPP table is 104 MB; the benchmark fills
    the frequencies between by log-log interpolation)."""
    atm = wl.scurve_atmosphere()
    nustart, nuhi = 9599, 30000
    # a token selection run, for WAVETAB
    pred, tio = wl.atlas_iteration_lists(100, 400.0, 600.0, seed=1)
    d = atlas_case(atm, nustart, nuhi, None, raw=(pred, tio))
    n = d["nrhox"]
    # diatomic slots (841..1006) the equilibrium populates most at the
    # photosphere, for the molecular share of the line store
    xd = d["xnfdop"]
    mol = np.arange(840, 1006)
    score = xd[mol, n // 2]
    mol_used = mol[np.argsort(score)[::-1][:6]] + 1
    atoms = np.array(wl.ATLAS_SPECIES)
    used = np.unique(np.concatenate([atoms, mol_used, [wl.TIO_NELION]]))
    sub = np.arange(0, nuhi, 100)
    keep = dict(nustart=nustart, teff=d["teff"], flux=d["flux"], mion=d["mion"],
                waveset=d["waveset"], rcoset=d["rcoset"], tabcont=d["tabcont"],
                wavetab=d["wavetab"],
                iwavetab=d["iwavetab"], h0=d["h0"], h1=d["h1"], h2=d["h2"],
                ion_idx=(used - 1).astype(np.int32), xnfdop=xd[used - 1], dopple=d["dopple"][used - 1],
                mol_nelion=mol_used.astype(np.int32), kapp_sub=sub.astype(np.int32))
    for k in ("t", "hkt", "hckt", "xne", "rhox", "xnf1", "xnf3", "xnf841"):
        keep[k] = d[k]
    for k in ("acont", "sigmac", "scont", "sline", "aline", "sigmal"):
        keep[k + "_sub"] = d[k][sub]
    np.savez_compressed(os.path.join(HERE, "atlas_plain_state.npz"), **keep)
    return {"atlas_plain_state": {"numnu": nuhi, "molecules": mol_used.tolist()}}


def golden_select():
    """SELECTLINES on the plain-SYNTHE atmosphere's first-iteration state:
    the raw lists of the one-iteration workload's generator (atomic and
    diatomic predicted records, TiO MOL2 records) over 95 nm - 20 um, and
    the store the reference keeps.  The raw lists are regenerated from the
    seed by the tests (their sha256 is recorded to catch generator drift)."""
    atm = wl.scurve_atmosphere()
    st = np.load(os.path.join(HERE, "atlas_plain_state.npz"))
    mols = tuple(int(m) for m in st["mol_nelion"])
    # TiO gf raised so that the TiO packing path keeps lines at this Teff
    pred, tio = wl.atlas_iteration_lists(80_000, 95.0, 20_000.0, seed=31, molecules=mols,
                                         tio_loggf=(4.0, 1.5))
    d = atlas_case(atm, 9599, 10, None, raw=(pred, tio))
    used = np.unique(np.concatenate([np.array(wl.ATLAS_SPECIES), np.array(mols),
                                     [wl.TIO_NELION]]))
    digest = hashlib.sha256(pred.tobytes() + tio.tobytes()).hexdigest()
    np.savez_compressed(os.path.join(HERE, "atlas_select.npz"), selected=d["selected"],
                        wavetab=d["wavetab"], tabcont=d["tabcont"], iwavetab=d["iwavetab"],
                        hckt=d["hckt"], teff=d["teff"], mion=d["mion"],
                        ion_idx=(used - 1).astype(np.int32), xnfdop=d["xnfdop"][used - 1],
                        n_raw=np.int64(80_000), seed=np.int64(31), wl_range=np.array([95.0, 20_000.0]),
                        tio_loggf=np.array([4.0, 1.5]),
                        mol_nelion=np.array(mols, np.int32), raw_sha256=np.array(digest))
    return {"atlas_select": {"raw": [len(pred), len(tio)], "selected": len(d["selected"]),
                             "raw_sha256": digest}}


SELECT_ALL = dict(seed=51, wl_range=(95.0, 20_000.0), n_pred=20_000, n_lowobs=8_000,
                  n_hilines=8_000, n_h2o=8_000, n_h3p=4_000, n_dia=2_500)


def golden_select_all():
    """SELECTLINES over all seven of its lists on the plain-SYNTHE
    atmosphere's first-iteration state, and read_diatomics_for_atlas's
    records for the ASCII lists.  The raw lists are regenerated from
    SELECT_ALL by the tests (sha256 recorded to catch generator drift)."""
    atm = wl.scurve_atmosphere()
    st = np.load(os.path.join(HERE, "atlas_plain_state.npz"))
    mols = tuple(int(m) for m in st["mol_nelion"])
    pred, tio, lowobs, hilines, h2o, h3p, dia = wl.atlas_all_lists(mols, **SELECT_ALL)
    files = {"lowobsat12.bin": lowobs.tobytes(), "hilines.bin": hilines.tobytes(),
             "h2opokazatel.bin": h2o.tobytes(), "h3plus.dat": h3p.tobytes()}
    files.update({k: v.encode() for k, v in dia.items()})
    d = atlas_case(atm, 9599, 10, None, raw=(pred, tio), sel_files=files,
                   sel_mols=tuple(dia))
    digest = hashlib.sha256(b"".join([pred.tobytes(), tio.tobytes()] +
                                     [files[k] for k in sorted(files)])).hexdigest()
    np.savez_compressed(os.path.join(HERE, "atlas_select_all.npz"), selected=d["selected"],
                        diatomic=d["diatomic"], wavetab=d["wavetab"], tabcont=d["tabcont"],
                        iwavetab=d["iwavetab"], hckt=d["hckt"], teff=d["teff"], mion=d["mion"],
                        xnfdop=d["xnfdop"], mol_nelion=np.array(mols, np.int32),
                        raw_sha256=np.array(digest), config=np.array(json.dumps(SELECT_ALL)))
    return {"atlas_select_all": {"selected": len(d["selected"]),
                                 "diatomic_records": len(d["diatomic"]), "raw_sha256": digest}}


def scale_atmosphere():
    """The plain-SYNTHE structure with a rising Rosseland opacity, a
    radiative acceleration and a PRADK card, so SCALE_MODEL's INTEG, MAP1
    and radiation-pressure scaling all act on non-trivial data."""
    atm = wl.scurve_atmosphere()
    s = np.arange(atm.nrhox) / (atm.nrhox - 1)
    atm.abross = np.round(10.0 ** (-2.0 + 4.0 * s ** 1.3), 3 - 0)
    atm.accrad = 10.0 ** atm.logg * 1.0e-3 * (1.0 + s)
    atm.pradk = 0.35
    return atm


def golden_scale():
    """SCALE_MODEL (src/atlas12_modules.f90:6967-7060) after READIN on one
    deck, for a regrid-only call, an unchanged (Teff, log g) and three
    rescalings."""
    atm = scale_atmosphere()
    cases = [(0.0, 0.0), (5777.0, 4.44), (4500.0, 2.5), (7200.0, 4.8), (3500.0, 1.0)]
    tmp = tempfile.mkdtemp(prefix="a12s_")
    g = os.path.join(tmp, "gfall.dat")
    open(g, "w").close()
    root = data_root(tmp, g)
    deck = os.path.join(tmp, "model.dat")
    with open(deck, "w") as fh:
        fh.write(atm.deck())
    rec = {k: [] for k in ("params", "inp", "nout", "taustd", "tauros", "out", "scal")}
    for tn, gn in cases:
        out = os.path.join(tmp, "o.bin")
        run([os.path.join(REFBIN, "ref_scale_dump"), deck, repr(tn), repr(gn), out], tmp,
            dict(os.environ, ATLAS12=root))
        b = open(out, "rb").read()
        n, kw = (int(x) for x in np.frombuffer(b, "<i4", 2))
        o = 8
        par = np.frombuffer(b, "<f8", 9, o); o += 72
        inp = np.frombuffer(b, "<f8", 7 * n, o).reshape(7, n); o += 56 * n
        nout = int(np.frombuffer(b, "<i4", 1, o)[0]); o += 4
        tt = np.frombuffer(b, "<f8", 2 * kw, o).reshape(2, kw); o += 16 * kw
        res = np.frombuffer(b, "<f8", 9 * kw, o).reshape(9, kw); o += 72 * kw
        sc = np.frombuffer(b, "<f8", 6, o)
        rec["params"].append(par); rec["inp"].append(inp); rec["nout"].append(nout)
        rec["taustd"].append(tt[0]); rec["tauros"].append(tt[1]); rec["out"].append(res)
        rec["scal"].append(sc)
    shutil.rmtree(tmp)
    np.savez_compressed(os.path.join(HERE, "scale_model.npz"),
                        **{k: np.array(v) for k, v in rec.items()})
    return {"scale_model": {"cases": cases, "nrhox": n, "kw": kw}}


def read_molec_dump(path):
    b = open(path, "rb").read()
    n, nequa, nummol, nloc = (int(x) for x in np.frombuffer(b, "<i4", 4))
    o = 16
    ints = np.frombuffer(b, "<i4", nummol + 1 + nloc + nequa, o).astype(np.int32)
    o += 4 * len(ints)
    d = {"locj": ints[:nummol + 1], "kcomps": ints[nummol + 1:nummol + 1 + nloc],
         "idequa": ints[nummol + 1 + nloc:]}

    def f8(cnt, shape=None):
        nonlocal o
        a = np.frombuffer(b, "<f8", cnt, o).copy()
        o += 8 * cnt
        return a if shape is None else a.reshape(shape)
    d["xnmolcode"] = f8(nummol)
    d["p"], d["tk"], d["t"] = f8(n), f8(n), f8(n)
    d["xabund"] = f8(99 * n, (99, n))            # (element, depth)
    d["equilj"] = f8(nummol * n, (n, nummol))    # (depth, molecule)
    d["xnatom"], d["xne"] = f8(n), f8(n)
    d["xnmol"] = f8(nummol * n, (nummol, n))     # (molecule, depth)
    assert o == len(b)
    return d


def golden_molec():
    """NMOLEC's gas-phase Newton solve (src/atlas12_modules.f90:9078-9217)
    on two decks: the plain-SYNTHE atmosphere (4400-9500 K) and a cool dwarf
    (2800-6000 K, log g 5) where molecules hold most of C, N, O and H.
    Inputs as NMOLEC forms them and the reference's converged outputs."""
    cases = {"solar": wl.scurve_atmosphere(),
             "cool": wl.scurve_atmosphere(t_top=2800.0, t_bot=6000.0, logg=5.0, teff=3500.0)}
    tmp = tempfile.mkdtemp(prefix="a12m_")
    g = os.path.join(tmp, "gfall.dat")
    open(g, "w").close()
    root = data_root(tmp, g)
    rec, info = {}, {}
    for name, atm in cases.items():
        deck = os.path.join(tmp, f"{name}.dat")
        with open(deck, "w") as fh:
            fh.write(atm.deck())
        out = os.path.join(tmp, f"{name}.bin")
        run([os.path.join(REFBIN, "ref_molec_dump"), deck, out], tmp, dict(os.environ, ATLAS12=root))
        d = read_molec_dump(out)
        for k, v in d.items():
            if k in ("locj", "kcomps", "idequa", "xnmolcode"):
                if k in rec:
                    assert np.array_equal(rec[k], v), k  # one table for both decks
                rec[k] = v
            else:
                rec[f"{name}_{k}"] = v
        info[name] = {"nrhox": int(len(d["p"])), "t_range": [float(d["t"][0]), float(d["t"][-1])]}
    shutil.rmtree(tmp)
    np.savez_compressed(os.path.join(HERE, "molec_ref.npz"), **rec)
    info["nequa"] = int(len(rec["idequa"]))
    info["nummol"] = int(len(rec["xnmolcode"]))
    return {"molec": info}


def golden_records():
    """SYNTHE's record preparation for all three binary/ASCII readers on a
    5 nm window at R = 500,000: a gfall list, a predicted-format list over
    every old species index (mapped, unmapped and molecular), and a MOL2
    list over the isotope/molecule codes molec_dispatch knows plus unknown
    ones, both label flags and classical/tabulated radiative damping.  The
    reference's run_mklinelist concatenates them (gfall, predict, mol); the
    fixture keeps the raw lists, the mol_broad.dat rows, and the records."""
    tmp = tempfile.mkdtemp(prefix="a12r_")
    grid = wl.Grid.make(500.0, 505.0, 500000.0)
    g = wl.synthetic_gfall(600, 499.5, 505.5, seed=71)
    gpath = os.path.join(tmp, "gfall.dat")
    wl.write_gfall(g, gpath)
    root = data_root(tmp, gpath)
    d = os.path.join(root, "data")
    rng = np.random.default_rng(72)
    pred = wl.atlas_linedata(6000, 498.0, 507.0, seed=73, species=tuple(range(1, 1006)))
    with open(os.path.join(d, "pred.bin"), "wb") as fh:
        fh.write(np.ascontiguousarray(pred, "<i4").tobytes())
    isos = np.array([1, 2, 12, 13, 14, 15, 16, 17, 18, 23, 24, 25, 26, 28, 29, 30, 33, 39, 40,
                     41, 42, 43, 44, 46, 47, 48, 49, 50, 51, 52, 53, 54, 56, 57, 58, 3, 60, 99])
    codes = np.array([606, 608, 106, 607, 813, 814, 812, 820, 120, 124, 156, 822, 101, 108, 126])
    n = 5000
    mol = np.zeros(n, wl.MOL2_DTYPE)
    mol["wl"] = np.sort(np.exp(rng.uniform(np.log(498.0), np.log(507.0), n))).astype(np.float32)
    mol["loggf"] = np.round(rng.normal(-2.0, 1.5, n), 3)
    e = np.round(rng.exponential(8000.0, n), 3)
    mol["e1"] = np.where(rng.random(n) < 0.5, e, e + 2.0e4)
    mol["e2"] = np.where(mol["e1"] == e, e + 2.0e4, e)
    mol["code"] = codes[rng.integers(0, len(codes), n)]
    mol["iso"] = isos[rng.integers(0, len(isos), n)]
    mol["gr100"] = np.where(rng.random(n) < 0.3, 0, rng.integers(700, 900, n))
    mol["flag"] = (rng.random(n) < 0.4).astype(np.int32)
    wl.write_mol2(mol, os.path.join(d, "molx.bin"))
    with open(os.path.join(d, "lines.list"), "a") as fh:
        fh.write("predict pred.bin\nmol molx.asc\n")
    deck = os.path.join(tmp, "rec.dat")
    with open(deck, "w") as fh:
        fh.write(wl.scurve_atmosphere().deck())
    dump = os.path.join(tmp, "dump.bin")
    run([os.path.join(REFBIN, "ref_synthe_dump"), deck, "500.0", "505.0", "500000.0", dump],
        tmp, dict(os.environ, ATLAS12=root))
    lines = read_dump(dump)["lines"]
    broad = [ln.rstrip("\n") for ln in open(os.path.join(REF, "data", "mol_broad.dat"))]
    shutil.rmtree(tmp)
    np.savez_compressed(os.path.join(HERE, "records_ref.npz"), gfall=g, pred=pred,
                        mol2=mol.view(np.uint8), lines=lines.view(np.uint8),
                        mol_broad=np.array(broad), grid=np.array([500.0, 505.0, 500000.0]))
    return {"records": {"gfall_raw": len(g), "pred_raw": len(pred), "mol2_raw": n,
                        "lte_records": int(len(lines))}}


RECORDS_MOL = dict(seed=81, n_gfall=200, n_mol=1500, n_poly=1500, n_polybin=3000, n_h2o=4000,
                   window=(500.0, 505.0, 500000.0), span=(498.0, 507.0), teff=3400.0)


def golden_records_mol():
    """SYNTHE's remaining readers on a 5 nm window: Kurucz molecular ASCII
    lists (read_molec_ascii; FeH under code 156, AlO, isotopes molec_dispatch
    does not know), a polyatomic ASCII list (read_molec_poly, CaOH and an
    unlisted code), a polyatomic binary list (read_molec_poly_bin) and an H2O
    list (read_h2o), with a small gfall list first as run_mklinelist
    requires; TEFF 3400 K so that the polyatomic lists are read.  The raw
    lists are regenerated from RECORDS_MOL (stored as "config") by the
    tests; the fixture keeps the reference's records and mol_broad.dat."""
    c = RECORDS_MOL
    tmp = tempfile.mkdtemp(prefix="a12m_")
    g, mol, poly, pbin, h2o = wl.synthe_mol_lists(**{k: c[k] for k in wl.SYNTHE_MOL_KEYS})
    gpath = os.path.join(tmp, "gfall.dat")
    wl.write_gfall(g, gpath)
    root = data_root(tmp, gpath)
    d = os.path.join(root, "data")
    for name, text in mol.items():
        with open(os.path.join(d, name), "w") as fh:
            fh.write(text)
    with open(os.path.join(d, "caoh.asc"), "w") as fh:
        fh.write(poly)
    with open(os.path.join(d, "caohb.bin"), "wb") as fh:
        fh.write(pbin.astype("<i4").tobytes())
    with open(os.path.join(d, "h2o.bin"), "wb") as fh:
        fh.write(h2o.tobytes())
    with open(os.path.join(d, "lines.list"), "a") as fh:
        for name in mol:
            fh.write(f"mol {name}\n")
        fh.write("polymol caoh.asc\npolymol caohb.bin\nh2o h2o.bin\n")
    atm = wl.scurve_atmosphere(teff=c["teff"])
    deck = os.path.join(tmp, "recm.dat")
    with open(deck, "w") as fh:
        fh.write(atm.deck())
    dump = os.path.join(tmp, "dump.bin")
    w = c["window"]
    run([os.path.join(REFBIN, "ref_synthe_dump"), deck, str(w[0]), str(w[1]), str(w[2]), dump],
        tmp, dict(os.environ, ATLAS12=root))
    lines = read_dump(dump)["lines"]
    broad = [ln.rstrip("\n") for ln in open(os.path.join(REF, "data", "mol_broad.dat"))]
    shutil.rmtree(tmp)
    digest = hashlib.sha256(g.tobytes() + "".join(mol.values()).encode() + poly.encode() +
                            pbin.tobytes() + h2o.tobytes()).hexdigest()
    np.savez_compressed(os.path.join(HERE, "records_mol_ref.npz"), lines=lines.view(np.uint8),
                        mol_broad=np.array(broad), config=np.array(json.dumps(c)),
                        raw_sha256=np.array(digest))
    return {"records_mol": {"lte_records": int(len(lines)), "raw_sha256": digest}}


STEPS = {"synthe_small": golden_synthe_small, "atlas_small": golden_atlas_small,
         "state_atlas": golden_state_atlas, "select": golden_select, "select_all": golden_select_all, "scale": golden_scale,
         "tcorr": golden_tcorr, "tcorr_conv": golden_tcorr_conv,
         "state_plain": golden_state_plain,
         "numerics": golden_numerics, "josh": golden_josh, "molec": golden_molec,
         "kapp": golden_kapp, "cplx": golden_cplx, "records": golden_records,
         "records_mol": golden_records_mol}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", nargs="*")
    args = ap.parse_args()
    man_path = os.path.join(HERE, "MANIFEST.json")
    man = json.load(open(man_path)) if os.path.exists(man_path) else {}
    comp = subprocess.run(["amdflang", "--version"], capture_output=True, text=True).stdout.splitlines()[0]
    man["compiler"] = comp
    man["fflags"] = "-O3 -ffp-contract=off"
    for name, fn in STEPS.items():
        if args.only and name not in args.only:
            continue
        print("golden:", name, flush=True)
        man.setdefault("cases", {}).update(fn())
    files = {}
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            files[f] = hashlib.sha256(open(os.path.join(HERE, f), "rb").read()).hexdigest()
    man["sha256"] = files
    with open(man_path, "w") as fh:
        json.dump(man, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
