"""CPU restatement of the library's model iteration (a12_model_converge).

TEST INFRASTRUCTURE ONLY: the checker of the GPU model iteration, imported by
tests/ and bench.py's CPU leg.  The product library never imports it.

It restates, in numpy, the closed-form EOS, continuum opacity and
hydrostatic integration the device kernels use (the same formulas, loop
counts and update rules; the FP64 sums run in another order and the
transcendental functions are numpy's, so agreement is to rounding), and
runs the ATLAS12 pieces through the oracle's C restatement of the reference
(or_rt_pass: RUN_RT_PASS, src/atlas12_modules.f90:3880-3951; or_integ:
INTEG :5293-5434; or_tcorr3: TCORR(3) :2528-2752).

The closed forms are this library's own reduced physics, not the
reference's KAPP/POPS: nothing in the reference pins them ("parity
unpinned" against the reference; tests/test_oracle_model.py pins them
against pure-hydrogen Saha and the analytic limits instead).
"""
from __future__ import annotations

import numpy as np

from . import oracle as orc

N = 72
NF = 1200
KB, MH, SIG, SIGT = 1.380649e-16, 1.6735575e-24, 5.670374419e-5, 6.6524587e-25
HP, CL, CNMS = 6.62607015e-27, 2.99792458e10, 2.99792458e17
CHI_H, CHI_M, SAHA, CHI_HM, HE = 157807.7, 88194.0, 2.4147e15, 8753.3, 0.085
TAUSTD = 10.0 ** (-6.875 + 0.125 * np.arange(N))


def freq_table():
    """(wave nm, freq Hz, RCOSET) of the model iteration's grid: the ATLAS12
    driver's WAVESET at every 25th point from NUSTART 9599
    (src/atlas12c.f90:470-490)."""
    nu = 1 + 25 * np.arange(NF)
    w = 10.0 ** (1.0 + 0.0001 * (nu + 9599 - 1))
    f = CNMS / w
    r = np.zeros(NF)
    r[0] = (f[0] - f[1]) * 1.5
    r[1:-1] = (f[:-2] - f[2:]) * 0.5
    r[-1] = (f[-2] + f[-1]) * 0.25
    return w, f, r


WAVE, FREQ, RCO = freq_table()


def _per_freq():
    lam = WAVE * 10.0
    hm = [1.99654, -1.18267e-5, 2.64243e-6, -4.40524e-10, 3.23992e-14, -1.39568e-18, 2.78701e-23]
    lc = np.clip(lam, 2250.0, 16000.0)
    a = sum(c * lc ** i for i, c in enumerate(hm))
    hmbf = np.where(lam < 16419.0, np.maximum(a, 0.0) * 1e-18, 0.0)
    ll = np.log10(np.clip(lam, 2600.0, 113900.0))
    f0 = -2.2763 - 1.6850 * ll + 0.76661 * ll**2 - 0.053346 * ll**3
    f1 = 15.2827 - 9.2846 * ll + 1.99381 * ll**2 - 0.142631 * ll**3
    f2 = -197.789 + 190.266 * ll - 67.9775 * ll**2 + 10.6913 * ll**3 - 0.625151 * ll**4
    lr = np.maximum(lam, 1300.0)
    ray = 5.799e-13 / lr**4 + 1.422e-6 / lr**6 + 2.784 / lr**8
    nmin = np.full(NF, 9)
    for lv in range(8, 0, -1):
        nmin = np.where(FREQ >= 3.2898e15 / lv**2, lv, nmin)
    mbf = np.where(FREQ >= 1.839e15, 1e-18 * (1.839e15 / FREQ) ** 2, 0.0)
    return dict(hk=HP * FREQ / KB, hmbf=hmbf, f0=f0, f1=f1, f2=f2, ray=ray, nmin=nmin,
                hbf=2.815e29 / FREQ**3, hff=3.69e8 / FREQ**3, mbf=mbf, f153=(FREQ / 1e15) ** 3)


PF = _per_freq()


def eos(T, pg, am):
    """Saha and charge balance of H, He and the metal donor at (T, P_gas):
    dict of n_e, n(H I), n(H II), n(H-), n(donor I), rho (arrays)."""
    T = np.asarray(T, float)
    ntot = pg / (KB * T)
    t15 = T * np.sqrt(T)
    sh = SAHA * t15 * np.exp(-CHI_H / T)
    sm = SAHA * t15 * np.exp(-CHI_M / T)
    lo = np.full_like(ntot, 1e-30)
    hi = np.array(ntot, float)
    for _ in range(64):
        ne = np.sqrt(lo * hi)
        nh = (ntot - ne) / (1.0 + HE + am)
        f = ne - nh * (sh / (ne + sh) + am * sm / (ne + sm))
        hi = np.where(f > 0, ne, hi)
        lo = np.where(f > 0, lo, ne)
    ne = np.sqrt(lo * hi)
    nh = (ntot - ne) / (1.0 + HE + am)
    fh = sh / (ne + sh)
    nh0 = nh * (1 - fh)
    return dict(ne=ne, nh0=nh0, nhp=nh * fh, rho=nh * MH * (1 + 4 * HE + 20 * am),
                nhm=nh0 * ne / (4 * SAHA * t15) * np.exp(CHI_HM / T), nm0=am * nh * ne / (ne + sm))


def _levels(T, nh0):
    """s[n-1] = sum over n' >= n (n' <= 8) of n(H I, n') / n'^5, s[8] = 0."""
    T = np.asarray(T, float)
    s = np.zeros((9,) + T.shape)
    for lv in range(8, 0, -1):
        n2 = float(lv * lv)
        s[lv - 1] = s[lv] + nh0 * n2 * np.exp(-CHI_H * (1 - 1 / n2) / T) / (n2 * n2 * lv)
    return s


def kappa(T, p):
    """(acont, sigmac, stim, bnu) per gram, each (NF, nlayers), for layer
    temperatures T and populations p (from eos)."""
    T = np.atleast_1d(np.asarray(T, float))
    ehv = np.exp(-PF["hk"][:, None] / T[None])
    st = 1 - ehv
    lth = np.log10(5040.0 / T)[None]
    pe = p["ne"] * KB * T
    s = _levels(T, p["nh0"])
    hmbf = PF["hmbf"][:, None] * p["nhm"][None]
    hmff = 1e-26 * pe[None] * 10.0 ** (PF["f0"][:, None] + PF["f1"][:, None] * lth
                                      + PF["f2"][:, None] * lth * lth) * p["nh0"][None]
    hbf = PF["hbf"][:, None] * s[PF["nmin"] - 1]
    hff = PF["hff"][:, None] * (p["ne"] * p["nhp"] / np.sqrt(T))[None]
    mbf = PF["mbf"][:, None] * p["nm0"][None]
    rho = p["rho"][None]
    acont = ((hmbf + hbf + hff + mbf) * st + hmff) / rho
    sigmac = (SIGT * p["ne"][None] + PF["ray"][:, None] * p["nh0"][None]) / rho
    bnu = 2.0 * HP / (CL * CL) * 1e45 * PF["f153"][:, None] * ehv / st
    return acont, sigmac, st, bnu


def kross(T, ptot, am, vt2, r):
    """(kappa_Ross, new r = rho / P_gas, populations) at one layer."""
    pg = ptot / (1 + 0.5 * vt2 * r)
    p = eos(np.array([T]), np.array([pg]), am)
    ac, sg, st, b = kappa(np.array([T]), p)
    dbdt = b[:, 0] * FREQ * (HP / KB / T) / T / st[:, 0]
    s = np.sum(dbdt / (ac[:, 0] + sg[:, 0]) * RCO)
    return 4 * SIG / np.pi * T**3 / s, p["rho"][0] / pg, {k: v[0] for k, v in p.items()}


def hydro(T, g, am, vt2):
    """P_tot, P_gas and the populations on the grid for T (the device's
    fixed-point and trapezoid passes, in the same order; the populations are
    those of each layer's last kappa_Ross evaluation)."""
    ptot, pgas = np.zeros(N), np.zeros(N)
    pops = {k: np.zeros(N) for k in ("ne", "nh0", "nhp", "rho", "nhm", "nm0")}
    r = 0.0
    p = g * TAUSTD[0] / 1e-2
    for _ in range(30):
        k, r, _ = kross(T[0], p, am, vt2, r)
        p = g * TAUSTD[0] / k
    _, r, pp = kross(T[0], p, am, vt2, r)
    ptot[0], pgas[0] = p, p / (1 + 0.5 * vt2 * r)
    for k in pops:
        pops[k][0] = pp[k]
    for j in range(1, N):
        k0, _, _ = kross(T[j - 1], ptot[j - 1], am, vt2, r)
        dl = np.log(TAUSTD[j] / TAUSTD[j - 1])
        f0 = g * TAUSTD[j - 1] / k0
        pj = ptot[j - 1] + f0 * dl
        for _ in range(4):
            k1, r, _ = kross(T[j], pj, am, vt2, r)
            pj = ptot[j - 1] + 0.5 * dl * (f0 + g * TAUSTD[j] / k1)
        _, r, pp = kross(T[j], pj, am, vt2, r)
        ptot[j], pgas[j] = pj, pj / (1 + 0.5 * vt2 * r)
        for k in pops:
            pops[k][j] = pp[k]
    return ptot, pgas, pops


def converge(par, t_start, itmax, coefj, coefh):
    """One model: par (teff, logg, mh, vturb km/s), starting T (72,).
    Returns (iters as a12_model_converge reports it, T, structure dict pg,
    ptot, rhox, abross, xne of the last iteration run, hist (k, 3))."""
    teff, logg, mh, vt = (float(x) for x in par)
    g, am, vt2 = 10.0 ** logg, 1e-4 * 10.0 ** mh, (vt * 1e5) ** 2
    flux = SIG / (4 * np.pi) * teff**4
    T = np.array(t_start, float)
    oldt1 = np.zeros(N)
    hist = []
    st = None
    for it in range(1, itmax + 1):
        ptot, pg, p = hydro(T, g, am, vt2)
        rhox = ptot / g
        rho = p["rho"]
        ac, sg, st, b = kappa(T, p)
        z = dict(waveset=WAVE, rcoset=RCO, t=T, hkt=HP / (KB * T), rhox=rhox, teff=teff,
                 acont=ac, sigmac=sg, scont=b, sline=b, aline=np.zeros((NF, N)),
                 sigmal=np.zeros((NF, N)))
        acc, _ = orc.rt_pass(z, np.zeros((NF, N)), coefj, coefh)
        abross = 4 * SIG / np.pi * T**3 / acc["abross"]
        st = dict(pg=pg, ptot=ptot, rhox=rhox, abross=abross, xne=p["ne"])
        tauros = orc.integ(rhox, abross, abross[0] * rhox[0])
        tz = dict(n=N, ifconv=0, iter=it, flux=flux, teff=teff, mixlth=1.25, tauros=tauros,
                  rhox=rhox, t=T, abross=abross, ptotal=ptot, rho=rho, flxrad=acc["flxrad"],
                  rjmins=acc["rjmins"], rdiagj=acc["rdiagj"])
        for k in ("dltdlp", "flxcnv", "flxcnv0", "grdadb", "dlrdlt", "heatcp", "hscale"):
            tz[k] = np.zeros(N)
        out, _, oldt1 = orc.tcorr3(tz, acc["rdabh"], oldt1=oldt1)
        e = np.max(np.abs(out["flxerr"]))
        d = np.max(np.abs(out["flxdrv"]))
        dt = np.max(np.abs(out["t1"]))
        hist.append((e, d, dt))
        if not (np.all(np.isfinite(out["t1"])) and np.all(np.isfinite(out["flxerr"]))):
            return -it, T, st, np.array(hist)
        if e <= 1.0 and d <= 10.0 and dt <= 1.0:
            return it, T, st, np.array(hist)
        T = np.maximum(T + out["t1"], 1500.0)
    return 0, T, st, np.array(hist)
