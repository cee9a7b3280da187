"""Learned starting structure: training set, training and evaluation of the
MLP (4 -> 4 x 256 SiLU -> 144) that proposes T and log Pgas on the first 72
points of the standard Rosseland grid TAUSTD (10^(-6.875 + 0.125 i)), the
start that SCALE_MODEL (src/atlas12_modules.f90:6967-7060) otherwise makes
by rescaling one model.

The models it learns come from a reduced-physics atmosphere (no
counterpart in the reference): a grey temperature law with the Hopf
function, T^4 = 3/4 Teff^4 (tau + q(tau)), and hydrostatic equilibrium
dP_tot/dtau = g / kappa(P, T) with an H^- -like Rosseland opacity that
grows with metallicity and temperature, an electron-scattering floor and
a turbulent pressure rho vturb^2 / 2.  It depends on all four inputs (Teff,
log g, [M/H], vturb) where a rescaled model carries only Teff and g.

Training and inference run on the GPU through the library's HIP kernels
(atlas12_amd.native.mlp_*); the scaled start through its SCALE_MODEL.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

TAU1LG, STEPLG, KW, NDEPTH = -6.875, 0.125, 80, 72
TAUSTD = 10.0 ** (TAU1LG + STEPLG * np.arange(KW))
K_BOLTZ, M_H, MU = 1.380649e-16, 1.6735575e-24, 1.3
RANGES = {"teff": (3500.0, 8000.0), "logg": (1.0, 5.0), "mh": (-2.0, 0.5), "vturb": (0.0, 3.0)}
SUN = (5777.0, 4.44, 0.0, 1.5)


def kappa(p, t, mh):
    """Rosseland opacity [cm^2/g] of the reduced-physics atmosphere."""
    return (10.0 ** mh) * 2.0e-2 * np.sqrt(p / 1.0e4) * (t / 5000.0) ** 9 + 2.0e-4 + 1.0e-3 * 10.0 ** mh


def grey_structure(teff, logg, mh, vturb_kms, tau=TAUSTD, sub=4):
    """T, Pgas, P_tot and kappa on the optical-depth grid tau."""
    g = 10.0 ** logg
    vt2 = (vturb_kms * 1.0e5) ** 2
    t = (0.75 * teff ** 4 * (tau + 0.7104 - 0.1331 * np.exp(-3.4488 * tau))) ** 0.25
    fturb = MU * M_H * vt2 / (2.0 * K_BOLTZ * t)  # P_turb / P_gas

    def rhs(lnpt, lntau, tt, ft):
        pt = math.exp(lnpt)
        return math.exp(lntau) * g / (kappa(pt / (1.0 + ft), tt, mh) * pt)

    # P_tot at the first depth: P_tot = g tau / kappa(P_gas), by fixed point in log space
    lnpt = math.log(g * tau[0] / kappa(1.0, t[0], mh))
    for _ in range(80):
        new = math.log(g * tau[0] / kappa(math.exp(lnpt) / (1.0 + fturb[0]), t[0], mh))
        lnpt = 0.5 * (lnpt + new)
    lnt = np.log(tau)
    pts = [lnpt]
    for i in range(len(tau) - 1):
        h = (lnt[i + 1] - lnt[i]) / sub
        for s in range(sub):
            x0 = lnt[i] + s * h
            f = lambda x, y: rhs(y, x, np.interp(x, lnt, t), np.interp(x, lnt, fturb))  # noqa: E731
            k1 = f(x0, lnpt)
            k2 = f(x0 + h / 2, lnpt + h * k1 / 2)
            k3 = f(x0 + h / 2, lnpt + h * k2 / 2)
            k4 = f(x0 + h, lnpt + h * k3)
            lnpt += h * (k1 + 2 * k2 + 2 * k3 + k4) / 6
        pts.append(lnpt)
    ptot = np.exp(np.array(pts))
    pgas = ptot / (1.0 + fturb)
    return t, pgas, ptot, kappa(pgas, t, mh)


def draw_inputs(n, seed):
    rng = np.random.default_rng(seed)
    return np.stack([rng.uniform(*RANGES[k], n) for k in ("teff", "logg", "mh", "vturb")], 1)


def targets(inputs):
    """(n, 144): log10 T, then log10 Pgas, on the first 72 TAUSTD points."""
    out = np.zeros((len(inputs), 2 * NDEPTH))
    for i, (te, lg, mh, vt) in enumerate(inputs):
        t, pg, _, _ = grey_structure(te, lg, mh, vt, TAUSTD[:NDEPTH])
        out[i, :NDEPTH] = np.log10(t)
        out[i, NDEPTH:] = np.log10(pg)
    return out


@dataclass
class Scaler:
    mean: np.ndarray
    std: np.ndarray

    @classmethod
    def fit(cls, a):
        return cls(a.mean(0), np.where(a.std(0) > 0, a.std(0), 1.0))

    def fwd(self, a):
        return (a - self.mean) / self.std

    def inv(self, a):
        return a * self.std + self.mean


@dataclass
class Surrogate:
    params: np.ndarray
    xs: Scaler
    ys: Scaler
    losses: list


def train(ctx, inputs, tgt, seed=0, schedule=((2000, 2e-3), (1500, 5e-4), (1500, 1e-4))):
    """Full-batch Adam on the GPU; returns the trained Surrogate."""
    from . import native
    xs, ys = Scaler.fit(inputs), Scaler.fit(tgt)
    native.mlp_init(ctx, len(inputs), seed)
    native.mlp_set_data(ctx, xs.fwd(inputs), ys.fwd(tgt))
    losses = [native.mlp_grads(ctx)[0]]
    for steps, lr in schedule:
        losses.append(native.mlp_train(ctx, steps, lr))
    return Surrogate(native.mlp_get_params(ctx), xs, ys, losses)


def predict(ctx, model: Surrogate, inputs):
    from . import native
    native.mlp_set_params(ctx, model.params)
    return model.ys.inv(native.mlp_infer(ctx, model.xs.fwd(inputs)).astype(np.float64))


def scaled_start(ctx, inputs):
    """SCALE_MODEL's start: the solar model of the same reduced physics,
    regridded and rescaled to each (Teff, log g) on the GPU.  (n, 144) as
    targets()."""
    from . import native
    te, lg, mh, vt = SUN
    t, pg, pt, kap = grey_structure(te, lg, mh, vt, TAUSTD)
    g = 10.0 ** lg
    n = len(inputs)
    fields = np.zeros((n, 7, KW))
    fields[:, 0] = pt / g                 # RHOX
    fields[:, 1] = t                      # T
    fields[:, 2] = pg                     # P
    fields[:, 3] = 1.0e10                 # XNE (not used by the start)
    fields[:, 4] = kap                    # ABROSS
    fields[:, 5] = 0.0                    # PRAD
    fields[:, 6] = vt * 1.0e5             # VTURB
    sc = np.zeros((n, 7))
    sc[:, 0], sc[:, 1] = te, g
    sc[:, 5], sc[:, 6] = inputs[:, 0], inputs[:, 1]
    out, _, _ = native.scale_model(ctx, sc, fields, kw=KW, tau1lg=TAU1LG, steplg=STEPLG)
    res = np.zeros((n, 2 * NDEPTH))
    res[:, :NDEPTH] = np.log10(out[:, 1, :NDEPTH])
    res[:, NDEPTH:] = np.log10(out[:, 2, :NDEPTH])
    return res


def structure_errors(pred, ref):
    """Per model: max over depth of |dT|/T and of |d log10 Pgas|."""
    dt = np.max(np.abs(10.0 ** (pred[:, :NDEPTH] - ref[:, :NDEPTH]) - 1.0), 1)
    dp = np.max(np.abs(pred[:, NDEPTH:] - ref[:, NDEPTH:]), 1)
    return dt, dp
