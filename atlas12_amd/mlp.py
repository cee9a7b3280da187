"""Learned starting structure: the training set, training and evaluation
of the MLP (4 -> 4 x 256 SiLU -> 144) that proposes T and log Pgas on the
first 72 points of the standard Rosseland grid TAUSTD = 10^(-6.875 +
0.125 i), the start that SCALE_MODEL (src/atlas12_modules.f90:6967-7060)
otherwise makes by rescaling one converged model.

The models come from the library's model iteration on the GPU
(a12_model_converge: the ATLAS12 loop with the device RT pass and TCORR(3)
and a reduced-physics KAPP/EOS).  The pipeline:

1. the solar model (5777 K, log g 4.44, [M/H] 0, 1.5 km/s) converged from
   the grey Eddington law;
2. each drawn (Teff, log g, [M/H], vturb) started from SCALE_MODEL of that
   solar model (the GPU restatement of the reference's routine) and
   iterated to the convergence criterion (max |ERROR| <= 1 %, max |DERIV|
   <= 10 %, max |T1| <= 1 K on the TCORR(3) table);
3. the network trained on the GPU (the library's BF16 MFMA kernels) on the
   converged log T and log Pgas of the training draw;
4. held-out models iterated from the network's T and compared, by the
   iteration at which each meets the criterion, with the same models
   started from SCALE_MODEL.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

TAU1LG, STEPLG, NDEPTH = -6.875, 0.125, 72
TAUSTD = 10.0 ** (TAU1LG + STEPLG * np.arange(NDEPTH))
RANGES = {"teff": (3500.0, 8000.0), "logg": (1.0, 5.0), "mh": (-2.0, 0.5), "vturb": (0.0, 3.0)}
SUN = (5777.0, 4.44, 0.0, 1.5)


def draw_inputs(n, seed):
    """(n, 4) Teff, log g, [M/H], vturb, uniform over RANGES."""
    rng = np.random.default_rng(seed)
    return np.stack([rng.uniform(*RANGES[k], n) for k in ("teff", "logg", "mh", "vturb")], 1)


def grey_t(teff):
    """The grey Eddington law T^4 = 3/4 Teff^4 (tau + 2/3) on the grid."""
    return teff * (0.75 * (TAUSTD + 2.0 / 3.0)) ** 0.25


@dataclass
class Converged:
    inputs: np.ndarray   # (n, 4)
    iters: np.ndarray    # (n,): iteration the criterion was met (0: not within itmax)
    t: np.ndarray        # (n, 72)
    st: dict             # structure arrays (n, 72): pg, ptot, rhox, abross, xne


def solar_model(ctx, itmax=100):
    """The solar model of the model iteration, from the grey law."""
    from . import native
    par = np.array([SUN])
    its, t, st = native.model_converge(ctx, par, grey_t(SUN[0])[None], itmax)
    if its[0] <= 0:
        raise RuntimeError(f"solar model did not converge ({its[0]})")
    return Converged(par, its, t, st)


def scaled_start(ctx, sun: Converged, inputs):
    """SCALE_MODEL's start for every (Teff, log g): the solar model regridded
    (INTEG of ABROSS, MAP1) and rescaled on the GPU.  Returns T (n, 72)."""
    from . import native
    n = len(inputs)
    te, lg, _, vt = SUN
    fields = np.zeros((n, 7, NDEPTH))
    fields[:, 0] = sun.st["rhox"][0]
    fields[:, 1] = sun.t[0]
    fields[:, 2] = sun.st["pg"][0]
    fields[:, 3] = sun.st["xne"][0]
    fields[:, 4] = sun.st["abross"][0]
    fields[:, 5] = 0.0                  # PRAD (not in the reduced physics)
    fields[:, 6] = vt * 1.0e5           # VTURB
    sc = np.zeros((n, 7))
    sc[:, 0], sc[:, 1] = te, 10.0 ** lg
    sc[:, 5], sc[:, 6] = inputs[:, 0], inputs[:, 1]
    out, _, _ = native.scale_model(ctx, sc, fields, kw=NDEPTH, tau1lg=TAU1LG, steplg=STEPLG)
    return out[:, 1, :]


def converge(ctx, inputs, t_start, itmax=150):
    from . import native
    its, t, st = native.model_converge(ctx, inputs, t_start, itmax)
    return Converged(np.asarray(inputs, float), its, t, st)


def targets(c: Converged):
    """(n, 144): log10 T, then log10 Pgas, of converged models."""
    return np.concatenate([np.log10(c.t), np.log10(c.st["pg"])], 1)


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


def train(ctx, inputs, tgt, seed=0, schedule=((2000, 2e-3), (2000, 5e-4), (2000, 1e-4))):
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
    """(n, 144) log10 T and log10 Pgas proposed for inputs."""
    from . import native
    native.mlp_set_params(ctx, model.params)
    return model.ys.inv(native.mlp_infer(ctx, model.xs.fwd(inputs)).astype(np.float64))


def structure_errors(pred, ref):
    """Per model: max over depth of |dT|/T and of |d log10 Pgas|."""
    dt = np.max(np.abs(10.0 ** (pred[:, :NDEPTH] - ref[:, :NDEPTH]) - 1.0), 1)
    dp = np.max(np.abs(pred[:, NDEPTH:] - ref[:, NDEPTH:]), 1)
    return dt, dp


def pipeline(ctx, n_train=2000, n_test=200, seed=10, itmax=150):
    """Steps 1-4 of the module docstring.  Returns a dict with the training
    and held-out sets, the network, and per held-out model the iterations
    from SCALE_MODEL's start and from the network's start."""
    sun = solar_model(ctx)
    x_tr = draw_inputs(n_train, seed)
    x_te = draw_inputs(n_test, seed + 1)
    tr = converge(ctx, x_tr, scaled_start(ctx, sun, x_tr), itmax)
    te = converge(ctx, x_te, scaled_start(ctx, sun, x_te), itmax)
    ok = tr.iters > 0
    model = train(ctx, x_tr[ok], targets(tr)[ok])
    pred = predict(ctx, model, x_te)
    ml = converge(ctx, x_te, 10.0 ** pred[:, :NDEPTH], itmax)
    return dict(sun=sun, train=tr, test=te, model=model, pred=pred, mlp_run=ml)
