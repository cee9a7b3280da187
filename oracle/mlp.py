"""CPU oracle of the starting-structure network (test infrastructure, the
checker only; the network has no counterpart in the reference).

The same 4 -> 256 -> 256 -> 256 -> 256 -> 144 network (SiLU hidden layers)
in numpy, forward and backward, with the library's arithmetic made
explicit: every GEMM operand rounded to BF16 (round to nearest even), sums
in float64, activations and gradients kept in float32.  With bf16=False it
is the plain float64 network that the finite-difference test pins.

Parameter order (as the C-ABI's a12_mlp_*_params): W1 (4 x 256, in x out),
b1, W2 (256 x 256), b2, W3, b3, W4, b4, W5 (256 x 144), b5.
"""
from __future__ import annotations

import numpy as np

SIZES = [(4, 256), (256, 256), (256, 256), (256, 256), (256, 144)]
N_PARAMS = sum(i * o + o for i, o in SIZES)  # 235,664


def bf16(a):
    """Round float32 values to BF16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


def unpack(params):
    p = np.asarray(params)
    out, o = [], 0
    for i, n in SIZES:
        w = p[o:o + i * n].reshape(i, n)
        o += i * n
        b = p[o:o + n]
        o += n
        out.append((w, b))
    assert o == len(p) == N_PARAMS
    return out


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, x, bf=True):
    """Returns (y, cache)."""
    f = np.float32 if bf else np.float64
    rnd = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    layers = unpack(params)
    h = rnd(np.asarray(x, f))
    hs, zs = [h], []
    for k, (w, b) in enumerate(layers):
        z = (rnd(h).astype(np.float64) @ rnd(np.asarray(w, f)).astype(np.float64) + b).astype(f)
        zs.append(z)
        if k < len(layers) - 1:
            h = rnd(_silu(z).astype(f))
            hs.append(h)
    return zs[-1], (hs, zs)


def loss_and_grads(params, x, t, bf=True):
    """Mean-squared error over the n x 144 outputs and its gradient in the
    parameter order."""
    f = np.float32 if bf else np.float64
    rnd = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    layers = unpack(params)
    y, (hs, zs) = forward(params, x, bf)
    n = len(x)
    d = (y - np.asarray(t, f)).astype(f)
    loss = float(np.sum(d.astype(np.float64) ** 2) / (n * 144))
    dz = (2.0 * d / (n * 144.0)).astype(f)
    grads = [None] * len(layers)
    for k in range(len(layers) - 1, -1, -1):
        w, _ = layers[k]
        dzb = rnd(dz).astype(np.float64)
        gw = (rnd(hs[k]).astype(np.float64).T @ dzb).astype(f)
        gb = np.sum(dz.astype(np.float64), axis=0).astype(f)
        grads[k] = (gw, gb)
        if k > 0:
            dh = dzb @ rnd(np.asarray(w, f)).astype(np.float64).T
            dz = (dh * _dsilu(zs[k - 1].astype(np.float64))).astype(f)
    flat = np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in grads])
    return loss, flat


def glorot(seed=0):
    rng = np.random.default_rng(seed)
    parts = []
    for i, n in SIZES:
        lim = np.sqrt(6.0 / (i + n))
        parts += [rng.uniform(-lim, lim, i * n), np.zeros(n)]
    return np.concatenate(parts).astype(np.float32)
