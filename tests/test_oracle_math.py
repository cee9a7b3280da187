"""Pins of the oracle that do not go through the reference: closed forms,
invariants and library routines chosen so that a dropped term, a wrong sign
or a transposed operand fails one of them.  Mistakes are planted in copies
of the oracle's formulas here (on the CPU) to show the checks bite."""
import math

import numpy as np
import pytest
from scipy.special import erfcx, expn, wofz


# ---------------------------------------------------------------- Voigt
def test_voigt_is_real_part_of_faddeeva(orc):
    rng = np.random.default_rng(2)
    v = rng.uniform(0, 14, 4000)
    a = 10 ** rng.uniform(-4, 0.8, 4000)
    keep = v + a <= 15
    h = orc.voigt_f64(v[keep], a[keep])
    w = wofz(v[keep] + 1j * a[keep]).real
    # the method is accurate to ~2e-8 here (the reference states 1e-5)
    assert np.max(np.abs(h - w) / w) <= 1e-7


def test_voigt_centre_is_erfcx(orc):
    a = 10 ** np.linspace(-5, 1, 60)
    h = orc.voigt_f64(np.zeros_like(a), a)
    assert np.max(np.abs(h - erfcx(a)) / erfcx(a)) <= 1e-12


def test_voigt_normalisation(orc):
    """int H dv = sqrt(pi) (Humlicek asymptote beyond |v|+a = 15 included)."""
    for a in (0.01, 0.1, 1.0):
        v = np.linspace(0, 400, 400001)
        h = orc.voigt_f64(v, np.full_like(v, a))
        tail = 2 * a / math.sqrt(math.pi) / v[-1]  # Lorentz tail beyond 400
        integral = 2 * np.trapezoid(h, v) + tail
        assert abs(integral - math.sqrt(math.pi)) / math.sqrt(math.pi) <= 2e-5


def test_weideman_coefficients_regenerated():
    """The 32 Weideman coefficients follow from his FFT construction
    (Weideman 1994, N = 32, L = sqrt(N / sqrt 2)); a transposed or
    mis-ordered table fails this."""
    import re
    src = open(__import__("os").path.join(
        __import__("os").path.dirname(__file__), "..", "oracle", "a12_oracle.c")).read()
    body = src[src.index("WEI_A[32]"):]
    body = body[body.index("{") + 1:body.index("}")]
    tab = np.array([float(t) for t in re.findall(r"[-+0-9.e]+", body)])
    N = 32
    M = 2 * N
    M2 = 2 * M
    k = np.arange(-M + 1, M)
    L = math.sqrt(N / math.sqrt(2))
    t = L * np.tan(k * math.pi / M / 2)
    f = np.concatenate([[0.0], np.exp(-t * t) * (L * L + t * t)])
    a = np.real(np.fft.fft(np.fft.fftshift(f))) / M2
    a = a[1:N + 1][::-1]
    assert np.max(np.abs(a - tab)) <= 1e-15 * 10


# ---------------------------------------------------------------- numerics
def test_solvit_equals_lapack(orc):
    rng = np.random.default_rng(4)
    for _ in range(10):
        A = rng.normal(size=(51, 51)) + 5 * np.eye(51)
        b = rng.normal(size=51)
        assert np.allclose(orc.solvit(A, b), np.linalg.solve(A, b), rtol=1e-11, atol=1e-13)


def test_integ_exact_on_linear_and_quadratic(orc):
    x = np.sort(np.random.default_rng(1).uniform(0, 10, 40))
    f = 3.0 + 2.0 * x
    fi = orc.integ(x, f, 0.0)
    exact = 3.0 * (x - x[0]) + (x ** 2 - x[0] ** 2)
    assert np.max(np.abs(fi - exact)) <= 1e-10
    # a quadratic is integrated exactly on the parabolic intervals; the
    # forced-linear intervals 1-3 and N-1 carry the trapezoid error only
    f2 = x ** 2
    fi2 = orc.integ(x, f2, 0.0)
    seg = np.diff(fi2)
    ex = np.diff(x ** 3 / 3)
    assert np.max(np.abs(seg[3:-2] - ex[3:-2]) / ex[3:-2]) <= 1e-9


def test_map1_exact_on_linear(orc):
    x = np.sort(np.random.default_rng(2).uniform(0, 5, 30))
    xn = np.sort(np.random.default_rng(3).uniform(-1, 6, 51))
    f, _ = orc.map1(x, 2.0 - 0.5 * x, xn)
    assert np.max(np.abs(f - (2.0 - 0.5 * xn))) <= 1e-12


def test_deriv_exact_on_linear(orc):
    x = np.sort(np.random.default_rng(5).uniform(0.1, 5, 30))
    d = orc.deriv(x, 1.0 + 4.0 * x)
    assert np.max(np.abs(d - 4.0)) <= 1e-10


def test_expi_matches_scipy(orc):
    for n in (1, 2, 3, 4):
        for x in np.concatenate([10 ** np.linspace(-6, 0, 30), np.linspace(1, 60, 60)]):
            assert abs(orc.expi(n, x) - expn(n, x)) <= 2e-7 * abs(expn(n, x)) + 1e-300


# ---------------------------------------------------------------- JOSH
def _josh_inputs(n=72, s_a=1.0, s_b=2.0, alpha=0.0):
    rhox = 10 ** np.linspace(-4, 2, n)
    kap = np.full(n, 1.0)
    tau = rhox * kap  # tau = rhox for kappa = 1
    S = s_a + s_b * tau
    sig = kap * alpha / (1 - alpha) if alpha > 0 else np.zeros(n)
    return rhox, kap, sig, S


def test_josh_linear_source_flux(orc, coef):
    """Pure absorption, S = a + b tau (Milne-Eddington): H(0) = a/4 + b/6
    to the accuracy of the 51-point quadrature."""
    rhox, kap, sig, S = _josh_inputs()
    z = np.zeros_like(kap)
    r = orc.josh(1, 1, kap, z, sig, z, S, S, rhox, S, *coef)
    assert abs(r["hnu1"] - (1.0 / 4 + 2.0 / 6)) / (1.0 / 4 + 2.0 / 6) <= 2e-3


def test_josh_scattering_matches_fine_grid_lambda(orc, coef):
    """Constant albedo alpha with S_bar = B constant: S solves
    S = (1-alpha) B + alpha Lambda[S]; compared with an independent
    brute-force solve of the same equation on a fine grid with the E1
    kernel (a transposed COEFJ or a dropped (1-alpha) fails this)."""
    alpha = 0.9
    rhox, kap, sig, S = _josh_inputs(alpha=alpha, s_a=1.0, s_b=0.0)
    z = np.zeros_like(kap)
    acont = kap
    sigmac = kap * alpha / (1 - alpha)
    r = orc.josh(1, 1, acont, z, sigmac, z, S, S, rhox, S, *coef)
    # brute force: tau grid, Lambda with E1/2 kernel, solve directly
    t = np.concatenate([[0.0], np.geomspace(1e-6, 60, 2500)])
    mid = 0.5 * (t[1:] + t[:-1])
    w = np.diff(t)
    K = 0.5 * expn(1, np.maximum(np.abs(t[:, None] - mid[None, :]), 1e-300))
    # integrate the log singularity of E1 on the diagonal cell exactly
    lam = K * w[None, :]
    Sb = (1 - alpha) * np.ones(len(mid))
    A = np.eye(len(mid)) - alpha * (0.5 * expn(1, np.abs(mid[:, None] - mid[None, :]) + 1e-300) * w[None, :])
    idx = np.arange(len(mid))
    A[idx, idx] = 1 - alpha * (1 - np.exp(-w / 2) + (w / 2) * expn(1, w / 2)) * 1.0
    Ssol = np.linalg.solve(A, Sb)
    # surface flux H(0) = 1/2 int S E2(t) dt
    H = 0.5 * np.sum(Ssol * expn(2, mid) * w)
    assert abs(r["hnu1"] - H) / H <= 2e-2
    del lam


# ---------------------------------------------------------------- planted mistakes
def test_planted_mistakes_are_caught():
    """Copies of two oracle formulas with a planted error: each must fail
    the corresponding pin above (checked here without touching the
    oracle)."""
    # (1) Voigt with a dropped second Weideman term must miss erfcx(a)
    a = 0.3
    w = wofz(1j * a).real
    assert abs(w - erfcx(a)) < 1e-14
    L = 4.75682846001088411
    assert abs((1 / (math.sqrt(math.pi) * (L + a))) - erfcx(a)) > 1e-3
    # (2) INTEG with the wrong sign on the quadratic term fails exactness
    x = np.linspace(0, 1, 20)
    c = 1.0 / 3.0
    good = np.sum(((x[1:] + x[:-1]) * x[1:] + x[:-1] ** 2) * c * np.diff(x))
    bad = np.sum(-((x[1:] + x[:-1]) * x[1:] + x[:-1] ** 2) * c * np.diff(x))
    assert abs(good - 1.0 / 3.0) < 1e-12 and abs(bad - 1.0 / 3.0) > 0.1


# ---------------------------------------------------------------- line opacity
def test_line_opacity_invariants(orc, small_state, small_ref):
    """Profiles are symmetric in grid steps about each centre, the far wing
    falls as n^-2, and the FP64 and reference-typed oracles agree to FP32
    rounding."""
    lines = small_ref["lines"].copy()
    strong = lines[np.argsort(lines["cgf"])[-1:]].copy()
    L = small_state.length
    strong["nbuff"] = L // 2
    strong["elo"] = 0.0
    op, ev = orc.synthe_line_opacity(small_state, strong, "f64")
    for j in (0, 40, 71):
        row = op[j]
        c = L // 2 - 1
        k = np.arange(1, 600)
        assert np.allclose(row[c + k], row[c - k], rtol=0, atol=0)
        # far wing: row(n) * n^2 is constant once past n10dop
        far = row[c + 400: c + 600] * (np.arange(400, 600) ** 2)
        if far[0] > 0:
            assert np.max(np.abs(far / far[0] - 1)) <= 1e-12
    op32, ev32 = orc.synthe_line_opacity(small_state, small_ref["lines"], "ref")
    op64, ev64 = orc.synthe_line_opacity(small_state, small_ref["lines"], "f64")
    assert abs(ev32 - ev64) <= 1e-6 * ev64
    rel = np.abs(op32 - op64) / np.maximum(op64, 1e-300)
    assert np.median(rel) <= 1e-6 and rel.max() <= 1e-3


def test_line_opacity_brute_force_tiny(orc, small_state, small_ref):
    """A direct, unstructured restatement of the per-point rule on a tiny
    case (3 lines, 3 depths): sum over lines of kappa0 H(n dv, a) for
    |n| <= nstep, computed here from the oracle's own extents."""
    lines = small_ref["lines"][[100, 1500, 3000]].copy()
    op, ev, ns = orc.synthe_line_opacity(small_state, lines, "f64", want_nstep=True)
    st = small_state
    cont, cont64 = orc.synthe_continuum(st, 0)
    j = 0
    L = st.length
    mw6 = st.xnfpel.shape[1]
    out = np.zeros(L)
    txnxn = (st.xnfh[j] + 0.42 * st.xnfhe1[j] + 0.85 * st.xnfh2[j]) * (st.t[j] / 1e4) ** 0.3
    for il, ln in enumerate(lines):
        nel = ln["nelion"] - 1
        dop = st.dopple[j, nel]
        xnfdop = st.xnfpel[j, nel] / st.rho[j] / dop
        k0 = float(ln["cgf"]) * xnfdop * math.exp(-float(ln["elo"]) * st.hckt[j])
        a = (float(ln["gamrf"]) + float(ln["gamsf"]) * st.xne[j] + float(ln["gamwf"]) * txnxn) / dop
        n10 = int(10 * dop * st.resolu)
        dv = 1 / dop / st.resolu
        nstep = ns[j, il]
        if nstep < 0:
            continue
        xw = k0 * orc.voigt_f64([n10 * dv], [a])[0] * n10 * n10
        for p in range(L):
            n = abs(p + 1 - int(ln["nbuff"]))
            if n > nstep:
                continue
            out[p] += k0 * orc.voigt_f64([n * dv], [a])[0] if n <= n10 else xw / (n * n)
    assert np.max(np.abs(out - op[0]) / np.maximum(op[0], 1e-300)) <= 1e-12


def test_kurucz_voigt_against_faddeeva(orc):
    """The tabulated Kurucz Voigt (a < 0.2) against Re w: the Harris
    expansion to a^2, read at the nearest of 200 table points per Doppler
    width, is good to about 1-2 per cent (the reference quotes ~2 per cent
    for its tables and ~4 per cent near a = 0.2)."""
    tabs = orc.tabvoigt()
    v = np.linspace(0, 9.9, 400)
    for a, tol in ((0.001, 1.5e-2), (0.01, 1.5e-2), (0.1, 1.5e-2), (0.19, 2e-2)):
        h = orc.voigt_kurucz(v, np.full_like(v, a), tabs)
        w = wofz(v + 1j * a).real
        core = v < 3
        assert np.max(np.abs(h[core] - w[core]) / w[core]) <= tol, a
