"""Pins of the model iteration's CPU restatement (oracle/model_iter.py)
against closed forms that do not reuse its own expressions: the reduced
physics is this library's, so no reference output pins it (parity
against the reference is unpinned for it; these are the checks)."""
import numpy as np

from oracle import model_iter as mi


def test_eos_conserves_particles_and_charge():
    """Charge balance n_e = n(H II) + n(donor II) and the ideal-gas count
    P = (n_H (1 + He + A_M) + n_e) k T, from the returned populations."""
    T = np.array([3500.0, 5000.0, 7000.0, 12000.0])
    pg = np.array([1e3, 1e5, 1e4, 1e5])
    am = 1e-4
    p = mi.eos(T, pg, am)
    nh = p["nh0"] + p["nhp"]
    nm = am * nh
    assert np.allclose(p["ne"], p["nhp"] + (nm - p["nm0"]), rtol=1e-10)
    assert np.allclose((nh * (1 + mi.HE + am) + p["ne"]) * mi.KB * T, pg, rtol=1e-12)


def test_eos_pure_hydrogen_saha():
    """Without the donor (A_M -> 0) the electrons are the hydrogen's:
    n_e n(H II) / n(H I) = 2.4147e15 T^1.5 exp(-157807.7 / T) (Saha for H
    with U(H I) = 2, U(H II) = 1)."""
    T = np.array([6000.0, 8000.0, 10000.0])
    p = mi.eos(T, np.full(3, 1e4), 1e-30)
    saha = 2.4147e15 * T ** 1.5 * np.exp(-157807.7 / T)
    assert np.allclose(p["ne"] * p["nhp"] / p["nh0"], saha, rtol=1e-9)
    assert np.allclose(p["ne"], p["nhp"], rtol=1e-12)


def test_planck_quadrature_on_the_grid():
    """sum over the grid of B_nu RCOSET = sigma T^4 / pi to the grid's
    trapezoid accuracy (1,200 points, 91 nm - 91 um): 0.5 % at 5,000 K."""
    T = np.array([5000.0])
    p = mi.eos(T, np.array([1e5]), 1e-4)
    _, _, _, b = mi.kappa(T, p)
    tot = np.sum(b[:, 0] * mi.RCO)
    assert abs(tot / (mi.SIG * 5000.0**4 / np.pi) - 1) < 5e-3


def test_opacity_edges():
    """H- bound-free vanishes beyond its 1.64 um threshold, the H I Lyman
    continuum appears only above 3.2898e15 Hz, and Rayleigh scattering falls
    as lambda^-4 in the red."""
    w = mi.WAVE * 10.0
    assert np.all(mi.PF["hmbf"][w >= 16419.0] == 0.0) and np.all(mi.PF["hmbf"][w < 16000.0] > 0)
    assert np.all(mi.PF["nmin"][mi.FREQ < 3.2898e15] >= 2)
    assert np.all(mi.PF["nmin"][mi.FREQ >= 3.2898e15] == 1)
    i, k = np.searchsorted(w, 20000.0), np.searchsorted(w, 40000.0)
    assert abs(mi.PF["ray"][i] / mi.PF["ray"][k] * (w[i] / w[k]) ** 4 - 1) < 1e-2  # lambda^-6 term: 0.5 %


def test_grid_is_the_driver_grid():
    """WAVESET(NU) = 10^(1 + 1e-4 (NU + NUSTART - 1)) at every 25th NU from
    NUSTART 9599 (91.2 nm), and the RCOSET end rules."""
    assert abs(mi.WAVE[0] - 10.0 ** (1.0 + 0.0001 * 9599)) < 1e-12
    assert np.allclose(mi.WAVE[1:] / mi.WAVE[:-1], 10.0 ** 0.0025)
    assert mi.RCO[0] == (mi.FREQ[0] - mi.FREQ[1]) * 1.5
    assert mi.RCO[-1] == (mi.FREQ[-2] + mi.FREQ[-1]) * 0.25
