"""The oracle's NMOLEC gas-phase Newton solve (src/atlas12_modules.f90:
9078-9217) pinned three ways:

* against the reference itself: NMOLEC's depth loop on two decks (solar
  4400-9500 K, cool dwarf 2800-6000 K), with the equilibrium constants the
  reference formed (tests/golden/molec_ref.npz), reproduces every
  converged XN and every XNMOL bit for bit;
* against closed forms (to 1e-4 relative, the reference's own
  convergence criterion: the last accepted step may be damped by 0.69): pure hydrogen ionisation (n_e^2 = K n_H) and pure
  H2 dissociation (n_H2 = K n_H^2), which fix the signs of the electron
  terms and the doubled Jacobian entries of a repeated component;
* against the physical equations it stands for: with H, H+, H- and
  electrons the solution conserves particles and nuclei and is neutral,
  which fails if the negative-ion correction has the wrong sign or count.
"""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def mz(golden):
    return golden("molec_ref.npz")


@pytest.mark.parametrize("case", ["solar", "cool"])
def test_chain_matches_reference_bitwise(orc, mz, case):
    xn, its, sts = orc.nmolec_chain(mz, case)
    assert np.all(sts == 0) and its.max() < 50
    assert np.array_equal(xn, orc.nmolec_reference_xn(mz, case))
    # XNMOL from the converged XN, as the reference forms it
    locj, kc, ne = mz["locj"], mz["kcomps"], len(mz["idequa"])
    eqj = mz[case + "_equilj"]
    for jm in range(len(locj) - 1):
        v = eqj[:, jm].copy()
        for l in range(locj[jm] - 1, locj[jm + 1] - 1):
            v = v / xn[:, ne - 1] if kc[l] == ne + 1 else v * xn[:, kc[l] - 1]
        assert np.array_equal(v, mz[case + "_xnmol"][jm]), jm


def test_molecules_matter_in_the_cool_deck(mz):
    """The cool deck exercises the molecular terms: at its top carbon is
    almost all in CO (code 608), so a dropped molecular term would show."""
    code = mz["xnmolcode"]
    co = mz["cool_xnmol"][int(np.flatnonzero(code == 608.0)[0])]
    c = mz["cool_xnmol"][int(np.flatnonzero(code == 6.0)[0])]
    assert co[0] > 1.0e4 * c[0]


def _table(entries, idequa):
    locj, kc = [1], []
    for comps in entries:
        kc += comps
        locj.append(len(kc) + 1)
    return {"locj": np.array(locj), "kcomps": np.array(kc), "idequa": np.array(idequa)}


@pytest.mark.parametrize("k", [1e10, 1e13, 1e16])
def test_hydrogen_ionisation_closed_form(orc, k):
    tb = _table([[2], [2, 4]], [0, 1, 100])
    xntot = 1.0e15
    xn, _, st = orc.molec_newton(tb, xntot, [0.0, 1.0, 0.0], [1.0, k],
                                 [xntot / 2, xntot / 2, xntot / 20])
    ne = xntot / (1.0 + np.sqrt(1.0 + xntot / k))
    nh = xntot - 2.0 * ne
    assert st == 0
    np.testing.assert_allclose(xn, [nh + ne, nh, ne], rtol=1e-4)


@pytest.mark.parametrize("k", [1e-17, 1e-15, 1e-12])
def test_h2_dissociation_closed_form(orc, k):
    tb = _table([[2], [2, 2]], [0, 1])
    xntot = 1.0e16
    xn, _, st = orc.molec_newton(tb, xntot, [0.0, 1.0], [1.0, k], [xntot, xntot / 10])
    nh = 2.0 * xntot / (1.0 + np.sqrt(1.0 + 4.0 * k * xntot))
    assert st == 0
    assert k * nh * nh > 0.01 * nh  # the molecule matters
    np.testing.assert_allclose(xn, [nh + 2.0 * k * nh * nh, nh], rtol=1e-4)


def test_negative_ion_is_neutral(orc):
    tb = _table([[2], [2, 4], [2, 3]], [0, 1, 100])
    xntot, k1, k2 = 1.0e15, 1.0e12, 2.0e-16
    xn, _, st = orc.molec_newton(tb, xntot, [0.0, 1.0, 0.0], [1.0, k1, k2],
                                 [xntot / 2, xntot / 2, xntot / 20])
    nuc, nh, ne = xn
    hp, hm = k1 * nh / ne, k2 * nh * ne
    assert st == 0 and hm > 0.02 * hp  # H- matters to the charge balance
    assert abs(nh + hp + hm + ne - xntot) <= 1e-4 * xntot
    assert abs(nh + hp + hm - nuc) <= 1e-4 * nuc
    assert abs(ne - (hp - hm)) <= 1e-4 * ne
