"""GPU NMOLEC gas-phase Newton (src/atlas12_modules.f90:9078-9217) through
the C-ABI against the oracle, which reproduces the reference bit for bit
(tests/test_oracle_molec.py):

* NMOLEC's depth loop as one chain per deck: XN, XNMOL and the iteration
  count of every depth equal the oracle's and the reference's exactly;
* the same depths as independent systems started where the chain starts
  them: identical results;
* a batch of 4,608 systems (64 perturbed copies of both decks' 72 depths)
  in one launch: every system equal to the oracle's solve of it;
* the closed-form hydrogen system, and the argument checks.

The solve is FP64 with contraction off and the reference's operation
order, so every comparison is exact."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mz(golden):
    return golden("molec_ref.npz")


def _set(gpu_ctx, tb):
    from atlas12_amd import native
    native.molec_set_table(gpu_ctx, tb["locj"], tb["kcomps"], tb["idequa"])


def _chain_inputs(orc, mz, case):
    xntot, xab, start = orc.nmolec_inputs(mz, case)
    p = mz[case + "_p"]
    pr = np.ones(len(p))
    pr[1:] = p[1:] / p[:-1]
    return xntot, xab, start, pr


@pytest.mark.parametrize("case", ["solar", "cool"])
def test_chain_matches_oracle_and_reference(gpu_ctx, orc, mz, case):
    from atlas12_amd import native
    _set(gpu_ctx, mz)
    xntot, xab, start, pr = _chain_inputs(orc, mz, case)
    n = len(xntot)
    x0 = np.zeros_like(xab)
    x0[0] = start
    xn, xm, it, st = native.molec_newton(gpu_ctx, xntot, xab, mz[case + "_equilj"], x0,
                                         chain_off=[0, n], pratio=pr)
    oxn, oit, _ = orc.nmolec_chain(mz, case)
    assert np.all(st == 0)
    assert np.array_equal(it, oit)
    assert np.array_equal(xn, oxn)
    assert np.array_equal(xn, orc.nmolec_reference_xn(mz, case))
    assert np.array_equal(xm.T, mz[case + "_xnmol"])


def test_independent_systems_equal_the_chain(gpu_ctx, orc, mz):
    from atlas12_amd import native
    _set(gpu_ctx, mz)
    xntot, xab, start, pr = _chain_inputs(orc, mz, "cool")
    oxn, oit, _ = orc.nmolec_chain(mz, "cool")
    x0 = np.vstack([start, oxn[:-1] * pr[1:, None]])
    xn, _, it, st = native.molec_newton(gpu_ctx, xntot, xab, mz["cool_equilj"], x0)
    assert np.all(st == 0) and np.array_equal(it, oit) and np.array_equal(xn, oxn)


def test_large_batch_matches_oracle(gpu_ctx, orc, mz):
    from atlas12_amd import native
    _set(gpu_ctx, mz)
    rng = np.random.default_rng(12)
    xt, xa, eq, x0 = [], [], [], []
    for case in ("solar", "cool"):
        xntot, xab, start, pr = _chain_inputs(orc, mz, case)
        oxn, _, _ = orc.nmolec_chain(mz, case)
        base = np.vstack([start, oxn[:-1] * pr[1:, None]])
        for _ in range(32):
            xt.append(xntot)
            xa.append(xab)
            eq.append(mz[case + "_equilj"])
            x0.append(base * rng.uniform(0.5, 2.0, base.shape))
    xt, xa, eq, x0 = (np.concatenate(a) for a in (xt, xa, eq, x0))
    xn, xm, it, st = native.molec_newton(gpu_ctx, xt, xa, eq, x0)
    assert len(xt) == 4608 and np.all(st == 0)
    for s in range(len(xt)):
        oxn, oit, ost = orc.molec_newton(mz, xt[s], xa[s], eq[s], x0[s])
        assert ost == 0 and oit == it[s] and np.array_equal(oxn, xn[s]), s


def test_hydrogen_closed_form(gpu_ctx, orc):
    from atlas12_amd import native
    tb = {"locj": [1, 2, 4], "kcomps": [2, 2, 4], "idequa": [0, 1, 100]}
    _set(gpu_ctx, tb)
    ks = np.array([1e10, 1e13, 1e16])
    xntot = np.full(3, 1.0e15)
    xab = np.tile([0.0, 1.0, 0.0], (3, 1))
    eqj = np.stack([np.ones(3), ks], 1)
    x0 = np.tile([5e14, 5e14, 5e13], (3, 1))
    xn, _, _, st = native.molec_newton(gpu_ctx, xntot, xab, eqj, x0)
    ne = xntot / (1.0 + np.sqrt(1.0 + xntot / ks))
    nh = xntot - 2.0 * ne
    assert np.all(st == 0)
    np.testing.assert_allclose(xn, np.stack([nh + ne, nh, ne], 1), rtol=1e-4)
    for s in range(3):
        assert np.array_equal(xn[s], orc.molec_newton(tb, xntot[s], xab[s], eqj[s], x0[s])[0])


def test_rejects_bad_input(gpu_ctx, mz):
    from atlas12_amd import native
    bad = dict(locj=[1, 2, 4], kcomps=[2, 2, 7], idequa=[0, 1, 100])
    with pytest.raises(native.A12Error):
        _set(gpu_ctx, bad)
    _set(gpu_ctx, mz)
    ne, nm = len(mz["idequa"]), len(mz["xnmolcode"])
    with pytest.raises(native.A12Error):  # chains longer than one need pratio
        native.molec_newton(gpu_ctx, np.ones(2), np.ones((2, ne)), np.ones((2, nm)),
                            np.ones((2, ne)), chain_off=[0, 2])
    with pytest.raises(native.A12Error):  # chain offsets must end at nsys
        native.molec_newton(gpu_ctx, np.ones(2), np.ones((2, ne)), np.ones((2, nm)),
                            np.ones((2, ne)), chain_off=[0, 1])
