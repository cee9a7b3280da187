"""GPU SOLVIT (src/atlas12_modules.f90:5739-5847) through a12_solvit_batched
against the oracle's restatement, which test_oracle_golden pins bit for bit
to the reference's own SOLVIT outputs.  Both device forms are covered: the
register form (order 51, as JOSH calls it) and the LDS form (any other
order, as NMOLEC calls it).  Cases: random systems with and without
diagonal dominance, pivot columns whose largest magnitudes tie (the
reference takes the first such row), integer matrices with ties at many
steps, a zero column (the 1e-30 pivot floor), a permutation matrix, and
JOSH's own I - diag(alpha) COEFJ.  Bar: bitwise (NaN where the oracle has
NaN)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cases(n, rng):
    out = []
    for _ in range(4):
        out.append(rng.standard_normal((n, n)) + n * rng.uniform(0.1, 1.0) * np.eye(n))
    for _ in range(4):
        out.append(rng.standard_normal((n, n)))
    a = rng.standard_normal((n, n))
    a[:, 0] = 0.5
    r1, r2 = (3, 7) if n > 7 else (0, n - 1)
    a[r1, 0], a[r2, 0] = 2.0, -2.0  # two rows tie for the first pivot
    out.append(a)
    a = rng.integers(-3, 4, (n, n)).astype(float)
    a += 7.0 * np.diag((rng.random(n) < 0.5).astype(float))
    out.append(a)
    a = rng.standard_normal((n, n))
    a[:, min(2, n - 1)] = 0.0  # a column with no pivot: floored to 1e-30
    out.append(a)
    out.append(np.eye(n)[::-1].copy())
    return out


def _check(gpu_ctx, orc, mats, rng):
    n = mats[0].shape[0]
    a = np.stack(mats)
    b = rng.standard_normal((len(mats), n))
    x = gpu_ctx.solvit_batched(a, b)
    for k in range(len(mats)):
        ref = orc.solvit(a[k], b[k])
        assert np.array_equal(x[k], ref, equal_nan=True), (n, k)


@pytest.mark.parametrize("n", [51, 33, 64, 2])
def test_solvit_matches_oracle_bitwise(gpu_ctx, orc, n):
    rng = np.random.default_rng(100 + n)
    _check(gpu_ctx, orc, _cases(n, rng), rng)


def test_solvit_josh_matrices(gpu_ctx, orc, coef):
    """The systems JOSH solves (:18550-18557): AMAT = I - diag(XALPHA) COEFJ
    with scattering fractions from 0 to 1, through the register form."""
    coefj = np.asarray(coef[0], float).reshape(51, 51).T  # [l, m] = COEFJ(l+1, m+1)
    rng = np.random.default_rng(7)
    mats = []
    for lo, hi in ((0.0, 1e-3), (0.0, 0.5), (0.5, 1.0), (0.99, 1.0)):
        al = rng.uniform(lo, hi, 51)
        mats.append(np.eye(51) - al[:, None] * coefj)
    _check(gpu_ctx, orc, mats, rng)


def test_solvit_argument_errors(gpu_ctx):
    from atlas12_amd import native
    with pytest.raises(native.A12Error):
        gpu_ctx.solvit_batched(np.zeros((1, 65, 65)), np.zeros((1, 65)))
    assert gpu_ctx.solvit_batched(np.zeros((0, 5, 5)), np.zeros((0, 5))).shape == (0, 5)
