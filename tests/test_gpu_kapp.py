"""KAPP on the GPU (a12_kapp_set_tables/_set_state/_eval, a12_atlas_kapp_run)
against the reference's own KAPP (src/atlas12_modules.f90:11202-11324):

* every per-source array at 349 wavelengths (8 nm - 300 um, clustered on
  the edges and H lines the sources switch on) on three decks
  (tests/golden/kapp_ref.npz);
* the six KAPP tables on the full 30,000-point grid of the ATLAS12 workload
  at every 100th frequency (tests/golden/atlas_plain_state.npz);
* all 240 frequencies of the small ATLAS12 case, and RUN_RT_PASS on the
  GPU's KAPP against the reference's accumulators (atlas_small.npz).

Tolerance: the kernels evaluate the reference's expressions in its order
(the same source built for the host reproduces every fixture bit for bit,
tests/test_kapp.py); the device's exp/log/pow may differ from the host libm
in the last bit, so 1e-12 relative per element (the RT accumulators 1e-12
of their largest element, as tests/test_gpu_atlas.py)."""
import os

import numpy as np
import pytest

from tests.test_kapp import DECKS, deck_state

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-12


def _rel(a, b):
    a, b = np.atleast_1d(np.asarray(a, float)), np.atleast_1d(np.asarray(b, float))
    d = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
    d[(a == 0) & (b == 0)] = 0.0
    return d


@pytest.fixture(scope="module")
def ref():
    return np.load(os.path.join(GOLDEN, "kapp_ref.npz"))


@pytest.fixture(scope="module")
def ktables():
    from atlas12_amd import native
    return native.kapp_tables(np.load(os.path.join(GOLDEN, "kapp_tables.npz")))


@pytest.mark.parametrize("deck", DECKS)
def test_every_source_matches_reference(gpu_ctx, ref, ktables, deck):
    from atlas12_amd import native
    native.kapp_set_tables(gpu_ctx, ktables)
    native.kapp_set_state(gpu_ctx, native.kapp_state(deck_state(ref, deck)))
    out = native.kapp_eval(gpu_ctx, ref["waves"], len(ref[deck + "_t"]))
    depths = ref["depths"]
    for k in native.KAPP_SRC:
        r = _rel(out[k][:, depths], ref[f"{deck}_{k}"])
        assert r.max() <= TOL, (k, r.max())


def _c2(gpu_ctx, ref, ktables):
    from atlas12_amd import native
    z = np.load(os.path.join(GOLDEN, "atlas_plain_state.npz"))
    small = np.load(os.path.join(GOLDEN, "atlas_small.npz"))
    st = deck_state(ref, "plain")
    assert np.array_equal(st["t"], z["t"]) and np.array_equal(st["xne"], z["xne"])
    d = native.atlas_state_from(z)
    native.atlas_load_lines(gpu_ctx, d, small["linedata"])
    native.atlas_set_state(gpu_ctx, d)
    native.kapp_set_tables(gpu_ctx, ktables)
    native.kapp_set_state(gpu_ctx, native.kapp_state(st))
    return z


def test_full_grid_matches_reference(gpu_ctx, ref, ktables):
    """a12_atlas_kapp_run on the 30,000-frequency grid (91.2 nm - 91.2 um)
    against the reference's KAPP at every 100th frequency."""
    from atlas12_amd import native
    z = _c2(gpu_ctx, ref, ktables)
    native.atlas_kapp_run(gpu_ctx)
    k = native.atlas_kapp_fetch(gpu_ctx)
    sub = z["kapp_sub"]
    for name in ("acont", "sigmac", "scont", "sline", "aline", "sigmal"):
        r = _rel(k[name][sub], z[name + "_sub"])
        assert r.max() <= TOL, (name, r.max())
    # the grid run and a12_kapp_eval at the same wavelengths agree bit for bit
    e = native.kapp_eval(gpu_ctx, z["waveset"][sub], len(z["t"]))
    for name in ("acont", "sigmac", "scont", "sline", "aline", "sigmal"):
        assert np.array_equal(e[name], k[name][sub]), name


def test_rt_pass_on_gpu_kapp_matches_reference(gpu_ctx, ref, ktables, coef):
    """The small ATLAS12 case (plain deck, 240 frequencies from 500 nm):
    KAPP at every frequency, and RUN_RT_PASS's accumulators on it against
    the reference's own pass (its KAPP, its frequency loop)."""
    from atlas12_amd import native
    z = np.load(os.path.join(GOLDEN, "atlas_small.npz"))
    st = deck_state(ref, "plain")
    assert np.array_equal(st["t"], z["t"])
    d = native.atlas_state_from(z)
    gpu_ctx.set_operators(*coef)
    native.atlas_load(gpu_ctx, d, z["linedata"])
    native.kapp_set_tables(gpu_ctx, ktables)
    native.kapp_set_state(gpu_ctx, native.kapp_state(st))
    native.atlas_kapp_run(gpu_ctx)
    k = native.atlas_kapp_fetch(gpu_ctx)
    for name in ("acont", "sigmac", "scont", "sline", "aline", "sigmal"):
        assert _rel(k[name], z[name]).max() <= TOL, name
    acc, h1 = native.atlas_rt_pass(gpu_ctx, z["xlines"])
    for name in ("rjmins", "flxrad", "raden", "accrad", "abross"):
        scale = np.max(np.abs(z["acc_" + name]))
        assert np.max(np.abs(acc[name] - z["acc_" + name])) <= TOL * scale, name
    assert _rel(acc["pradk0"], z["acc_pradk0"]) <= TOL
    assert _rel(h1[-1], z["last_hnu"][0]) <= TOL


def test_state_of_another_atmosphere_rejected(gpu_ctx, ref, ktables):
    from atlas12_amd import native
    _c2(gpu_ctx, ref, ktables)
    native.kapp_set_state(gpu_ctx, native.kapp_state(deck_state(ref, "hot")))
    with pytest.raises(native.A12Error, match="not the ATLAS12 state"):
        native.atlas_kapp_run(gpu_ctx)


def test_tables_required(gpu_ctx, ref):
    from atlas12_amd import native
    ctx = native.Context(0)
    try:
        with pytest.raises(native.A12Error, match="set_tables"):
            native.kapp_set_state(ctx, native.kapp_state(deck_state(ref, "plain")))
    finally:
        ctx.close()
