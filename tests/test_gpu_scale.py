"""GPU SCALE_MODEL (src/atlas12_modules.f90:6967-7060) through the C-ABI,
batched, against the reference's own outputs (tests/golden/scale_model.npz)
and the oracle.  Bar: bit for bit (INTEG and MAP1 are restated bit-faithfully
and TAUSTD, GNEW use the host libm as the reference does)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_scale_model_matches_reference(gpu_ctx, golden):
    from atlas12_amd import native
    z = golden("scale_model.npz")
    par = z["params"]
    sc = np.stack([np.r_[p[:5], p[7], p[8]] for p in par])
    out, so, tau = native.scale_model(gpu_ctx, sc, z["inp"], kw=80, tau1lg=par[0][5],
                                      steplg=par[0][6])
    assert np.array_equal(tau, z["taustd"][0])
    for k in range(len(par)):
        ref = z["out"][k]
        assert np.array_equal(out[k, :8], ref[:8]), k  # 7 fields + PRADK
        if k >= 2:  # rescaled cases
            assert np.array_equal(out[k, 8], ref[8]), k
            assert np.array_equal(so[k], z["scal"][k]), k
        else:
            assert np.array_equal(so[k, [0, 1, 4]], z["scal"][k][[0, 1, 4]]), k


def test_scale_model_batch_vs_oracle(gpu_ctx, golden, orc):
    """256 models of mixed sizes of rescale: every one equals the oracle."""
    from atlas12_amd import native
    z = golden("scale_model.npz")
    rng = np.random.default_rng(5)
    base = z["inp"][0]
    nm = 256
    fields = np.repeat(base[None], nm, 0) * (1.0 + 0.01 * rng.standard_normal((nm, 7, 1)))
    fields[:, 0] = np.sort(fields[:, 0], axis=1)  # RHOX ascending
    sc = np.zeros((nm, 7))
    sc[:, :5] = z["params"][0][:5]
    sc[:, 5] = rng.uniform(3500, 8000, nm)
    sc[:, 6] = rng.uniform(1.0, 5.0, nm)
    sc[::7, 5] = 0.0  # some regrid-only
    out, so, _ = native.scale_model(gpu_ctx, sc, fields)
    for m in range(nm):
        o, s, resc = orc.scale_model(fields.shape[2], 80, -6.875, 0.125, sc[m, :5], sc[m, 5],
                                     sc[m, 6], fields[m])
        assert np.array_equal(out[m, :8], o[1:9]), m
        if resc:
            assert np.array_equal(out[m, 8], o[9]) and np.array_equal(so[m], s), m
