"""The network oracle (oracle/mlp.py) against finite differences: its
backward pass is pinned in float64 (no BF16 rounding) on a random parameter
sample, and a planted mistake (the SiLU derivative replaced by the SiLU
itself) is caught by the same check."""
import numpy as np


def _fd_check(loss_and_grads, params, x, t, idx, h=1e-6):
    _, g = loss_and_grads(params, x, t, bf=False)
    fd = []
    for i in idx:
        p1 = params.copy()
        p1[i] += h
        p2 = params.copy()
        p2[i] -= h
        l1, _ = loss_and_grads(p1, x, t, bf=False)
        l2, _ = loss_and_grads(p2, x, t, bf=False)
        fd.append((l1 - l2) / (2 * h))
    return g[idx], np.array(fd)


def test_oracle_gradients_match_finite_differences():
    from oracle import mlp
    rng = np.random.default_rng(0)
    params = mlp.glorot(1).astype(np.float64) + 0.01 * rng.standard_normal(mlp.N_PARAMS)
    x = rng.standard_normal((6, 4))
    t = rng.standard_normal((6, 144))
    # indices in every layer's weights and biases
    idx = np.unique(np.concatenate([rng.integers(0, mlp.N_PARAMS, 40),
                                    [0, 1023, 1024 + 5, 1280 + 17, mlp.N_PARAMS - 1,
                                     mlp.N_PARAMS - 144 - 3]]))
    g, fd = _fd_check(mlp.loss_and_grads, params, x, t, idx)
    assert np.allclose(g, fd, rtol=1e-5, atol=1e-9)


def test_planted_mistake_is_caught(monkeypatch):
    from oracle import mlp
    rng = np.random.default_rng(1)
    params = mlp.glorot(2).astype(np.float64)
    x = rng.standard_normal((4, 4))
    t = rng.standard_normal((4, 144))
    idx = rng.integers(0, 1280, 20)  # first-layer parameters feel every SiLU'
    monkeypatch.setattr(mlp, "_dsilu", mlp._silu)
    g, fd = _fd_check(mlp.loss_and_grads, params, x, t, idx)
    assert not np.allclose(g, fd, rtol=1e-3, atol=1e-9)


def test_bf16_rounding():
    from oracle import mlp
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 0.0], np.float32)
    r = mlp.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0  # exactly half-way: ties go to the even mantissa
    assert r[2] == 1.0078125
    assert abs(r[3] + 3.140625) < 1e-7 and r[4] == 0.0
