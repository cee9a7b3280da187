"""GPU starting-structure network through the C-ABI: the MFMA tile kernel,
forward pass and gradients against the numpy oracle (oracle/mlp.py, which
emulates the BF16 operand rounding and is itself pinned by finite
differences), determinism of the graph-replayed training, and what the
trained network gains over SCALE_MODEL's rescaled start.

Tolerances: the GEMM sums exact BF16 products in FP32 (1e-5 relative to
the float64 sum); forward and gradients differ from the oracle only by FP32
vs float64 accumulation and the occasional BF16 rounding that flips with
it (relative L2 2e-3 / 5e-3)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bits(a):
    from oracle import mlp
    return mlp.bf16(a).view(np.uint32).__rshift__(16).astype(np.uint16)


@pytest.mark.parametrize("m,n,k,split", [(80, 144, 48, 1), (64, 64, 16, 1), (130, 70, 37, 1),
                                         (256, 144, 2048, 16), (16, 256, 1000, 7)])
def test_mfma_gemm_matches_numpy(gpu_ctx, m, n, k, split):
    from atlas12_amd import native
    from oracle import mlp
    rng = np.random.default_rng(m + n + k)
    a = mlp.bf16(rng.standard_normal((m, k)).astype(np.float32))
    b = mlp.bf16(rng.standard_normal((k, n)).astype(np.float32))
    c = native.mlp_gemm(gpu_ctx, _bits(a), _bits(b), split)
    ref = a.astype(np.float64) @ b.astype(np.float64)
    assert np.max(np.abs(c - ref)) <= 1e-5 * np.max(np.abs(ref)) + 1e-6


def _rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, float) - b) / np.linalg.norm(b))


def test_forward_and_gradients_match_oracle(gpu_ctx):
    from atlas12_amd import native
    from oracle import mlp
    rng = np.random.default_rng(3)
    params = (mlp.glorot(4) + 0.02 * rng.standard_normal(mlp.N_PARAMS)).astype(np.float32)
    x = rng.standard_normal((100, 4)).astype(np.float32)
    t = rng.standard_normal((100, 144)).astype(np.float32)
    native.mlp_init(gpu_ctx, 128, 0)
    native.mlp_set_params(gpu_ctx, params)
    assert np.array_equal(native.mlp_get_params(gpu_ctx), params)
    y = native.mlp_infer(gpu_ctx, x)
    y_ref, _ = mlp.forward(params, x)
    assert _rel_l2(y, y_ref) <= 2e-3
    native.mlp_set_data(gpu_ctx, x, t)
    loss, g = native.mlp_grads(gpu_ctx)
    loss_ref, g_ref = mlp.loss_and_grads(params, x, t)
    assert abs(loss - loss_ref) <= 1e-4 * loss_ref
    o = 0
    for i, nn in mlp.SIZES:
        for size in (i * nn, nn):
            assert _rel_l2(g[o:o + size], g_ref[o:o + size]) <= 5e-3, (i, nn, size)
            o += size


def test_training_is_deterministic(gpu_ctx):
    """The same seed and data give bit-identical weights, whether the steps
    come in one call or several (one graph replay per step)."""
    from atlas12_amd import native
    rng = np.random.default_rng(5)
    x = rng.standard_normal((200, 4)).astype(np.float32)
    t = np.tanh(x @ rng.standard_normal((4, 144))).astype(np.float32)
    runs = []
    for split in ((40,), (10, 10, 20)):
        native.mlp_init(gpu_ctx, 256, 7)
        native.mlp_set_data(gpu_ctx, x, t)
        l0 = native.mlp_grads(gpu_ctx)[0]
        for s in split:
            loss = native.mlp_train(gpu_ctx, s, 1e-3)
        runs.append(native.mlp_get_params(gpu_ctx))
        assert loss < l0
    assert np.array_equal(runs[0], runs[1])


def test_surrogate_start_halves_iterations(gpu_ctx, coef):
    """The learned start against SCALE_MODEL's, by iterations of the model
    iteration (a12_model_converge) to the convergence criterion (max |ERROR|
    <= 1 %, max |DERIV| <= 10 %, max |T1| <= 1 K on TCORR(3)'s table):
    trained on the GPU-converged models of 2,000 seeded draws, evaluated on
    200 held-out draws each started from SCALE_MODEL of the converged solar
    model and from the network.  The median over the held-out models (a model
    that does not converge within 150 iterations counts as 150) is at most
    half the scaled start's, and no more models fail to converge from the
    network's start than from the scaled one."""
    from atlas12_amd import mlp
    gpu_ctx.set_operators(*coef)
    r = mlp.pipeline(gpu_ctx, 2000, 200, seed=10, itmax=150)
    assert r["sun"].iters[0] > 0
    n_tr = int(np.sum(r["train"].iters > 0))
    assert n_tr >= 1800, n_tr
    assert r["model"].losses[-1] < 1e-3 * r["model"].losses[0]
    it_s = np.where(r["test"].iters > 0, r["test"].iters, 150)
    it_n = np.where(r["mlp_run"].iters > 0, r["mlp_run"].iters, 150)
    assert np.median(it_n) <= 0.5 * np.median(it_s), (np.median(it_n), np.median(it_s))
    assert np.sum(r["mlp_run"].iters <= 0) <= np.sum(r["test"].iters <= 0)
    # and the network's structure is close to the converged held-out models
    ok = r["test"].iters > 0
    dt, dp = mlp.structure_errors(r["pred"][ok], mlp.targets(r["test"])[ok])
    assert np.median(dt) < 0.02 and np.median(dp) < 0.1
