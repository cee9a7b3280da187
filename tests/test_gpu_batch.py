"""Batched SYNTHE through the C-ABI: five models with their own windows on
one log-lambda grid (one starting before, two inside and one past the
reference case's window), their own layer counts (72, 48, 36, 72, 60) and
one with no lines, over one shared line list; every model's opacity
(mixed and FP64), evaluation count and transfer output equal the
single-model pipeline's on the same inputs exactly, for one group or
several (a small work budget), eager or replayed from the hipGraph, and for
every gather tile shape."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch_case(small_state, small_ref):
    st = small_state
    lines = np.sort(small_ref["lines"], order="nbuff", kind="stable")
    # (window offset, length, layers)
    spec = [(0, st.length, np.arange(72)), (1000, 2000, np.arange(48)),
            (-300, 1500, np.arange(0, 72, 2)), (2500, 2475, np.arange(72)),
            (900, 600, np.arange(6, 66))]
    states, shift, lo, hi = [], [], [], []
    nb = lines["nbuff"]
    for k, (off, length, depths) in enumerate(spec):
        states.append(st.window(off, length).depths(depths))
        shift.append(off)
        if k == 4:  # no lines at all
            lo.append(0)
            hi.append(0)
        else:
            lo.append(int(np.searchsorted(nb, off - 400)))
            hi.append(int(np.searchsorted(nb, off + length + 400, side="right")))
    return states, lines, np.array(shift), np.array(lo), np.array(hi)


def _single(ctx, st, lines, prec, transfer):
    ctx.synthe_load(st, lines)
    ctx.synthe_opacity_run(prec)
    op = ctx.synthe_opacity_fetch()
    ev = ctx.eval_split()
    hl = hc = None
    if transfer:
        ctx.synthe_transfer_run()
        hl, hc = ctx.synthe_transfer_fetch()
    return op, ev, hl, hc


@pytest.fixture(scope="module")
def singles(gpu_ctx, batch_case, coef):
    from atlas12_amd import native
    states, lines, shift, lo, hi = batch_case
    gpu_ctx.set_operators(*coef)
    gpu_ctx.synthe_set_shape(0)
    out = {}
    for prec in (native.PREC_MIXED, native.PREC_FP64):
        for k, st in enumerate(states):
            sub = lines[lo[k]:hi[k]].copy()
            sub["nbuff"] -= shift[k]
            out[prec, k] = _single(gpu_ctx, st, sub, prec, prec == native.PREC_MIXED)
    return out


def _check(ctx, states, singles, prec, transfer):
    for k in range(len(states)):
        op, hl, hc, ev = ctx.synthe_batch_fetch(k)
        sop, sev, shl, shc = singles[prec, k]
        assert np.array_equal(op, sop), k
        assert tuple(ev) == tuple(sev), k
        if transfer:
            assert np.array_equal(hl, shl) and np.array_equal(hc, shc), k


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_equals_single_models(gpu_ctx, batch_case, singles, prec):
    states, lines, shift, lo, hi = batch_case
    gpu_ctx.synthe_set_shape(0)
    gpu_ctx.synthe_batch_load(states, lines, shift, lo, hi)
    transfer = prec == 0
    gpu_ctx.synthe_batch_run(prec, transfer=transfer)
    assert gpu_ctx.synthe_batch_info()["groups"] == 1
    _check(gpu_ctx, states, singles, prec, transfer)
    assert np.all(gpu_ctx.synthe_batch_fetch(4)[0] == 0)


def test_batch_groups_and_graph_replay(gpu_ctx, batch_case, singles):
    states, lines, shift, lo, hi = batch_case
    gpu_ctx.synthe_set_shape(0)
    gpu_ctx.synthe_batch_load(states, lines, shift, lo, hi)
    gpu_ctx.synthe_batch_run(1)  # FP64 work arrays are the larger
    gpu_ctx.synthe_batch_set_budget(gpu_ctx.synthe_batch_info()["largest_model_work"])
    for _ in range(3):  # capture, then two replays with eager work between
        gpu_ctx.synthe_batch_run(0, transfer=True, graph=True)
        assert gpu_ctx.synthe_batch_info()["groups"] > 1
        _check(gpu_ctx, states, singles, 0, True)
        gpu_ctx.synthe_batch_run(1, transfer=False, graph=False)
        assert gpu_ctx.synthe_batch_info()["groups"] > 1
        _check(gpu_ctx, states, singles, 1, False)


def test_every_tile_shape_gives_the_same_batch(gpu_ctx, batch_case, singles):
    from atlas12_amd import native
    states, lines, shift, lo, hi = batch_case
    gpu_ctx.synthe_batch_load(states, lines, shift, lo, hi)
    for shape in range(len(native.synthe_shapes())):
        gpu_ctx.synthe_set_shape(shape)
        gpu_ctx.synthe_batch_run(0)
        _check(gpu_ctx, states, singles, 0, False)
    gpu_ctx.synthe_set_shape(0)


def test_batch_rejects_unsorted_lines(gpu_ctx, batch_case):
    from atlas12_amd import native
    states, lines, shift, lo, hi = batch_case
    with pytest.raises(native.A12Error):
        gpu_ctx.synthe_batch_load(states, lines[::-1].copy(), shift, lo, hi)
