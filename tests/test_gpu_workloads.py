"""GPU runs at the shapes of the batched and large SYNTHE workloads, with
the generators bench.py uses and reduced line lists:

* batched SYNTHE: six models of the 256-model draw (own layer counts,
  Teff, log g, [M/H], windows at R = 300,000) over a 300,000-line list --
  every model's opacity and flux equal the single-model pipeline's bit for
  bit, and mixed against FP64 emergent flux stays within 1e-4;
* large SYNTHE: the cool-dwarf structure over 300-1000 nm at R = 500,000
  (601,987 points) with 500,000 lines, 70 % in molecular band progressions
  -- mixed against FP64 emergent flux within 1e-4 and the molecular lines
  adding opacity.

The 1e-4 bound is the metric's tolerance on emergent flux."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def base():
    from atlas12_amd.state import SyntheState
    return SyntheState.load(os.path.join(GOLDEN, "synthe_plain_state.npz"))


def test_batch_workload_models(gpu_ctx, base, coef):
    from atlas12_amd import native
    from atlas12_amd import workloads as wl
    g, models, states, lines, sh, lo, hi = wl.synthe_batch_workload(base, 256, 300_000)
    sel = sorted(range(256), key=lambda k: models[k].length)[:6]
    gpu_ctx.set_operators(*coef)
    gpu_ctx.synthe_set_shape(0)
    gpu_ctx.synthe_batch_load([states[k] for k in sel], lines, sh[sel], lo[sel], hi[sel])
    out = {}
    for prec in (native.PREC_MIXED, native.PREC_FP64):
        gpu_ctx.synthe_batch_run(prec, transfer=True, graph=prec == native.PREC_MIXED)
        out[prec] = [gpu_ctx.synthe_batch_fetch(i) for i in range(len(sel))]
    for i, k in enumerate(sel):
        hm, hd = out[native.PREC_MIXED][i][1], out[native.PREC_FP64][i][1]
        assert np.max(np.abs(hm - hd) / np.abs(hd)) <= 1e-4, k
        assert out[native.PREC_MIXED][i][3][0] > 0
    for i in (0, len(sel) - 1):  # the smallest and the largest of the six, alone
        k = sel[i]
        sub = lines[lo[k]:hi[k]].copy()
        sub["nbuff"] -= sh[k]
        gpu_ctx.synthe_load(states[k], sub)
        gpu_ctx.synthe_opacity_run(native.PREC_MIXED)
        assert np.array_equal(gpu_ctx.synthe_opacity_fetch(), out[native.PREC_MIXED][i][0])
        gpu_ctx.synthe_transfer_run()
        assert np.array_equal(gpu_ctx.synthe_transfer_fetch()[0], out[native.PREC_MIXED][i][1])


def test_large_workload_shape(gpu_ctx, base, coef):
    from atlas12_amd import native
    from atlas12_amd import workloads as wl
    grid = wl.Grid.make(300.0, 1000.0, 500000.0)
    st = wl.cool_dwarf_state(base, grid)
    lines = wl.large_synthe_lines(grid, 500_000)
    mol = np.isin(lines["nelion"], [s for s, _ in wl.LARGE_MOLECULES])
    assert 0.65 < mol.mean() < 0.75
    gpu_ctx.set_operators(*coef)
    gpu_ctx.synthe_set_shape(0)
    res = {}
    for prec in (native.PREC_MIXED, native.PREC_FP64):
        gpu_ctx.synthe_load(st, lines)
        gpu_ctx.synthe_opacity_run(prec)
        gpu_ctx.synthe_transfer_run()
        res[prec] = (gpu_ctx.synthe_transfer_fetch()[0], gpu_ctx.eval_split()[0])
    hm, hd = res[native.PREC_MIXED][0], res[native.PREC_FP64][0]
    assert np.max(np.abs(hm - hd) / np.abs(hd)) <= 1e-4
    # the molecular bands add opacity: without them the line flux is higher
    gpu_ctx.synthe_load(st, lines[~mol])
    gpu_ctx.synthe_opacity_run(native.PREC_MIXED)
    gpu_ctx.synthe_transfer_run()
    h_atomic = gpu_ctx.synthe_transfer_fetch()[0]
    assert np.all(h_atomic >= hm * (1 - 1e-6)) and np.mean(h_atomic > hm * (1 + 1e-4)) > 0.05
