"""GPU parity at the full sizes of the benchmark configurations, with the
generators bench.py uses, through the C-ABI, against the oracle:

* C2, one ATLAS12 iteration: 5,000,000 raw lines through SELECTLINES on the
  GPU (the kept store equals the oracle's, record for record: 4,110,587
  lines), LINOP1 over the whole store on the 30,000-point grid (FP64
  XLINES within 1e-13 of the oracle's and the exact contribution count;
  mixed within 5e-6 with the same count), and the full 30,000-frequency
  RT pass on the GPU's FP64 XLINES (accumulators and HNU(1) within 1e-12);
* C1, plain SYNTHE: 200,000 lines x 279,808 points x 72 layers, checked on
  8 layers spread over the atmosphere -- FP64 opacity within 1e-11 of the
  FP64 oracle with its exact count, mixed against the reference-typed
  oracle (FP32 Voigt: median 1e-6, max 2e-4) with its exact count, and
  mixed-vs-FP64 emergent flux within 1e-4 at every point;
* C4, the large run: all 50,000,000 lines over 601,987 points, two layers
  against the oracle (its per-depth loop split over lines across threads,
  same values, sums in chunk order): FP64 within 1e-11 with the FP64
  count; mixed against the FP64 oracle (median 2e-6, max 2e-4) with the
  reference-typed count; mixed-vs-FP64 emergent flux within 1e-4;
* C3, the batch: all 256 models over the 10,000,000-line list in four
  groups of 64, mixed-vs-FP64 emergent flux within 1e-4 for every model;
  the 4 largest models alone equal their batch results bit for bit; two
  models against the oracle on sampled layers.

The FP64 tolerances cover the device exp/pow differing from libm in the
last bit and the different summation order; counts are exact."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
THREADS = min(16, os.cpu_count() or 1)


def _rel(a, b, floor=1e-300):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


@pytest.fixture(scope="module")
def base():
    from atlas12_amd.state import SyntheState
    return SyntheState.load(os.path.join(GOLDEN, "synthe_plain_state.npz"))


# ------------------------------------------------------------------ C2
def test_c2_full_iteration_matches_oracle(gpu_ctx, coef, orc):
    from atlas12_amd import native
    from atlas12_amd import workloads as wl
    z = np.load(os.path.join(GOLDEN, "atlas_plain_state.npz"))
    w = np.asarray(z["waveset"])
    pred, tio = wl.atlas_iteration_lists(5_000_000, float(w[0]), float(w[-1]), seed=2,
                                         molecules=tuple(int(m) for m in z["mol_nelion"]))
    store = native.atlas_select_lines(gpu_ctx, z, pred, tio)
    assert np.array_equal(store, orc.atlas_select(z, pred, tio))
    assert len(store) == 4_110_587
    zz = {k: z[k] for k in z.files}
    zz["linedata"] = store
    orc.lib().or_omp_threads(THREADS)
    oxl, oev = orc.linop1(zz, threads=THREADS)
    d = native.atlas_state_from(z)
    native.atlas_load_lines(gpu_ctx, d, store)
    native.atlas_set_state(gpu_ctx, d)
    xl64, ev64 = native.atlas_linop1(gpu_ctx, native.PREC_FP64)
    assert ev64 == oev == 3_848_489_111
    assert _rel(xl64, oxl).max() <= 1e-13
    xlm, evm = native.atlas_linop1(gpu_ctx, native.PREC_MIXED)
    assert evm == oev
    assert _rel(xlm, oxl).max() <= 5e-6
    # the full 30,000-frequency RT pass on the GPU's FP64 XLINES
    gpu_ctx.set_operators(*coef)
    kap = wl.kapp_standin(z, len(w))
    native.atlas_set_kapp(gpu_ctx, *kap)
    acc, h1 = native.atlas_rt_pass(gpu_ctx, xl64)
    for k, a in zip(("acont", "sigmac", "scont", "sline", "aline", "sigmal"), kap):
        zz[k] = a
    oacc, oh1 = orc.rt_pass(zz, xl64, *coef)
    for k in native.ACC_NAMES:
        scale = np.max(np.abs(oacc[k]))
        assert np.max(np.abs(acc[k] - oacc[k])) <= 1e-12 * scale, k
    assert _rel(acc["pradk0"], oacc["pradk0"]) <= 1e-12
    assert _rel(h1, oh1).max() <= 1e-12


# ------------------------------------------------------------------ C1
def test_c1_full_synthe_matches_oracle(gpu_ctx, coef, orc, base):
    from atlas12_amd import native
    from atlas12_amd import workloads as wl
    grid = wl.Grid.make(400.0, 700.0, 500000.0)
    assert grid.length == base.length == 279_808
    lines = wl.synthe_plain_lines(grid, 200_000, seed=1)
    gpu_ctx.set_operators(*coef)
    idx = np.unique(np.linspace(0, base.nrhox - 1, 8).round().astype(int))
    res = {}
    for prec in (native.PREC_FP64, native.PREC_MIXED):
        gpu_ctx.synthe_load(base, lines)
        gpu_ctx.synthe_opacity_run(prec)
        op = gpu_ctx.synthe_opacity_fetch()
        gpu_ctx.synthe_transfer_run()
        hl, _ = gpu_ctx.synthe_transfer_fetch()
        # the count over the sampled layers: the same layers alone (their
        # rows agree with the full run's to rounding -- a line's extent
        # class, which orders the sums, is its widest over the layers given)
        gpu_ctx.synthe_load(base.depths(idx), lines)
        gpu_ctx.synthe_opacity_run(prec)
        sub = gpu_ctx.synthe_opacity_fetch()
        assert _rel(sub, op[idx], 1e-30).max() <= (1e-12 if prec == native.PREC_FP64 else 1e-5)
        res[prec] = (op[idx], gpu_ctx.synthe_evals(), hl)
    o64, oev64 = orc.synthe_depths(base, lines, idx, "f64", THREADS)
    assert res[native.PREC_FP64][1] == oev64
    assert _rel(res[native.PREC_FP64][0], o64).max() <= 1e-11
    o32, oev32 = orc.synthe_depths(base, lines, idx, "ref", THREADS)
    assert res[native.PREC_MIXED][1] == oev32
    r = _rel(res[native.PREC_MIXED][0], o32, 1e-30)
    assert np.median(r) <= 1e-6 and r.max() <= 2e-4, (np.median(r), r.max())
    hm, hd = res[native.PREC_MIXED][2], res[native.PREC_FP64][2]
    assert _rel(hm, hd).max() <= 1e-4


# ------------------------------------------------------------------ C4
def test_c4_full_large_run(gpu_ctx, coef, orc, base):
    from atlas12_amd import native
    from atlas12_amd import workloads as wl
    grid = wl.Grid.make(300.0, 1000.0, 500000.0)
    st = wl.cool_dwarf_state(base, grid)
    lines = wl.large_synthe_lines(grid, 50_000_000)
    assert len(lines) == 50_000_000
    gpu_ctx.set_operators(*coef)
    gpu_ctx.synthe_set_shape(0)
    idx = [40, 71]
    flux, rows, counts = {}, {}, {}
    for prec in (native.PREC_MIXED, native.PREC_FP64):
        gpu_ctx.synthe_load(st, lines)
        gpu_ctx.synthe_opacity_run(prec)
        gpu_ctx.synthe_transfer_run()
        flux[prec] = gpu_ctx.synthe_transfer_fetch()[0]
        gpu_ctx.synthe_load(st.depths(idx), lines)
        gpu_ctx.synthe_opacity_run(prec)
        rows[prec] = gpu_ctx.synthe_opacity_fetch()
        counts[prec] = gpu_ctx.synthe_evals()
    assert _rel(flux[native.PREC_MIXED], flux[native.PREC_FP64]).max() <= 1e-4
    ev64 = ev32 = 0
    for i, j in enumerate(idx):
        o64, e64 = orc.synthe_depth_split(st, lines, j, "f64", THREADS)
        assert _rel(rows[native.PREC_FP64][i], o64).max() <= 1e-11, j
        # mixed against the FP64 values: at 50M lines the reference's own
        # plain FP32 sums drift by ~5e-5 (thousands of additions per point),
        # so the compensated mixed path is checked against FP64 instead
        r = _rel(rows[native.PREC_MIXED][i], o64, 1e-30)
        assert np.median(r) <= 2e-6 and r.max() <= 2e-4, (j, np.median(r), r.max())
        _, e32 = orc.synthe_depth_split(st, lines, j, "ref", THREADS)
        ev64 += e64
        ev32 += e32
    assert counts[native.PREC_FP64] == ev64
    assert counts[native.PREC_MIXED] == ev32


# ------------------------------------------------------------------ C3
@pytest.fixture(scope="module")
def batch_workload(base):
    from atlas12_amd import workloads as wl
    return wl.synthe_batch_workload(base, 256, 10_000_000)


@pytest.mark.parametrize("quarter", [0, 1, 2, 3])
def test_c3_full_batch(gpu_ctx, coef, orc, batch_workload, quarter):
    from atlas12_amd import native
    g, models, states, lines, sh, lo, hi = batch_workload
    sel = list(range(64 * quarter, 64 * (quarter + 1)))
    gpu_ctx.set_operators(*coef)
    gpu_ctx.synthe_set_shape(0)
    gpu_ctx.synthe_batch_load([states[k] for k in sel], lines, sh[sel], lo[sel], hi[sel])
    out = {}
    for prec in (native.PREC_MIXED, native.PREC_FP64):
        gpu_ctx.synthe_batch_run(prec, transfer=True, graph=prec == native.PREC_MIXED)
        out[prec] = [gpu_ctx.synthe_batch_fetch(i, opacity=False) for i in range(len(sel))]
    for i, k in enumerate(sel):
        hm, hd = out[native.PREC_MIXED][i][1], out[native.PREC_FP64][i][1]
        assert _rel(hm, hd).max() <= 1e-4, k
    if quarter == 0:
        # the 4 largest models of the draw, alone, equal their batch results
        big = sorted(range(256), key=lambda k: models[k].length)[-4:]
        gpu_ctx.synthe_batch_load([states[k] for k in big], lines, sh[big], lo[big], hi[big])
        gpu_ctx.synthe_batch_run(native.PREC_MIXED, transfer=True, graph=False)
        for i, k in enumerate(big):
            op_b, hl_b, _, ev_b = gpu_ctx.synthe_batch_fetch(i)
            sub = lines[lo[k]:hi[k]].copy()
            sub["nbuff"] -= sh[k]
            gpu_ctx.synthe_load(states[k], sub)
            gpu_ctx.synthe_opacity_run(native.PREC_MIXED)
            assert np.array_equal(gpu_ctx.synthe_opacity_fetch(), op_b), k
            gpu_ctx.synthe_transfer_run()
            assert np.array_equal(gpu_ctx.synthe_transfer_fetch()[0], hl_b), k
    if quarter == 1:
        # two models of this group against the oracle on four layers each
        for k in (sel[3], sel[40]):
            sub = lines[lo[k]:hi[k]].copy()
            sub["nbuff"] -= sh[k]
            st = states[k]
            idx = np.unique(np.linspace(0, st.nrhox - 1, 4).round().astype(int))
            gpu_ctx.synthe_load(st.depths(idx), sub)
            gpu_ctx.synthe_opacity_run(native.PREC_FP64)
            op = gpu_ctx.synthe_opacity_fetch()
            o64, oev = orc.synthe_depths(st, sub, idx, "f64", THREADS)
            assert gpu_ctx.synthe_evals() == oev, k
            assert _rel(op, o64).max() <= 1e-11, k
