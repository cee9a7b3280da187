#!/usr/bin/env python3
"""Benchmark of the SYNTHE/ATLAS12 hot path on one MI355X.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]

A step is one line-opacity build of the plain-SYNTHE workload (72 layers,
200,000 seeded atomic lines over 400-700 nm at R = 500,000, 279,808
wavelength points) with inputs resident in HBM, through the library's
staged entry point a12_synthe_opacity_run.  value = line-opacity evals/s of
the mixed-precision path (exact accumulated-contribution count / time).
Also reported: the FP64 GPU path on the same step, the emergent-flux error
of mixed against FP64 after the GPU transfer, the transfer throughput, the
roofline of the dominant kernel (HIP events on the library's stream) and a
CPU baseline (the oracle on the host cores, a bounded depth sample).
Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = ("line-opacity evals/s + s/ATLAS12 iter, mixed vs FP64 GPU at ≤1e-4 rel flux")
PEAK_FP32_TFLOPS = 157.3   # MI355X, published vector/matrix FP32
PEAK_FP64_TFLOPS = 78.6
FLOP_CORE = 247 + 1        # Weideman Voigt eval + accumulation
FLOP_WING = 3 + 1          # x_wing / n^2 + accumulation


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-depths", type=int, default=24)
    ap.add_argument("--profile-only", action="store_true",
                    help="warmup+steps mixed builds, one FP64 build, one transfer and a "
                         "device copy of --calib-mib MiB (for rocprofv3 runs)")
    ap.add_argument("--calib-mib", type=int, default=256,
                    help="size of the plain device-to-device copy used to scale HBM counters")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(st, lines, ndepth):
    """The oracle (C restatement of the reference loop) on the host cores:
    a bounded sample of ndepth of the 72 depths spread over the atmosphere,
    all lines and all wavelength points, one depth per OpenMP thread."""
    import ctypes

    from oracle import oracle as orc
    orc.build()
    L = orc.lib()
    threads = L.or_omp_threads(min(os.cpu_count() or 1, 16))
    a, keep = orc._atm(st)
    idx = np.unique(np.linspace(0, st.nrhox - 1, ndepth).round().astype(np.int32))
    out = np.zeros((st.nrhox, st.length), np.float32)
    lines = np.ascontiguousarray(lines)
    e = ctypes.c_int64(0)
    t0 = time.perf_counter()
    L.or_synthe_line_opacity_omp_list(ctypes.byref(a), lines.ctypes.data_as(ctypes.c_void_p),
                                      ctypes.c_int(len(lines)), ctypes.c_int(0),
                                      idx.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(len(idx)),
                                      orc._p(out, orc.P_F), None, ctypes.byref(e))
    dt = time.perf_counter() - t0
    return {"value": e.value / dt, "unit": "evals/s", "cores": int(min(threads, len(idx))),
            "kind": "port",
            "sample": f"{len(idx)} of {st.nrhox} depths (one per thread), all {len(lines)} lines "
                      f"and {st.length} points: {e.value} evals in {dt:.1f} s"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the path would shard by wavelength tiles across GPUs, which is not built.",
              file=sys.stderr)
        sys.exit(2)

    from atlas12_amd import native
    from atlas12_amd import workloads as wl
    from atlas12_amd.state import SyntheState, load_coef

    g = os.path.join(REPO, "tests", "golden")
    st = SyntheState.load(os.path.join(g, "synthe_plain_state.npz"))
    grid = wl.Grid.make(400.0, 700.0, 500000.0)
    assert grid.length == st.length
    t0 = time.perf_counter()
    lines = wl.synthe_plain_lines(grid, 200_000, seed=1)
    log(f"workload: {len(lines)} lines, {st.nrhox} layers, {st.length} points "
        f"({time.perf_counter() - t0:.1f}s to generate)")
    cj, ch = load_coef(os.path.join(g, "josh_coef.npz"))

    import torch
    torch.cuda.init()

    ctx = native.Context(0)
    ctx.set_operators(cj, ch)
    ctx.set_profiling(True)
    ctx.synthe_load(st, lines)

    def run(prec, k):
        stages = np.zeros(5)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(k):
            ctx.synthe_opacity_run(prec)
            ms = ctx.stage_ms()  # waits on this run's events
            stages[:4] += ms[:4]
        torch.cuda.synchronize()
        return time.perf_counter() - t, stages / max(k, 1)

    if args.profile_only:
        run(native.PREC_MIXED, args.warmup + args.steps)
        run(native.PREC_FP64, 1)
        ctx.synthe_transfer_run()
        ctx.synthe_transfer_fetch()
        # known-size copy: FETCH_SIZE/WRITE_SIZE of this kernel calibrate the counters
        nel = args.calib_mib * (1 << 20) // 4
        src = torch.ones(nel, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"profile-only run done (calibration copy {args.calib_mib} MiB)")
        return

    log("warmup")
    run(native.PREC_MIXED, args.warmup)
    log("timed mixed")
    t_mix, st_mix = run(native.PREC_MIXED, args.steps)
    evals, core = ctx.eval_split()
    ctx.synthe_transfer_run()
    hl_mix, hc_mix = ctx.synthe_transfer_fetch()
    t_tr = ctx.stage_ms()[4]
    log("timed fp64")
    run(native.PREC_FP64, 1)
    t_64, st_64 = run(native.PREC_FP64, max(1, args.steps // 2))
    k64 = max(1, args.steps // 2)
    evals64, core64 = ctx.eval_split()
    ctx.synthe_transfer_run()
    hl_64, hc_64 = ctx.synthe_transfer_fetch()
    flux_err = float(np.max(np.abs(hl_mix - hl_64) / np.abs(hl_64)))
    ctx.close()

    ms_step = 1e3 * t_mix / args.steps
    value = evals * args.steps / t_mix
    value64 = evals64 * k64 / t_64
    wing = evals - core
    flops = FLOP_CORE * core + FLOP_WING * wing
    tile_ms = float(st_mix[3])
    achieved = flops / (tile_ms * 1e-3) / 1e12
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded lines; atmosphere state from the reference's own set-up "
                "on a seeded deck)",
        "config": {"workload": "synthe_plain: plain-SYNTHE line-opacity build, 72 layers, "
                               "200,000 atomic lines, 400-700 nm at R=500,000 (279,808 points)",
                   "layers": st.nrhox, "lines": int(len(lines)), "points": st.length,
                   "evals_per_step": int(evals), "core_evals": int(core),
                   "wing_evals": int(wing)},
        "fp64_gpu": {"value": value64, "unit": "evals/s", "ms_per_step": 1e3 * t_64 / k64,
                     "evals_per_step": int(evals64)},
        "mixed_over_fp64": value / value64,
        "flux_max_rel_err_mixed_vs_fp64": flux_err,
        "transfer": {"points_per_s": st.length / (t_tr * 1e-3) if t_tr > 0 else None,
                     "ms": t_tr, "josh_calls_per_s": 2 * st.length / (t_tr * 1e-3) if t_tr > 0 else None},
        "stage_ms_mixed": {"prep_continuum": st_mix[0], "line_setup": st_mix[1],
                           "partition": st_mix[2], "tile_gather": st_mix[3]},
        "stage_ms_fp64": {"prep_continuum": st_64[0], "line_setup": st_64[1],
                          "partition": st_64[2], "tile_gather": st_64[3]},
        "roofline": {"bound": "valu", "kernel": "k_tile<mixed>", "achieved": achieved,
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                     "flops_per_launch": flops, "avg_launch_ms": tile_ms},
    }
    if not args.no_cpu:
        log("cpu baseline")
        result["cpu_baseline"] = cpu_baseline(st, lines, args.cpu_depths)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
