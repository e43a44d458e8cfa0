"""Benchmark: cell-updates/s of the FP64 finite-volume update on one MI355X.

Workload: 256^3 periodic single-material gas (rho = p = 1 plus 16 seeded Fourier modes of
amplitude 1e-2 and a Gaussian pressure pulse), gamma = 1.4, MUSCL + van Albada, HLLC, RK3,
CFL 0.4, FP64 -- the plain GPU drop-in configuration. A "step" is one full time step: three
fused stage launches (reconstruct + flux + accumulate + stage combine), the fused CFL reduction
and the device-side dt update.

Prints one JSON line with the throughput, the roofline of the fused stage kernel (algorithmic
bytes per launch divided by its average duration, measured with HIP events on the library's own
stream) and a CPU baseline (the C oracle timed on the host's cores on a bounded sample).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per cell-update, FP64 single material RK3: stage 1 reads V_n and writes U_1
# (80 B), stages 2 and 3 read U_s and V_n and write their output (120 B each)
BYTES_PER_CELL_UPDATE = 320.0


def cpu_baseline(settings, V0, ID0, nsteps):
    """Time the C oracle (OpenMP over z) on a bounded sample of the same workload."""
    from tests.cases import oracle_config
    from oracle import pyoracle as po
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    cfg = oracle_config(settings, vanalbada_pow=True, nthreads=threads)
    o = po.Oracle(cfg)
    V, ID = o.ghosted(V0, ID0)
    t0 = time.perf_counter()
    st, t, n = o.run(V, ID, 1e300, nsteps)
    el = time.perf_counter() - t0
    cells = V0.shape[0] * V0.shape[1] * V0.shape[2]
    o.close()
    return {"value": cells * n / el, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{n} RK3 step(s) of the same {V0.shape[2]}x{V0.shape[1]}x{V0.shape[0]} workload, "
                      f"{el:.1f} s wall, C oracle -O3 with {threads} OpenMP threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); the multi-GPU "
              "decomposition is a design, see the sharding section of the design notes", file=sys.stderr)
        sys.exit(2)

    import m2c_amd as M
    from m2c_amd import workloads as W

    settings, V0, ID0 = W.periodic_modes(n=args.n)
    cells = V0.shape[0] * V0.shape[1] * V0.shape[2]
    s = M.Solver(W.config_from(settings))
    s.upload(V0, ID0)
    print(f"bench: {M.version()} {args.n}^3 uploaded, warmup {args.warmup}", file=sys.stderr, flush=True)
    if args.warmup:
        s.run(1e300, args.warmup, sync_every=0)
    # timed region: K CFL-adaptive steps, device synchronised on both sides
    s.L.m2c_synchronize(s.h)
    t0 = time.perf_counter()
    st = s.run(1e300, args.steps, sync_every=0)
    s.L.m2c_synchronize(s.h)
    el = time.perf_counter() - t0
    if st.error != 0:
        print(f"bench: device reported status {st.error}", file=sys.stderr)
        sys.exit(3)
    value = cells * args.steps / el
    print(f"bench: {args.steps} steps in {el:.3f} s", file=sys.stderr, flush=True)
    # stage-kernel duration with HIP events on the library's stream
    nprof = max(5, min(args.steps, 50))
    ms_stage, ms_step_ev = s.time_steps(nprof)
    copy_gbs, copy_ms = M.stream_copy_GBs(s, 1 << 30, 20)
    s.close()
    achieved = BYTES_PER_CELL_UPDATE / 3.0 * cells / (ms_stage * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None,
            "kernel": "m2c::stage_kernel (fused MUSCL+HLLC+RK stage)", "avg_launch_ms": ms_stage}
    line = {"metric": "cell-updates/s on 1 GPU: plain FP64 vs mixed form at L1 err<=1e-5; % HBM peak",
            "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": el / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"periodic_modes_{args.n}^3_rk3_hllc_vanalbada_fp64", "cells": cells,
                       "precision": "fp64", "hbm_frac_whole_step": value * BYTES_PER_CELL_UPDATE / 1e9 / HBM_PEAK_GBS,
                       "measured_copy_GBs": copy_gbs, "copy_kernel_ms_per_GiB": copy_ms},
            "roofline": roof}
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(settings, V0, ID0, args.cpu_steps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
