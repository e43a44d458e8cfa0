"""Benchmark: cell-updates/s of the FP64 finite-volume update on one MI355X.

Default workload: 256^3 periodic single-material gas (rho = p = 1 plus 16 seeded Fourier modes of
amplitude 1e-2 and a Gaussian pressure pulse), gamma = 1.4, MUSCL + van Albada, HLLC, RK3,
CFL 0.4, FP64 -- the plain GPU drop-in configuration. A "step" is one full time step: three
fused stage launches (reconstruct + flux + accumulate + stage combine), the fused CFL reduction
and the device-side dt update.

--workload shock_bubble: the 384x192x192 water/air shock-bubble (exact two-material Riemann
solves at the interface faces, level-set advection, reinitialisation every 5 steps, material-ID
and state update of swept cells), RK2, CFL 0.5, FP64; a step includes all of it.

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
# two materials + level set, RK2: 200 B of conserved state, the int8 ID read by both stages (2 B),
# the level-set RK2 (phi read/write per stage: 8+8 and 8+8+8 B), the material-ID update (phi_n,
# phi_n+1, ID read/write: 18 B) -- 260 B; reinitialisation adds 88 B per iteration every 5 steps
BYTES_PER_CELL_UPDATE_MP = 260.0
BYTES_PER_REINIT_ITER = 88.0


def cpu_baseline(settings, V0, ID0, nsteps, phi0=None):
    """Time the C oracle (OpenMP over z) on a bounded sample of the same workload."""
    from tests.cases import oracle_config
    from oracle import pyoracle as po
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    cfg = oracle_config(settings, vanalbada_pow=True, nthreads=threads)
    o = po.Oracle(cfg)
    V, ID = o.ghosted(V0, ID0)
    phi = o.ghosted_phi(phi0) if phi0 is not None else None
    t0 = time.perf_counter()
    st, t, n = o.run(V, ID, 1e300, nsteps, phi=phi)
    el = time.perf_counter() - t0
    cells = V0.shape[0] * V0.shape[1] * V0.shape[2]
    o.close()
    return {"value": cells * n / el, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{n} step(s) of the same {V0.shape[2]}x{V0.shape[1]}x{V0.shape[0]} workload, "
                      f"{el:.1f} s wall, C oracle -O3 with {threads} OpenMP threads"}


def mixed_form(M, W, settings, V0, ID0, phi0, args, V64, value64):
    """The mixed-precision form on the same workload and step count: throughput and the relative
    L1 error per field of its final state against the FP64 run's."""
    cells = V0.shape[0] * V0.shape[1] * V0.shape[2]
    s = M.Solver(W.config_from(settings, precision=M.PREC_MIXED))
    s.upload(V0, ID0)
    if phi0 is not None:
        s.upload_levelset(phi0)
    if args.warmup:
        s.run(1e300, args.warmup, sync_every=0)
    s.L.m2c_synchronize(s.h)
    t0 = time.perf_counter()
    st = s.run(1e300, args.steps, sync_every=0)
    s.L.m2c_synchronize(s.h)
    el = time.perf_counter() - t0
    V32, _ = s.download()
    ms_stage, _ = s.time_steps(max(5, min(args.steps, 20)))
    s.close()
    names = ["rho", "u", "v", "w", "p"]
    l1 = {}
    for f in range(5):
        den = float(np.abs(V64[..., f]).sum())
        l1[names[f]] = float(np.abs(V32[..., f] - V64[..., f]).sum() / den) if den > 0 else 0.0
    v = cells * args.steps / el
    print(f"bench: mixed form {args.steps} steps in {el:.3f} s, L1(rho) {l1['rho']:.2e}", file=sys.stderr, flush=True)
    return {"value": v, "unit": "cell-updates/s", "speedup_vs_fp64": v / value64, "status": st.error,
            "l1_rel_err_vs_fp64": l1, "l1_bar": 1e-5, "avg_stage_launch_ms": ms_stage,
            "form": "fp32 reconstruction/limiters/flux; fp64 fetch+C2P, accumulation, RK combination, CFL, "
                    "exact Riemann"}


def bench_batched(M, W, args):
    """64 independent Riemann-type instances (edges 48..160 per axis), mixed precision, RK2:
    one persistent launch per RK stage, a hipGraph per step, autotuned launch shape. Compared with
    the same instances advanced one context at a time."""
    insts = W.batched_riemann(n=64)
    cells = sum(V.shape[0] * V.shape[1] * V.shape[2] for _, V, _ in insts)
    cfgs = [W.config_from(s, precision=M.PREC_MIXED) for s, _, _ in insts]
    b = M.Batch(cfgs)
    for i, (s, V0, ID0) in enumerate(insts):
        b.upload(i, V0, ID0)
    print(f"bench: batch of {len(insts)} instances, {cells} cells uploaded", file=sys.stderr, flush=True)
    tune = b.autotune(trial_steps=2)
    print(f"bench: autotune kchunk {tune.kchunk} blocks {tune.blocks} ({tune.ms_per_step:.2f} ms/step)",
          file=sys.stderr, flush=True)
    if args.warmup:
        b.run(args.warmup)
    t0 = time.perf_counter()
    ms = b.run(args.steps)
    el = time.perf_counter() - t0
    errs = [b.status(i).error for i in range(len(insts))]
    b.close()
    value = cells * args.steps / el
    # the same instances one context at a time (a bounded sample of 8 instances)
    nsolo, solo_steps = 8, 10
    tsolo, csolo = 0.0, 0
    for s, V0, ID0 in insts[:nsolo]:
        sv = M.Solver(W.config_from(s, precision=M.PREC_MIXED))
        sv.upload(V0, ID0)
        sv.run(1e300, 2, sync_every=0)
        t1 = time.perf_counter()
        sv.run(1e300, solo_steps, sync_every=0)
        tsolo += time.perf_counter() - t1
        csolo += V0.shape[0] * V0.shape[1] * V0.shape[2] * solo_steps
        sv.close()
    line = {"metric": "cell-updates/s on 1 GPU: plain FP64 vs mixed form at L1 err<=1e-5; % HBM peak",
            "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": el / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64+f32", "data": "synthetic",
            "config": {"workload": "batched_riemann_64x[48..160]^3_rk2_hllc_vanalbada_mixed", "cells": cells,
                       "instances": len(insts), "errors": int(sum(1 for e in errs if e)),
                       "autotune": {"kchunk": tune.kchunk, "blocks": tune.blocks,
                                    "ms_per_step": tune.ms_per_step, "candidates": tune.candidates},
                       "graph_ms_per_step": ms / args.steps,
                       "one_context_at_a_time": {"value": csolo / tsolo, "instances": nsolo, "steps": solo_steps}}}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=0, help="oracle steps for the CPU baseline (0: 8 for the periodic case, 10 for the shock-bubble)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", default="periodic_modes", choices=["periodic_modes", "shock_bubble", "batched_riemann"])
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision comparison run")
    args = ap.parse_args()
    if args.cpu_steps <= 0:
        args.cpu_steps = 10 if args.workload == "shock_bubble" else 8
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); the multi-GPU "
              "decomposition is a design, see the sharding section of the design notes", file=sys.stderr)
        sys.exit(2)

    import m2c_amd as M
    from m2c_amd import workloads as W

    if args.workload == "batched_riemann":
        return bench_batched(M, W, args)
    mp = args.workload == "shock_bubble"
    if mp:
        settings, V0, ID0, phi0 = W.shock_bubble()
    else:
        settings, V0, ID0 = W.periodic_modes(n=args.n)
    cells = V0.shape[0] * V0.shape[1] * V0.shape[2]
    s = M.Solver(W.config_from(settings))
    s.upload(V0, ID0)
    if mp:
        s.upload_levelset(phi0)
    print(f"bench: {M.version()} {args.workload} {V0.shape[2]}x{V0.shape[1]}x{V0.shape[0]} uploaded, "
          f"warmup {args.warmup}", file=sys.stderr, flush=True)
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
    reinit_its = st.reinit_iterations
    V64, ID64 = s.download()
    print(f"bench: {args.steps} steps in {el:.3f} s (reinit iterations {reinit_its})", file=sys.stderr, flush=True)
    # stage-kernel duration with HIP events on the library's stream
    nprof = max(5, min(args.steps, 50))
    ms_stage, ms_step_ev = s.time_steps(nprof)
    copy_gbs, copy_ms = M.stream_copy_GBs(s, 1 << 30, 20)
    s.close()
    nstage = 2.0 if mp else 3.0
    bytes_stage = (200.0 if mp else BYTES_PER_CELL_UPDATE) / nstage
    achieved = bytes_stage * cells / (ms_stage * 1e-3) / 1e9
    total_steps = args.warmup + args.steps
    bpcu = BYTES_PER_CELL_UPDATE_MP + BYTES_PER_REINIT_ITER * reinit_its / max(total_steps, 1) if mp \
        else BYTES_PER_CELL_UPDATE
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None,
            "kernel": "m2c::stage_kernel (fused MUSCL+HLLC+RK stage)", "avg_launch_ms": ms_stage}
    line = {"metric": "cell-updates/s on 1 GPU: plain FP64 vs mixed form at L1 err<=1e-5; % HBM peak",
            "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": el / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": (f"shock_bubble_{V0.shape[2]}x{V0.shape[1]}x{V0.shape[0]}_rk2_hllc_vanalbada_"
                                    "exact_riemann_levelset_reinit5_fp64") if mp
                       else f"periodic_modes_{args.n}^3_rk3_hllc_vanalbada_fp64", "cells": cells,
                       "precision": "fp64", "bytes_per_cell_update": bpcu,
                       "hbm_frac_whole_step": value * bpcu / 1e9 / HBM_PEAK_GBS,
                       "measured_copy_GBs": copy_gbs, "copy_kernel_ms_per_GiB": copy_ms},
            "roofline": roof}
    if not args.no_mixed:
        line["mixed"] = mixed_form(M, W, settings, V0, ID0, phi0 if mp else None, args, V64, value)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(settings, V0, ID0, args.cpu_steps, phi0 if mp else None)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
