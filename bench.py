"""Benchmark: cell-updates/s of the finite-volume update on one MI355X.

Default workload: the 384x192x192 water/air shock-bubble (two stiffened gases, exact two-material
Riemann solves at the interface faces, level-set advection, reinitialisation every 5 steps,
material-ID and state update of swept cells), RK2, HLLC + MUSCL (van Albada), CFL 0.5, FP64 -- the
configuration the metric is quoted on. A "step" is one full time step: both fused RK stage launches
(reconstruct + flux + accumulate + stage combine), the interface Riemann pass, the level-set stages,
the CFL reduction, the end-of-step material-ID / state update and (every 5th step) the
reinitialisation. The timed region is --steps steps after --warmup steps from the initial state.
Beside it, in the same line:
* mixed, fp64_surrogate, mixed_surrogate: the same steps in the mixed-precision form and with the
  learned interface solver;
* full_run_300: every form over the full 300 steps from the initial state, the FP64 state at steps
  150 and 300 checked against the oracle's digest (tests/golden/shock_bubble_384x192x192_digest.json),
  and per-field relative L1 and L-infinity of the other forms against it at both steps;
* periodic_256: the 256^3 periodic single-material case (RK3), its stage-kernel roofline and the
  HBM traffic measured for it;
* cpu_baseline: the C oracle on the host's cores over a bounded sample (CPU model recorded).

--workload periodic_modes: the 256^3 periodic single-material gas (rho = p = 1 plus 16 seeded
Fourier modes of amplitude 1e-2 and a Gaussian pressure pulse), gamma = 1.4, MUSCL + van Albada,
HLLC, RK3, CFL 0.4, FP64, as its own line.

--workload batched_riemann: 64 independent instances (edges 48..160), every 4th one two-material
(level set, exact interface Riemann solves), mixed precision, one persistent launch per RK stage,
hipGraph step replay, autotuned launch shape.

--workload learned: the learned interface solver's kernels (inference roofline against the bf16
MFMA peak, training steps) and its per-problem cost against the exact solver.

--workload viscous: the 256^3 periodic case with viscous and heat-diffusion terms (constant
viscosity and diffusivity), RK3, FP64: per stage a face pass (velocity gradients, stresses, heat
flux per face) before the fused stage kernel, which adds the face terms after the advective sum.

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
FP64_PEAK_TFLOPS = 78.6
# FP64 flops the fused stage kernel issues per cell and stage on the 256^3 periodic HLLC workload:
# 64 x SQ_INSTS_VALU_FLOPS_FP64 per launch / cells, from a counter pass over the three RK3 stage
# variants (full-wave count: masked lanes included)
ISSUED_FP64_FLOP_PER_CELL_STAGE = 2640.0
# algorithmic bytes per cell-update, FP64 single material RK3: stage 1 reads V_n and writes U_1
# (80 B), stages 2 and 3 read U_s and V_n and write their output (120 B each)
BYTES_PER_CELL_UPDATE = 320.0
# two materials + level set, RK2: 200 B of conserved state, the int8 ID read by both stages (2 B),
# the level-set RK2 (phi read/write per stage: 8+8 and 8+8+8 B), the material-ID update (phi_n,
# phi_n+1, ID read/write: 18 B) -- 260 B; reinitialisation adds 88 B per iteration every 5 steps
BYTES_PER_CELL_UPDATE_MP = 260.0
BYTES_PER_REINIT_ITER = 88.0
# viscous + heat diffusion, per stage: the face pass reads the stage field (40 B/cell) and writes 3
# faces x (10 viscous + 1 heat) doubles (264 B/cell); the stage kernel reads them back, 6 doubles on
# each side of 3 faces per cell (288 B/cell)
BYTES_DIFF_PASS = 304.0
BYTES_DIFF_STAGE_READ = 288.0
# HBM bytes per launch of the fused stage kernel on the 256^3 periodic RK3 FP64 workload, from
# rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of `bench.py --steps 3 --warmup 1 --no-cpu
# --no-mixed` (FETCH scaled by 2.0, WRITE by 1.0 after a copy-kernel calibration; summary in
# profiles/stage_kernel_periodic256_rk3_fp64_traffic.md). Re-measure after changing the kernel.
STAGE_TRAFFIC_PERIODIC256_FP64 = 3.413e9
# The same for the two RK2 stage launches of the 384x192x192 shock-bubble FP64 run (mean of the two),
# from `bench.py --steps 3 --warmup 1 --no-cpu --no-mixed --no-full` under the same two counters
# (summary in profiles/stage_kernel_shock_bubble_384x192x192_fp64_traffic.md).
STAGE_TRAFFIC_SHOCK_BUBBLE_FP64 = 2.935e9


def host_cpu():
    """The host's CPU model, sockets and physical cores (lscpu), and the logical CPUs of this process."""
    info = {"model": None, "sockets": None, "cores_per_socket": None, "logical_cpus": os.cpu_count()}
    try:
        import subprocess
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for ln in txt.splitlines():
            k, _, v = ln.partition(":")
            k, v = k.strip(), v.strip()
            if k == "Model name":
                info["model"] = v
            elif k == "Socket(s)":
                info["sockets"] = int(v)
            elif k == "Core(s) per socket":
                info["cores_per_socket"] = int(v)
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        pass
    return info


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
            "sample": f"{n} step(s) of the same {V0.shape[2]}x{V0.shape[1]}x{V0.shape[0]} workload from its initial "
                      f"state, {el:.1f} s wall, C oracle (gcc -O3 -march=x86-64-v3 -ffp-contract=off) with {threads} "
                      f"OpenMP threads over z",
            "host": host_cpu()}


def l1_fields(Vx, Vref):
    """Relative L1 difference per field; velocity components are measured against the L1 norm of
    the whole velocity (a component that is zero in the reference would otherwise divide by ~0)."""
    names = ["rho", "u", "v", "w", "p"]
    vel = float(np.abs(Vref[..., 1:4]).sum())
    out = {}
    for f in range(5):
        den = vel if 1 <= f <= 3 else float(np.abs(Vref[..., f]).sum())
        out[names[f]] = float(np.abs(Vx[..., f] - Vref[..., f]).sum() / den) if den > 0 else 0.0
    return out


def linf_fields(Vx, Vref):
    """Relative L-infinity difference per field (velocities against the whole velocity's maximum)."""
    names = ["rho", "u", "v", "w", "p"]
    vel = float(np.abs(Vref[..., 1:4]).max())
    out = {}
    for f in range(5):
        den = vel if 1 <= f <= 3 else float(np.abs(Vref[..., f]).max())
        out[names[f]] = float(np.abs(Vx[..., f] - Vref[..., f]).max() / den) if den > 0 else 0.0
    return out


def state_digest(V, ID, phi):
    """SHA-256 of the state as the oracle's fixture writes it (negative zeros folded into +0)."""
    import hashlib
    return {"sha256_V": hashlib.sha256((np.ascontiguousarray(V, dtype=np.float64) + 0.0).tobytes()).hexdigest(),
            "sha256_ID": hashlib.sha256(np.ascontiguousarray(ID, dtype=np.int32).tobytes()).hexdigest(),
            "sha256_phi": hashlib.sha256((np.ascontiguousarray(phi, dtype=np.float64) + 0.0).tobytes()).hexdigest()}


DIGEST = os.path.join(ROOT, "tests", "golden", "shock_bubble_384x192x192_digest.json")


def full_run(M, W, settings, V0, ID0, phi0, nsteps, checkpoints, forms=True):
    """Every form over the whole run from the initial state: throughput (run segments only, state
    downloads at the checkpoints excluded), the FP64 state at each checkpoint against the oracle's
    digest, and per-field relative L1 / L-infinity of the other forms against the FP64 state (which
    stands for the oracle's where the digests match)."""
    cells = V0.shape[0] * V0.shape[1] * V0.shape[2]
    fixture = None
    if os.path.exists(DIGEST) and (settings["nx"], settings["ny"], settings["nz"]) == (384, 192, 192):
        with open(DIGEST) as f:
            fixture = json.load(f)
    out = {"steps": nsteps, "checkpoints": list(checkpoints)}

    def one(precision, net):
        s = M.Solver(W.config_from(settings, precision=precision))
        if net is not None:
            s.set_surrogate(net, newton_max=3)
        s.upload(V0, ID0)
        s.upload_levelset(phi0)
        el, done, states = 0.0, 0, {}
        for chk in checkpoints:
            s.L.m2c_synchronize(s.h)
            t0 = time.perf_counter()
            s.run(1e300, chk - done, sync_every=0)
            s.L.m2c_synchronize(s.h)
            el += time.perf_counter() - t0
            done = chk
            V, ID = s.download()
            states[chk] = (V, ID, s.download_levelset())
        st = s.status()
        s.close()
        return {"value": cells * nsteps / el, "ms_per_step": el / nsteps * 1e3, "status": st.error,
                "reinit_iterations": st.reinit_iterations, "unresolved_cells": st.unresolved_cells,
                "riemann_errors": st.riemann_errors}, states, st

    res64, ref, st64 = one(M.PREC_FP64, None)
    res64["t_final"] = st64.t
    if fixture is not None:
        chk = {}
        for c in checkpoints:
            d = state_digest(*ref[c])
            fx = fixture["checkpoints"].get(str(c))
            chk[str(c)] = {"bit_identical_to_oracle": bool(fx is not None and all(d[k] == fx[k] for k in d)),
                           "t_equal": bool(fx is not None and fx["t"] == (st64.t if c == checkpoints[-1] else fx["t"]))}
        res64["vs_oracle_digest"] = chk
        if "pow_mode" in fixture:
            res64["oracle_pow_vs_dqdq"] = fixture["pow_mode"]
    print(f"bench: full run fp64 {nsteps} steps, {res64['ms_per_step']:.2f} ms/step, digests "
          f"{res64.get('vs_oracle_digest')}", file=sys.stderr, flush=True)
    out["fp64"] = res64
    if forms:
        net = M.load_surrogate() if os.path.exists(M.SURROGATE_PARAMS) else None
        for tag, prec, nn in (("mixed", M.PREC_MIXED, None), ("fp64_surrogate", M.PREC_FP64, net),
                              ("mixed_surrogate", M.PREC_MIXED, net)):
            if tag.endswith("surrogate") and net is None:
                continue
            r, states, st = one(prec, nn)
            r["speedup_vs_fp64"] = r["value"] / res64["value"]
            for c in checkpoints:
                r[f"step_{c}"] = {"l1_rel": l1_fields(states[c][0], ref[c][0]),
                                  "linf_rel": linf_fields(states[c][0], ref[c][0]),
                                  "id_cells_differing": int((states[c][1] != ref[c][1]).sum())}
            if nn is not None:
                r["network_solves"], r["exact_fallbacks"] = st.surrogate_accepted, st.surrogate_fallbacks
            print(f"bench: full run {tag}: {r['ms_per_step']:.2f} ms/step, L1(rho) at {checkpoints[-1]} "
                  f"{r[f'step_{checkpoints[-1]}']['l1_rel']['rho']:.2e}", file=sys.stderr, flush=True)
            out[tag] = r
            del states
        if net is not None:
            net.close()
        out["l1_bar"] = 1e-5
    return out


def variant_form(M, W, settings, V0, ID0, phi0, args, V64, value64, precision, net=None):
    """The same workload and step count in another form (mixed precision and/or the learned interface
    solver): throughput and the relative L1 / L-infinity difference per field of its final state
    against the FP64 run's."""
    cells = V0.shape[0] * V0.shape[1] * V0.shape[2]
    s = M.Solver(W.config_from(settings, precision=precision))
    if net is not None:
        s.set_surrogate(net, newton_max=3)
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
    Vx, _ = s.download()
    ms_stage, _ = s.time_steps(max(5, min(args.steps, 20)))
    s.close()
    l1 = l1_fields(Vx, V64)
    v = cells * args.steps / el
    tag = ("mixed" if precision == M.PREC_MIXED else "fp64") + ("+surrogate" if net is not None else "")
    print(f"bench: {tag} form {args.steps} steps in {el:.3f} s, L1(rho) {l1['rho']:.2e}", file=sys.stderr, flush=True)
    out = {"value": v, "unit": "cell-updates/s", "ms_per_step": el / args.steps * 1e3, "speedup_vs_fp64": v / value64,
           "status": st.error, "l1_rel_err_vs_fp64": l1, "linf_rel_err_vs_fp64": linf_fields(Vx, V64), "l1_bar": 1e-5,
           "avg_stage_launch_ms": ms_stage}
    if precision == M.PREC_MIXED:
        out["form"] = ("fp32 reconstruction/limiters/flux; fp64 fetch+C2P, accumulation, RK combination, CFL, "
                       "exact Riemann")
    if net is not None:
        out["network_solves"] = st.surrogate_accepted
        out["exact_fallbacks"] = st.surrogate_fallbacks
    return out


def _run_batch(M, W, args, insts):
    """Autotune, warm up and time one batch; returns (cells/s, ms per step, graph ms per step,
    tuning, errors)."""
    cells = sum(x[1].shape[0] * x[1].shape[1] * x[1].shape[2] for x in insts)
    b = M.Batch([W.config_from(x[0], precision=M.PREC_MIXED) for x in insts])
    for i, x in enumerate(insts):
        b.upload(i, x[1], x[2])
        if len(x) == 4:
            b.upload_levelset(i, x[3])
    tune = b.autotune(trial_steps=2)
    print(f"bench: autotune kchunk {tune.kchunk} blocks {tune.blocks} tile {tune.tile_x}x{tune.tile_y} "
          f"({tune.ms_per_step:.2f} ms/step)", file=sys.stderr, flush=True)
    if args.warmup:
        b.run(args.warmup)
    t0 = time.perf_counter()
    ms = b.run(args.steps)
    el = time.perf_counter() - t0
    errs = [b.status(i).error for i in range(len(insts))]
    b.close()
    tuning = {"kchunk": tune.kchunk, "blocks": tune.blocks, "tile": [tune.tile_x, tune.tile_y],
              "ms_per_step": tune.ms_per_step, "candidates": tune.candidates}
    return cells * args.steps / el, el / args.steps * 1e3, ms / args.steps, tuning, int(sum(1 for e in errs if e))


def bench_batched(M, W, args):
    """64 independent Riemann-type instances (edges 48..160 per axis), every 4th one two-material
    (a second gas behind its split plane, level set reinitialised every 5 steps, exact Riemann solves
    at the interface faces), mixed precision, RK2: one persistent launch per RK stage, hipGraph step
    replay, autotuned tile shape and launch shape. Beside it: the same 64 instances with one gas
    each, and a sample of the instances advanced one context at a time."""
    insts = W.batched_riemann(n=64, two_every=4)
    cells = sum(x[1].shape[0] * x[1].shape[1] * x[1].shape[2] for x in insts)
    ntwo = sum(len(x) == 4 for x in insts)
    print(f"bench: batch of {len(insts)} instances ({ntwo} two-material), {cells} cells", file=sys.stderr,
          flush=True)
    value, ms_step, graph_ms, tuning, nerr = _run_batch(M, W, args, insts)
    single = W.batched_riemann(n=64)
    v1, ms1, g1, t1, e1 = _run_batch(M, W, args, single)
    # the same instances one context at a time (a bounded sample of 8 instances, two of them
    # two-material)
    nsolo, solo_steps = 8, 10
    tsolo, csolo = 0.0, 0
    for x in insts[:nsolo]:
        s, V0, ID0 = x[:3]
        sv = M.Solver(W.config_from(s, precision=M.PREC_MIXED))
        sv.upload(V0, ID0)
        if len(x) == 4:
            sv.upload_levelset(x[3])
        sv.run(1e300, 2, sync_every=0)
        t0 = time.perf_counter()
        sv.run(1e300, solo_steps, sync_every=0)
        tsolo += time.perf_counter() - t0
        csolo += V0.shape[0] * V0.shape[1] * V0.shape[2] * solo_steps
        sv.close()
    line = {"metric": "cell-updates/s on 1 GPU: plain FP64 vs mixed form at L1 err<=1e-5; % HBM peak",
            "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64+f32", "data": "synthetic",
            "config": {"workload": "batched_riemann_64x[48..160]^3_rk2_hllc_vanalbada_mixed", "cells": cells,
                       "instances": len(insts), "two_material_instances": ntwo, "errors": nerr,
                       "autotune": tuning, "graph_ms_per_step": graph_ms,
                       "single_material_batch": {"value": v1, "ms_per_step": ms1, "graph_ms_per_step": g1,
                                                 "autotune": t1, "errors": e1},
                       "one_context_at_a_time": {"value": csolo / tsolo, "instances": nsolo, "steps": solo_steps}}}
    print(json.dumps(line), flush=True)


MLP_FLOP_PER_SAMPLE = 2.0 * (10 * 256 + 3 * 256 * 256 + 256 * 3)   # forward, 399 872
BF16_MFMA_PEAK_TFLOPS = 2500.0


def bench_learned(M, W, args):
    """The learned interface solver: fused bf16 MFMA inference over 2^20 device-resident rows
    (roofline against the dense bf16 MFMA peak), training steps (batch 8192), and the per-problem
    cost of the checked network path against the exact solver on water/air interface problems."""
    net = M.load_surrogate()
    n = 1 << 20
    ms_inf = net.bench_infer(n, reps=max(args.steps // 10, 10))
    inf_tflops = MLP_FLOP_PER_SAMPLE * n / (ms_inf * 1e-3) / 1e12
    trainer = M.Mlp(seed=3)
    B = 8192
    ms_train = trainer.bench_train(B, steps=max(args.steps // 4, 20))
    trainer.close()
    train_tflops = 3.0 * MLP_FLOP_PER_SAMPLE * B / (ms_train * 1e-3) / 1e12
    print(f"bench: inference {ms_inf:.3f} ms per 2^20 rows ({inf_tflops:.0f} TFLOP/s), training step "
          f"{ms_train:.3f} ms (batch {B})", file=sys.stderr, flush=True)
    settings, _, _, _ = W.shock_bubble(scale=0.125)
    s = M.Solver(W.config_from(settings))
    nb = 1 << 16
    Vm, idm, Vp, idp = W.interface_problems(nb)
    dirs = np.zeros(nb, np.int32)
    s.riemann(dirs[:64], Vm[:64], idm[:64], Vp[:64], idp[:64])   # warm
    t0 = time.perf_counter()
    st_e = s.riemann(dirs, Vm, idm, Vp, idp)[0]
    t1 = time.perf_counter()
    st_s, _, _, _, _, path = s.surrogate_riemann(net, dirs, Vm, idm, Vp, idp, newton_max=3)
    t2 = time.perf_counter()
    # the problems the network accepts, through both paths (whole-batch times are set by the few
    # slowest exact solves, which both paths run)
    sel = np.nonzero(path == 1)[0]
    t3 = time.perf_counter()
    s.riemann(dirs[sel], Vm[sel], idm[sel], Vp[sel], idp[sel])
    t4 = time.perf_counter()
    s.surrogate_riemann(net, dirs[sel], Vm[sel], idm[sel], Vp[sel], idp[sel], newton_max=3)
    t5 = time.perf_counter()
    s.close()
    net.close()
    line = {"metric": "learned interface solver: inference samples/s (bf16 MFMA, FP32 accumulation)",
            "value": n / (ms_inf * 1e-3), "unit": "samples/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms_inf, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "bf16", "data": "synthetic",
            "config": {"workload": "mlp_10-256x4-3_silu_inference_2^20_rows", "rows": n,
                       "flop_per_row": MLP_FLOP_PER_SAMPLE},
            "roofline": {"bound": "mfma", "achieved": inf_tflops, "peak": BF16_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": inf_tflops / BF16_MFMA_PEAK_TFLOPS, "traffic": None,
                         "kernel": "m2c::k_mlp_infer (fused 5-layer MLP, 128-row tiles)", "avg_launch_ms": ms_inf},
            "training": {"batch": B, "ms_per_step": ms_train, "samples_per_s": B / (ms_train * 1e-3),
                         "tflops_3x_forward": train_tflops},
            "interface_problems": {"n": nb, "exact_batch_ms": 1e3 * (t1 - t0), "network_batch_ms": 1e3 * (t2 - t1),
                                   "accepted_only": {"n": int(sel.size), "exact_batch_ms": 1e3 * (t4 - t3),
                                                     "network_batch_ms": 1e3 * (t5 - t4)},
                                   "network_share": float((path == 1).mean()),
                                   "exact_failures": int((st_e != 0).sum()), "checked_failures": int((st_s != 0).sum()),
                                   "note": "host-inclusive pointwise batches (copies in and out)"}}
    if not args.no_cpu:
        from oracle import mlp_ref as R
        p = M.load_surrogate()
        params = p.params().astype(np.float64)
        p.close()
        X = np.random.default_rng(1).normal(size=(1 << 16, 10))
        t0 = time.perf_counter()
        R.forward(params, X, emulate=False)
        el = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": X.shape[0] / el, "unit": "samples/s",
                                "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                                "kind": "port", "sample": f"{X.shape[0]} rows through the numpy oracle's FP64 forward "
                                                          f"({el:.2f} s)"}
    print(json.dumps(line), flush=True)


def bench_grid(M, W, args, workload, variants=True):
    """The timed run of one grid workload: the JSON line without the CPU baseline, and what the
    baseline needs (settings, V0, ID0, phi0)."""
    mp = workload == "shock_bubble"
    visc = workload == "viscous"
    god = workload == "godunov"
    phi0 = None
    n = args.n
    if god and n == 256:
        n = 128   # an exact Riemann solve per face: 128^3 keeps the CPU sample near 20 s
    if mp:
        settings, V0, ID0, phi0 = W.shock_bubble()
    elif visc:
        settings, V0, ID0 = W.periodic_viscous(n=n)
    else:
        settings, V0, ID0 = W.periodic_modes(n=n)
        if god:
            settings = dict(settings, flux=M.FLUX_GODUNOV)
    cells = V0.shape[0] * V0.shape[1] * V0.shape[2]
    s = M.Solver(W.config_from(settings))
    s.upload(V0, ID0)
    if mp:
        s.upload_levelset(phi0)
    print(f"bench: {M.version()} {workload} {V0.shape[2]}x{V0.shape[1]}x{V0.shape[0]} uploaded, "
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
    ms_diff = s.diffusion_pass_ms()
    copy_gbs, copy_ms = M.stream_copy_GBs(s, 1 << 30, 20)
    s.close()
    nstage = 2.0 if mp else 3.0
    bytes_stage = (200.0 if mp else BYTES_PER_CELL_UPDATE) / nstage + (BYTES_DIFF_STAGE_READ if visc else 0.0)
    achieved = bytes_stage * cells / (ms_stage * 1e-3) / 1e9
    total_steps = args.warmup + args.steps
    bpcu = BYTES_PER_CELL_UPDATE_MP + BYTES_PER_REINIT_ITER * reinit_its / max(total_steps, 1) if mp \
        else BYTES_PER_CELL_UPDATE + (3 * (BYTES_DIFF_PASS + BYTES_DIFF_STAGE_READ) if visc else 0.0)
    traffic = None
    if mp:
        traffic = STAGE_TRAFFIC_SHOCK_BUBBLE_FP64
    elif not visc and not god and n == 256:
        traffic = STAGE_TRAFFIC_PERIODIC256_FP64
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "kernel": "m2c::stage_kernel (fused MUSCL+" + ("Godunov-flux read" if god else "HLLC") + "+RK stage)",
            "avg_launch_ms": ms_stage}
    if traffic is not None and not mp:
        # the bound the kernel actually meets: FP64 issue and latency, against the vector FP64 peak
        tf = ISSUED_FP64_FLOP_PER_CELL_STAGE * cells / (ms_stage * 1e-3) / 1e12
        roof["fp64_valu"] = {"achieved": tf, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                             "frac": tf / FP64_PEAK_TFLOPS,
                             "flop_per_cell_stage": ISSUED_FP64_FLOP_PER_CELL_STAGE,
                             "counted": "issued FP64 flops (64 x SQ_INSTS_VALU_FLOPS_FP64), masked lanes included"}
    line = {"metric": "cell-updates/s on 1 GPU: plain FP64 vs mixed form at L1 err<=1e-5; % HBM peak",
            "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": el / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": (f"shock_bubble_{V0.shape[2]}x{V0.shape[1]}x{V0.shape[0]}_rk2_hllc_vanalbada_"
                                    "exact_riemann_levelset_reinit5_fp64") if mp
                       else (f"periodic_viscous_{n}^3_rk3_hllc_vanalbada_mu0.01_kappa0.02_fp64" if visc
                             else f"periodic_modes_{n}^3_rk3_{'godunov' if god else 'hllc'}_vanalbada_fp64"),
                       "cells": cells,
                       "precision": "fp64", "bytes_per_cell_update": bpcu,
                       "hbm_frac_whole_step": value * bpcu / 1e9 / HBM_PEAK_GBS,
                       "measured_copy_GBs": copy_gbs, "copy_kernel_ms_per_GiB": copy_ms},
            "roofline": roof}
    if visc:
        # the face pass (three face kernels per stage) beside the stage kernel
        ach = BYTES_DIFF_PASS * cells / (ms_diff * 1e-3) / 1e9
        line["diffusion_pass"] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": ach / HBM_PEAK_GBS, "avg_pass_ms": ms_diff,
                                  "kernels": "k_diff_faces<0..2> (one thread per face)",
                                  "bytes_per_cell": BYTES_DIFF_PASS}
        line["ms_per_step_parts"] = {"stage_kernels": 3 * ms_stage, "diffusion_passes": 3 * ms_diff,
                                     "whole_step": ms_step_ev}
    if god:
        # the Godunov face pass: an exact Riemann solve per face, three face kernels per stage
        line["godunov_pass"] = {"bound": "latency (iterative exact solver per face)", "avg_pass_ms": ms_diff,
                                "faces_per_pass": 3 * cells, "faces_per_s": 3 * cells / (ms_diff * 1e-3),
                                "kernels": "k_godunov_faces<0..2> (one thread per face, grid-stride)"}
        line["ms_per_step_parts"] = {"stage_kernels": 3 * ms_stage, "godunov_passes": 3 * ms_diff,
                                     "whole_step": ms_step_ev}
    if variants and not god:
        line["mixed"] = variant_form(M, W, settings, V0, ID0, phi0 if mp else None, args, V64, value, M.PREC_MIXED)
        if mp and os.path.exists(M.SURROGATE_PARAMS):
            net = M.load_surrogate()
            line["fp64_surrogate"] = variant_form(M, W, settings, V0, ID0, phi0, args, V64, value, M.PREC_FP64, net)
            line["mixed_surrogate"] = variant_form(M, W, settings, V0, ID0, phi0, args, V64, value, M.PREC_MIXED,
                                                   net)
            net.close()
    return line, (settings, V0, ID0, phi0 if mp else None)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0,
                    help="timed steps (0: 20 for the shock-bubble, 200 for the periodic case, 100 batched)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=0, help="oracle steps for the CPU baseline (0: 8 for the periodic case, 10 for the shock-bubble)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", default="shock_bubble",
                    choices=["periodic_modes", "shock_bubble", "batched_riemann", "learned", "viscous", "godunov"])
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision / surrogate comparison runs")
    ap.add_argument("--no-full", action="store_true", help="shock-bubble: skip the full 300-step leg and the periodic case")
    args = ap.parse_args()
    if args.steps <= 0:
        args.steps = {"shock_bubble": 20, "batched_riemann": 100, "godunov": 20}.get(args.workload, 200)
    if args.cpu_steps <= 0:
        args.cpu_steps = {"shock_bubble": 10, "viscous": 4, "godunov": 2}.get(args.workload, 8)
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); the multi-GPU "
              "decomposition is a design, see the sharding section of the design notes", file=sys.stderr)
        sys.exit(2)

    import m2c_amd as M
    from m2c_amd import workloads as W

    if args.workload == "batched_riemann":
        return bench_batched(M, W, args)
    if args.workload == "learned":
        return bench_learned(M, W, args)
    line, (settings, V0, ID0, phi0) = bench_grid(M, W, args, args.workload, variants=not args.no_mixed)
    if args.workload == "shock_bubble" and not args.no_full:
        line["full_run_300"] = full_run(M, W, settings, V0, ID0, phi0, 300, (150, 300), forms=not args.no_mixed)
        pline, _ = bench_grid(M, W, args, "periodic_modes", variants=False)
        line["periodic_256"] = {k: pline[k] for k in ("value", "unit", "ms_per_step", "config", "roofline")}
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(settings, V0, ID0, args.cpu_steps, phi0)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
