#!/usr/bin/env python3
"""Benchmark of the HIP Flow Lenia step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
    python bench.py --population [--pop-steps K]
    python bench.py --grad [--grad-steps T] [--grad-every k]

Workload: the 2D multi-species configuration, 1024 x 1024, 3 channels, 9 kernels
(every source->target pair, 1-3 rings each), kernel_radius 15, dt 0.2, synthetic
parameters and initial blobs (autoverse_amd/workloads.py).  A step is one full
Flow Lenia time step over the device-resident state (FFT convolution + growth,
Sobel, flow, reintegration); `value` is cell-steps/s = W*H*steps / seconds.

Prints ONE JSON line.  `roofline` is computed from per-kernel HIP events recorded
on the library's own stream during a second pass over the same K steps, from the
same start state (the timed pass runs from hipGraphs without events); `host_boundary` times hp_step and hp_run on host
buffers (PCIe included, never `value`); `cpu_baseline` times the CPU oracle (oracle/, a
C restatement of the reference CPU propagator with its thread decomposition) on
a bounded number of steps of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
PROF_REPEATS = 10


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-oracle baseline and the FP64 error")
    ap.add_argument("--no-population", action="store_true",
                    help="skip the 256-member population object (batched forward+adjoint vs per-instance)")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the plan autotuner (run the built-in default plan)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU baseline sample length")
    ap.add_argument("--plan", type=str, default=None,
                    help='fix plan fields instead of tuning, as JSON, e.g. \'{"stencil_tile": 2}\'')
    ap.add_argument("--population", action="store_true",
                    help="population mode: 256 uneven members stepped as one batch vs one lone "
                         "propagator per member (prints its own JSON line)")
    ap.add_argument("--pop-steps", type=int, default=200)
    ap.add_argument("--no-lone", action="store_true",
                    help="population mode: skip the per-instance comparison pass (for profiling the batch)")
    ap.add_argument("--grad", action="store_true",
                    help="differentiation mode: 32 members of 512^2 x 2 channels, forward tape of "
                         "--grad-steps steps with a checkpoint every --grad-every, then the backward pass")
    ap.add_argument("--grad-steps", type=int, default=64)
    ap.add_argument("--grad-every", type=int, default=8)
    ap.add_argument("--sphere3d", action="store_true",
                    help="3D mode: 128^3, one channel, one shell kernel (prints its own JSON line)")
    ap.add_argument("--embedded", action="store_true",
                    help="embedded-parameter mode: multispecies --emb-size^2 x 3 with embedding enabled "
                         "(prints its own JSON line)")
    ap.add_argument("--emb-size", type=int, default=1024)
    ap.add_argument("--emb-linear", action="store_true", help="embedded mode: linear mixing instead of softmax")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain device copy of 256 MiB ten times (calibrates rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE against a kernel of known traffic)")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the step is single-GPU by design",
              file=sys.stderr)
        return 2
    if args.population:
        return population_bench(args)
    if args.grad:
        return grad_bench(args)
    if args.sphere3d:
        return sphere3d_bench(args)
    if args.embedded:
        return embedded_bench(args)

    import numpy as np

    import autoverse_amd as av
    from autoverse_amd import workloads

    cfg, seed = workloads.multispecies()
    N = cfg.width * cfg.height
    state = av.SimulationState.from_seed(seed, cfg)
    prop = av.HipPropagator(cfg)
    prop.set_state(state)
    prop.advance(args.warmup)
    _, mass0 = prop.mass_f64()
    tune = None
    if args.plan:
        prop.set_plan(json.loads(args.plan))
    elif not args.no_tune:  # untimed: pick the fastest plan on this device, state restored after
        best, table = prop.autotune(steps=20)
        tune = {"plan": best, "candidates": len(table),
                "default_us_per_step": table[0][1] if table else None,
                "best_us_per_step": min(us for _, us in table if us is not None) if table else None}
    plan = prop.plan()

    # the state the timed pass starts from, so that the per-kernel pass replays the same K steps
    start = prop.get_state()
    # untimed: launch every replay graph of both parities once (16-step and single-step graphs
    # of each parity: 16 + 1 + 16 + 1 steps end on the start parity), then back to the start
    prop.advance(34)
    prop.set_state(start)

    # timed pass: K steps replayed from hipGraphs, bracketed by device synchronisation
    t0 = time.perf_counter()
    prop.advance(args.steps)  # returns after the stream has drained
    t1 = time.perf_counter()
    secs = t1 - t0
    value = N * args.steps / secs
    radius, general = prop.stats()
    # FP64 mass of the device state after the timed pass (conservation check, reported)
    _, mass = prop.mass_f64()
    final = prop.get_state()  # warmup + K steps from the seed

    # per-kernel pass: the same K steps from the same start state (the stencil skips empty tiles,
    # so its time depends on how far the pattern has spread), with an event pair around every
    # pass; each pass (all of them write buffers their inputs do not alias) is launched
    # PROF_REPEATS times back to back inside its pair, so the event overhead is spread over the
    # repeats
    prop.set_state(start)
    prop.profile_enable(True, repeats=PROF_REPEATS)
    prop.advance(args.steps)
    prof = prop.profile_read()
    prop.profile_enable(False)
    kbytes = prop.kernel_bytes()
    per_kernel = {}
    for k, (ms, n) in prof.items():
        if n:
            avg_us = ms * 1e3 / n
            per_kernel[k] = {"launches": n, "avg_us": round(avg_us, 3),
                             "bytes_per_launch": kbytes[k],
                             "GBs": round(kbytes[k] / (avg_us * 1e-6) / 1e9, 1) if avg_us > 0 else None}
    dom = max(per_kernel, key=lambda k: per_kernel[k]["avg_us"] * per_kernel[k]["launches"])
    d = per_kernel[dom]
    achieved = d["GBs"]
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tfile):
        try:
            with open(tfile) as f:
                traffic = json.load(f).get(dom)
        except Exception:
            traffic = None
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom,
                "per_kernel": per_kernel}

    # the opt-in exact-conservation reintegration over the same K steps from the same start
    # (reported beside `value`, never as it): its step time and its FP64 mass drift
    prop.set_conservation("exact")
    prop.set_state(start)
    prop.advance(34)
    prop.set_state(start)
    t0 = time.perf_counter()
    prop.advance(args.steps)
    t1 = time.perf_counter()
    _, mass_ex = prop.mass_f64()
    prop.set_conservation("reference")
    exact_mode = {"ms_per_step": round((t1 - t0) * 1e3 / args.steps, 5),
                  "cell_steps_per_s": N * args.steps / (t1 - t0),
                  "mass_rel_drift": (mass_ex - mass0) / mass0 if mass0 else None,
                  "steps": args.steps}

    # the boundary with host buffers: hp_step and hp_run copy the state in from the caller's
    # (pageable) per-channel arrays and back out on every call.  PCIe-inclusive rates, reported
    # beside `value`, never as it.
    hs = prop.get_state()
    prop.step(hs)  # untimed first call
    n_call = 20
    t0 = time.perf_counter()
    for _ in range(n_call):
        prop.step(hs)
    t1 = time.perf_counter()
    prop.run(hs, 100)  # untimed first call
    t2 = time.perf_counter()
    prop.run(hs, 100)
    t3 = time.perf_counter()
    host_boundary = {
        "hp_step_ms_per_call": round((t1 - t0) * 1e3 / n_call, 3),
        "hp_step_cell_steps_per_s": N * n_call / (t1 - t0),
        "hp_run_100_ms_per_call": round((t3 - t2) * 1e3, 3),
        "hp_run_100_cell_steps_per_s": N * 100 / (t3 - t2),
        "state_bytes_each_way": 4 * N * cfg.channels,
        "note": "host arrays copied in and out on every call (pageable numpy buffers, as the "
                "reference's Vec<f32> channels)"}
    prop.close()
    drift = (mass - mass0) / mass0 if mass0 else None

    # the whole step against its algorithmic byte model and against a plain device copy
    step_bytes = step_model_bytes(cfg)
    step_gbs = step_bytes / (secs / args.steps) / 1e9
    copy_gbs = copy_bandwidth_gbs()
    step_model = {"bytes_per_step": step_bytes, "GBs": round(step_gbs, 1),
                  "frac_of_peak": round(step_gbs / HBM_PEAK_GBS, 4),
                  "copy_GBs": round(copy_gbs, 1) if copy_gbs else None,
                  "frac_of_copy": round(step_gbs / copy_gbs, 4) if copy_gbs else None,
                  "copy": "hipMemcpy device to device, 32 MiB each way, read + write bytes"}

    cpu = fp64 = None
    if not args.no_cpu:
        fp64 = fp64_error(cfg, seed, final, args.warmup + args.steps)
        cpu = cpu_baseline(cfg, seed, args.cpu_seconds)

    population = None
    if not args.no_population:
        population = population_measure(args.pop_steps, args.warmup, tune=not args.no_tune, lone=True)

    out = {
        "metric": "cell-steps/s at 1024^2x3ch (plain HipPropagator forward step)",
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": secs * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "multispecies 1024x1024x3ch, 9 kernels, R=15, dt=0.2",
                   "width": cfg.width, "height": cfg.height, "channels": cfg.channels,
                   "kernels": len(cfg.kernels), "kernel_radius": cfg.kernel_radius,
                   "max_gather_radius": radius, "general_path_steps": general,
                   "mass_f64_after": mass, "mass_rel_drift": drift,
                   "mass_drift_steps": args.steps, "plan": plan, "autotune": tune},
        "roofline": roofline,
        "host_boundary": host_boundary,
        "step_model": step_model,
        "exact_conservation": exact_mode,
        "fp64_error": fp64,
        "cpu_baseline": cpu,
        "population": population,
    }
    print(json.dumps(out), flush=True)
    return 0


def fp64_error(cfg, seed, final, steps_done: int, cap: int = 60):
    """Max |device - FP64 oracle| after the run's steps from the seed.  The device state after
    warmup + K steps is compared when that is at most `cap` steps; otherwise a fresh device
    run of `cap` steps from the seed is.  The f32 oracle's own distance from FP64 is given
    beside it (the reference's f32 arithmetic drifts from FP64 by that much)."""
    import numpy as np

    import autoverse_amd as av
    from oracle import oracle as O

    n = steps_done
    if n > cap:
        n = cap
        with av.HipPropagator(cfg) as p:
            st = av.SimulationState.from_seed(seed, cfg)
            p.run(st, n)
            final = st
    gpu = [np.asarray(c, np.float64) for c in final.channels]
    threads = max(1, min(os.cpu_count() or 1, 16, len(cfg.kernels)))
    st = O.seed_generate(seed, cfg.width, cfg.height, 1, cfg.channels)
    t0 = time.perf_counter()
    ref, _, _ = O.Propagator(cfg, np.float64, threads=threads).run([c.astype(np.float64) for c in st], n)
    t1 = time.perf_counter()
    o32, _, _ = O.Propagator(cfg, np.float32, threads=threads).run(st, n)
    err = max(float(np.abs(g - r).max()) for g, r in zip(gpu, ref))
    e32 = max(float(np.abs(o.astype(np.float64) - r).max()) for o, r in zip(o32, ref))
    g32 = max(float(np.abs(g - o).max()) for g, o in zip(gpu, o32))
    scale = max(float(np.abs(r).max()) for r in ref)
    return {"steps_from_seed": n, "gpu_vs_fp64_max_abs": err, "f32_oracle_vs_fp64_max_abs": e32,
            "gpu_vs_f32_oracle_max_abs": g32, "fp64_field_max": scale,
            "fp64_oracle_s": round(t1 - t0, 2)}


def population_bench(args) -> int:
    """256 members, sides U{64,128,192,256}, 1-3 channels (autoverse_amd/workloads.population):
    the batched step (one launch per pass per size class, hipGraph replay) against the plain
    per-instance path (one HipPropagator per member, advanced one after the other)."""
    r = population_measure(args.pop_steps, args.warmup, tune=not args.no_tune, lone=not args.no_lone)
    out = {
        "metric": "cell-steps/s, 256-member population (batched forward step)",
        "value": r["batched_forward"]["value"],
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": r["steps"],
        "warmup": args.warmup,
        "ms_per_step": r["batched_forward"]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": r["workload"], "members": r["members"], "cells": r["cells"]},
        **{k: v for k, v in r.items() if k not in ("workload", "members", "cells", "steps")},
    }
    print(json.dumps(out), flush=True)
    return 0


def population_measure(K: int, warmup: int, tune: bool = True, lone: bool = True) -> dict:
    """The 256-member population: batched forward step, per-instance forward step (one
    HipPropagator per member, advanced in turn), and batched forward + adjoint (gradient tape
    of K steps, then the backward pass).  `ratio_fwd_adjoint_vs_per_instance` is the second
    headline: batched forward+adjoint cell-steps/s over per-instance forward cell-steps/s."""
    import numpy as np

    import autoverse_amd as av
    from autoverse_amd import workloads
    from autoverse_amd.population import HipPopulation

    pop_def = workloads.population(256)
    cfgs = [c for c, _ in pop_def]
    seeds = [s for _, s in pop_def]
    cells = sum(c.width * c.height for c in cfgs)
    print(f"[population] {len(cfgs)} members, {cells} cells, {K} steps", file=sys.stderr, flush=True)

    pop = HipPopulation(cfgs)
    pop.set_seeds(seeds)
    pop.advance(warmup)
    plan = None
    if tune:  # untimed, states restored after
        plan, _ = pop.autotune(steps=10)
    start = [pop.get_state(i) for i in range(len(cfgs))]
    pop.advance(2)  # untimed: both step graphs launched once, then back to the start states
    for i, st in enumerate(start):
        pop.set_state(i, st)
    m0 = pop.mass_f64()
    c0, _ = pop.centroids()
    t0 = time.perf_counter()
    pop.advance(K)
    secs_b = time.perf_counter() - t0
    c1, _ = pop.centroids()
    m1 = pop.mass_f64()
    radius, general = pop.stats()
    # the persistent step (one launch per step, grid barriers between the passes) over the same
    # K steps from the same start states; it runs FFT radix order 0
    persistent = None
    from autoverse_amd._abi import HipError
    try:
        if pop.plan()["radix_order"] != 0:
            pop.set_plan({"radix_order": 0})
        pop.set_step_mode("persistent")
        for rep in range(2):  # untimed first: both step graphs launched once
            for i, st in enumerate(start):
                pop.set_state(i, st)
            t0 = time.perf_counter()
            pop.advance(2 if rep == 0 else K)
            secs_p = time.perf_counter() - t0
        _, timeouts = pop.step_mode()
        persistent = {"value": cells * K / secs_p, "unit": "cell-steps/s", "ms_per_step": secs_p * 1e3 / K,
                      "barrier_timeouts": timeouts, "plan": pop.plan(),
                      "path": "one persistent launch per step (grid barriers between the passes), hipGraph replay"}
        pop.set_step_mode("multi_launch")
    except HipError as e:
        persistent = {"error": str(e)}
    pop.close()
    disp = np.hypot(c1[:, 0] - c0[:, 0], c1[:, 1] - c0[:, 1])
    drift = float(np.max(np.abs(m1 - m0) / np.maximum(m0, 1e-30)))

    secs_i = None
    if lone:
        props = []
        for c, s in pop_def:
            p = av.HipPropagator(c)
            p.set_state(av.SimulationState.from_seed(s, c))
            p.advance(warmup)
            props.append(p)
        print("[population] per-instance pass", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        for p in props:
            p.advance(K)
        secs_i = time.perf_counter() - t0
        for p in props:
            p.close()

    # batched forward + adjoint over the same population: tape of K steps (checkpoint interval
    # chosen by the library from free HBM), then the backward pass for L = sum <seed, A_T>
    # with a fixed random seed
    pop = HipPopulation(cfgs)
    if plan is not None:
        pop.set_plan(plan)
    rng = np.random.default_rng(7)
    dl = [[rng.standard_normal(c.width * c.height).astype(np.float32) for _ in range(c.channels)] for c in cfgs]
    # warm, untimed, as the other two paths are warmed: the first tape allocates its states and
    # convolutions (tens of GB) and builds its launch lists; then back to the same start state
    pop.set_seeds(seeds)
    pop.advance(warmup)
    pop.tape_forward(K, 0)
    pop.tape_backward(dl)
    pop.set_seeds(seeds)
    pop.advance(warmup)
    print("[population] forward + adjoint pass", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    every = pop.tape_forward(K, 0)
    t1 = time.perf_counter()
    _, dth = pop.tape_backward(dl)
    secs_fa = time.perf_counter() - t0
    secs_f = t1 - t0
    pop.close()
    finite = all(np.all(np.isfinite(t)) for t in dth)

    value = cells * K / secs_b
    per_inst = cells * K / secs_i if secs_i else None
    fwd_adj = cells * K / secs_fa
    return {
        "workload": "population of 256: side U{64,128,192,256}, 1-3 channels, every source->target "
                    "kernel, R=15, dt=0.2", "members": len(cfgs), "cells": cells, "steps": K,
        "max_gather_radius": radius, "general_path_steps": general, "plan": plan,
        "batched_forward": {"value": value, "unit": "cell-steps/s", "ms_per_step": secs_b * 1e3 / K,
                            "path": "one launch per pass per size class, hipGraph replay"},
        "batched_forward_persistent": persistent,
        "per_instance_forward": None if per_inst is None else {
            "value": per_inst, "unit": "cell-steps/s", "ms_per_step": secs_i * 1e3 / K,
            "path": "one HipPropagator per member, advanced in turn"},
        "batched_forward_adjoint": {
            "value": fwd_adj, "unit": "cell-steps/s", "checkpoint_every": every,
            "ms_per_step": secs_fa * 1e3 / K, "tape_forward_ms_per_step": secs_f * 1e3 / K,
            "gradients_finite": finite},
        "speedup_batched_vs_per_instance": None if per_inst is None else value / per_inst,
        "ratio_fwd_adjoint_vs_per_instance": None if per_inst is None else fwd_adj / per_inst,
        "fitness": {"displacement_mean": float(disp.mean()), "displacement_max": float(disp.max())},
        "mass_drift_max_rel_fp64": drift,
    }


def grad_bench(args) -> int:
    """Differentiation workload: 32 members of 512^2 x 2 channels, 4 kernels each
    (autoverse_amd/workloads.differentiation with seeds 0x5EED4 + j), forward tape of T
    steps with a checkpoint every k, backward pass for L = sum <seed, A_T>."""
    import numpy as np

    from autoverse_amd import workloads
    from autoverse_amd.population import HipPopulation

    members = [workloads.differentiation(seed=0x5EED4 + j) for j in range(32)]
    cfgs = [c for c, _ in members]
    cells = sum(c.width * c.height for c in cfgs)
    T, k = args.grad_steps, args.grad_every
    print(f"[grad] {len(cfgs)} members, {cells} cells, T={T}, k={k}", file=sys.stderr, flush=True)
    pop = HipPopulation(cfgs)
    pop.set_seeds([s for _, s in members])
    rng = np.random.default_rng(11)
    dl = [[rng.standard_normal(c.width * c.height).astype(np.float32) for _ in range(c.channels)] for c in cfgs]
    pop.tape_forward(T, k)  # warm: allocations, graphs, first touches
    pop.tape_backward(dl)
    pop.set_seeds([s for _, s in members])
    t0 = time.perf_counter()
    pop.advance(T)
    secs_plain = time.perf_counter() - t0
    pop.set_seeds([s for _, s in members])
    t0 = time.perf_counter()
    pop.tape_forward(T, k)
    t1 = time.perf_counter()
    _, dth = pop.tape_backward(dl)
    t2 = time.perf_counter()
    pop.close()
    out = {
        "metric": "cell-steps/s, forward + adjoint (gradient tape) over 32 x 512^2 x 2ch",
        "value": cells * T / (t2 - t0),
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": T,
        "higher_is_better": True,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "differentiation: 32 members of 512^2, 2 channels, 4 kernels, R=15, dt=0.2",
                   "members": len(cfgs), "checkpoint_every": k},
        "ms_per_step": {"forward_plain": secs_plain * 1e3 / T, "tape_forward": (t1 - t0) * 1e3 / T,
                        "backward": (t2 - t1) * 1e3 / T, "forward_plus_adjoint": (t2 - t0) * 1e3 / T},
        "backward_over_forward": (t2 - t1) / secs_plain,
        "gradients_finite": all(bool(np.all(np.isfinite(t))) for t in dth),
    }
    print(json.dumps(out), flush=True)
    return 0


def sphere3d_bench(args) -> int:
    """3D workload: 128^3, one channel, one shell kernel, R=10, dt=0.1, Gaussian sphere seed
    (autoverse_amd/workloads.sphere_3d).  Steps default to 200 here."""
    import numpy as np

    import autoverse_amd as av
    from autoverse_amd import workloads

    cfg, seed = workloads.sphere_3d(128)
    N = cfg.width * cfg.height * cfg.depth
    K = args.steps if args.steps != 1000 else 200
    prop = av.HipPropagator(cfg)
    prop.set_state(av.SimulationState.from_seed(seed, cfg))
    prop.advance(args.warmup)
    tune = None
    if not args.no_tune:
        best, table = prop.autotune(steps=10)
        tune = {"plan": best, "candidates": len(table)}
    t0 = time.perf_counter()
    prop.advance(K)
    secs = time.perf_counter() - t0
    radius, general = prop.stats()
    prop.profile_enable(True)
    prop.advance(min(K, 50))
    prof = prop.profile_read()
    prop.profile_enable(False)
    kbytes = prop.kernel_bytes()
    per_kernel = {k: {"launches": n, "avg_us": round(ms * 1e3 / n, 3), "bytes_per_launch": kbytes[k]}
                  for k, (ms, n) in prof.items() if n}
    _, mass = prop.mass_f64()
    prop.close()
    out = {"metric": "cell-steps/s at 128^3 (3D HipPropagator forward step)", "value": N * K / secs,
           "unit": "cell-steps/s", "n_gpus": 1, "steps": K, "warmup": args.warmup,
           "ms_per_step": secs * 1e3 / K, "higher_is_better": True, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "sphere 128^3, 1 channel, shell kernel R=10, dt=0.1", "plan": tune,
                      "max_gather_radius": radius, "general_path_steps": general, "mass_f64_after": mass},
           "per_kernel": per_kernel}
    print(json.dumps(out), flush=True)
    return 0


def embedded_bench(args) -> int:
    """Embedded-parameter mode: the 1024^2 x 3 multispecies workload with embedding enabled
    (softmax mixing, temperature 1), every cell starting at EmbeddedState::from_seed's
    parameters; the device step against the oracle's (direct convolution, the reference's
    algorithm) on a bounded sample of steps on the host."""
    import numpy as np

    from autoverse_amd import workloads
    from autoverse_amd.embedded import EmbeddedState, HipEmbeddedPropagator
    from autoverse_amd.schema import EmbeddingConfig

    size = args.emb_size
    cfg, seed = workloads.multispecies(size=size)
    cfg.embedding = EmbeddingConfig(True, 1.0, args.emb_linear)
    N = cfg.width * cfg.height
    K = args.steps if args.steps != 1000 else 100
    st = EmbeddedState.from_seed(seed, cfg)
    hp = HipEmbeddedPropagator(cfg)
    hp.set_state(st)
    hp.advance(args.warmup)
    start = hp.get_state()
    t0 = time.perf_counter()
    hp.advance(K)
    secs = time.perf_counter() - t0
    radius = hp.last_radius()
    _, mass = hp.mass_f64()
    hp.close()
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        threads = min(os.cpu_count() or 1, 16)
        o = O.EmbeddedPropagator(cfg, cfg.embedding.__dict__, threads=threads)
        ch = [c.copy() for c in start.channels]
        ps = [p.data.copy() for p in start.params]
        n_cpu = 1
        t0 = time.perf_counter()
        o.run(ch, ps, n_cpu)
        cs = time.perf_counter() - t0
        cpu = {"value": N * n_cpu / cs, "unit": "cell-steps/s", "cores": threads, "kind": "port",
               "sample": f"{n_cpu} step of the same workload, f32, direct convolution as the reference, "
                         f"{cs * 1e3:.0f} ms/step"}
    out = {"metric": f"cell-steps/s at {size}^2x3ch (embedded-parameter step, softmax mixing)",
           "value": N * K / secs, "unit": "cell-steps/s", "n_gpus": 1, "steps": K, "warmup": args.warmup,
           "ms_per_step": secs * 1e3 / K, "higher_is_better": True, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"multispecies {size}x{size}x3ch, 9 kernels, R=15, dt=0.2, embedding enabled",
                      "linear_mixing": args.emb_linear, "search_radius": radius, "mass_f64_after": mass},
           "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)
    return 0


def step_model_bytes(cfg) -> int:
    """Algorithmic bytes of one step, with the half spectrum (S = 8 (W/2+1) H D bytes) taking
    two passes: C_s (4N + 3S) + 4 K S + 4 C N + 12 C N.  Per source channel: its N f32 cells read
    and its spectrum written, read and written again; per kernel: a spectrum read; per channel:
    U written once, then A, U read and A' written by the stencil (12 C N counts those three)."""
    n = cfg.width * cfg.height * max(cfg.depth, 1)
    s = 8 * (cfg.width // 2 + 1) * cfg.height * max(cfg.depth, 1)
    cs = len({k.source_channel for k in cfg.kernels})
    k = len(cfg.kernels)
    c = cfg.channels
    return cs * (4 * n + 3 * s) + 4 * k * s + 4 * c * n + 12 * c * n


def copy_bandwidth_gbs(nbytes: int = 32 << 20, reps: int = 20):
    """Read + write GB/s of hipMemcpy device to device (the runtime's copy kernel) at a size that,
    like one step's buffers, fits the Infinity Cache; None if the runtime is not loadable."""
    import ctypes
    from ctypes import byref, c_float, c_int, c_size_t, c_void_p

    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        return None
    hip.hipMalloc.argtypes = [ctypes.POINTER(c_void_p), c_size_t]
    hip.hipFree.argtypes = [c_void_p]
    hip.hipMemcpyAsync.argtypes = [c_void_p, c_void_p, c_size_t, c_int, c_void_p]
    hip.hipEventCreate.argtypes = [ctypes.POINTER(c_void_p)]
    hip.hipEventRecord.argtypes = [c_void_p, c_void_p]
    hip.hipEventSynchronize.argtypes = [c_void_p]
    hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(c_float), c_void_p, c_void_p]
    hip.hipEventDestroy.argtypes = [c_void_p]
    a, b, e0, e1 = c_void_p(), c_void_p(), c_void_p(), c_void_p()
    if hip.hipMalloc(byref(a), nbytes) != 0 or hip.hipMalloc(byref(b), nbytes) != 0:
        return None
    try:
        hip.hipEventCreate(byref(e0))
        hip.hipEventCreate(byref(e1))
        for _ in range(3):
            hip.hipMemcpyAsync(b, a, nbytes, 3, None)  # 3 = hipMemcpyDeviceToDevice
        hip.hipEventRecord(e0, None)
        for _ in range(reps):
            hip.hipMemcpyAsync(b, a, nbytes, 3, None)
        hip.hipEventRecord(e1, None)
        hip.hipEventSynchronize(e1)
        ms = c_float()
        if hip.hipEventElapsedTime(byref(ms), e0, e1) != 0 or ms.value <= 0:
            return None
        return 2.0 * nbytes * reps / (ms.value * 1e-3) / 1e9
    finally:
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
        hip.hipFree(a)
        hip.hipFree(b)


def copy_probe() -> int:
    """A plain copy kernel of known size: 64 Mi float32 (256 MiB) read and written, ten times
    (the calibration run for the traffic counters), then copies of the sizes one step's passes
    move, timed with events: the time a trivial kernel needs for the same bytes."""
    import torch

    a = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(10):
        b.copy_(a)
    torch.cuda.synchronize()
    out = {"copy_probe_bytes_each_way": a.numel() * 4, "launches": 10, "copy_us": {}}
    for mb in (4, 12.6, 19, 25, 30, 64, 256):
        n = int(mb * 1e6 / 4)
        x, y = a[:n], b[:n]
        for _ in range(3):
            y.copy_(x)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(20):
            y.copy_(x)
        e1.record()
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / 20
        out["copy_us"][f"{mb}MB_each_way"] = {"us": round(us, 2), "TBps_read_plus_write": round(2 * n * 4 / us / 1e6, 2)}
    print(json.dumps(out), flush=True)
    return 0


def cpu_baseline(cfg, seed, target_s: float):
    """CPU oracle (C restatement of CpuPropagator, reference thread decomposition: one thread
    per kernel for the convolutions, one per channel for the stencil) on a bounded sample."""
    import numpy as np

    from oracle import oracle as O

    ncpu = os.cpu_count() or 1
    threads = max(1, min(ncpu, 16, len(cfg.kernels)))
    st = O.seed_generate(seed, cfg.width, cfg.height, 1, cfg.channels)
    p = O.Propagator(cfg, np.float32, threads=threads)
    t0 = time.perf_counter()
    st, _, _ = p.run(st, 1)  # first step also measures the per-step cost
    one = time.perf_counter() - t0
    n = max(1, min(200, int(target_s / max(one, 1e-6))))
    t0 = time.perf_counter()
    st, _, _ = p.run(st, n)
    secs = time.perf_counter() - t0
    # the same oracle on one thread, 2 steps
    p.set_threads(1)
    t0 = time.perf_counter()
    p.run(st, 2)
    secs1 = (time.perf_counter() - t0) / 2
    p.close()
    return {"value": cfg.width * cfg.height * n / secs, "unit": "cell-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"{n} steps of the same 1024x1024x3ch workload, f32, {threads} threads "
                      f"(host has {ncpu} logical CPUs), {secs * 1e3 / n:.1f} ms/step",
            "one_thread": {"value": cfg.width * cfg.height / secs1, "unit": "cell-steps/s",
                           "ms_per_step": round(secs1 * 1e3, 1), "sample": "2 steps, 1 thread"},
            "fft_vs_scipy": fft_microbench(cfg.width, cfg.height)}


def fft_microbench(w: int, h: int, reps: int = 5):
    """The oracle's own complex 2D FFT (f32) against scipy.fft.fft2 (pocketfft, complex64, one
    worker) on the same W x H complex array: ms per transform, so the CPU baseline's FFT can be
    seen not to be a straw man."""
    import numpy as np

    from oracle import oracle as O
    try:
        import scipy.fft as sf
    except Exception:
        return None
    rng = np.random.default_rng(0)
    x = (rng.normal(size=(h, w)) + 1j * rng.normal(size=(h, w))).astype(np.complex64)
    buf = np.empty(2 * x.size, dtype=np.float32)
    buf[0::2], buf[1::2] = x.real.ravel(), x.imag.ravel()
    f = O.lib().orc_fft2d_f32
    f(O._p(buf), w, h, 0)
    t0 = time.perf_counter()
    for _ in range(reps):
        f(O._p(buf), w, h, 0)
    own = (time.perf_counter() - t0) / reps
    sf.fft2(x, workers=1)
    t0 = time.perf_counter()
    for _ in range(reps):
        sf.fft2(x, workers=1)
    ref = (time.perf_counter() - t0) / reps
    return {"size": f"{w}x{h} complex64", "oracle_ms": round(own * 1e3, 2), "scipy_pocketfft_ms": round(ref * 1e3, 2),
            "oracle_over_scipy": round(own / ref, 2), "threads": 1}


if __name__ == "__main__":
    sys.exit(main())
