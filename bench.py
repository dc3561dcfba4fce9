#!/usr/bin/env python3
"""Benchmark of the MCRG hot path on one MI355X.

Workload (plain, HBM-resident): one 2D Ising lattice L = 2048, periodic,
|K| = 0.4406868, int8 spins, Philox seed 7, checkerboard Metropolis, every
sweep measured on 6 levels (5 b=2 majority blockings, 2048 -> 64) with the
NN/NNN/plaquette operators and their cross-correlations in int64.

One step = one sweep of the lattice plus its measurement.  W untimed warmup
steps, then exactly K timed steps bracketed by device synchronisation.
value = spin updates / ns over the timed steps (inputs resident in HBM).

A second, separately timed pass records HIP events around every half-sweep
launch on the chain's own stream to get the dominant kernel's average
duration for the roofline entry.  The CPU baseline runs the oracle (a scalar C
restatement of the same algorithm) on a bounded sample on the host.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "spin-updates/ns (batched vs plain) at matched y_t error; p50/p99 job latency"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(L, J, seed, nl, n_sweeps):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as orc
    orc.lib()
    t0 = time.perf_counter()
    orc.run_chain(L, J, seed, 0, n_sweeps, nl, n_sweeps)
    dt = time.perf_counter() - t0
    return {"value": L * L * n_sweeps / (dt * 1e9), "unit": "spin-updates/ns", "cores": 1,
            "kind": "port",
            "sample": f"oracle (scalar C, -O2) L={L} {n_sweeps} measured sweeps, {nl} levels: "
                      f"{dt:.2f} s on 1 host core"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--timing-steps", type=int, default=200)
    ap.add_argument("--cpu-sweeps", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the per-GPU work is "
              "independent chains, so there is nothing to exchange across GPUs. Run with --gpus 1.",
              file=sys.stderr)
        return 2

    import mcrg_amd

    L, J, seed, nl = 2048, 0.4406868, 7, 5
    spb = 1000
    total = args.warmup + args.steps + args.timing_steps
    ctx = mcrg_amd.Context(0)
    ch = mcrg_amd.Chain(ctx, L, J, seed, n_levels=nl, sweeps_per_bin=spb,
                        max_bins=max(1, math.ceil(total / spb)))
    log(f"bench: L={L} warmup={args.warmup} steps={args.steps}")
    ch.sweep(args.warmup, measure=True)
    ch.sync()
    t0 = time.perf_counter()
    ch.sweep(args.steps, measure=True)
    ch.sync()
    dt = time.perf_counter() - t0
    updates = L * L * args.steps
    value = updates / (dt * 1e9)
    log(f"bench: timed {args.steps} steps in {dt:.3f} s -> {value:.1f} updates/ns")

    # dominant kernel: the half-sweep, timed with events on the chain's stream
    ctx.set_timing(True)
    ch.sweep(args.timing_steps, measure=True)
    ch.sync()
    n_launch, ms = ctx.kernel_time(0)
    n_m0, ms_m0 = ctx.kernel_time(1)
    n_lv, ms_lv = ctx.kernel_time(2)
    ctx.set_timing(False)
    avg_us = 1e3 * ms / max(1, n_launch)
    bytes_per_launch = 1.5 * L * L  # read all L^2 sites, write L^2/2 (int8)
    achieved = bytes_per_launch / (avg_us * 1e-6) / 1e9
    log(f"bench: half-sweep avg {avg_us:.2f} us over {n_launch} launches -> {achieved:.1f} GB/s")

    ch.close()
    ctx.close()

    cpu = None
    if not args.no_cpu:
        log(f"bench: cpu baseline {args.cpu_sweeps} sweeps")
        cpu = cpu_baseline(L, J, seed, nl, args.cpu_sweeps)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "spin-updates/ns",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8",
        "data": "synthetic",
        "config": {"workload": "plain HBM int8 L=2048, |K|=0.4406868, seed 7, 5 blockings (2048->64), "
                               "measure every sweep", "L": L, "coupling": J, "levels": nl,
                   "layout": "int8", "schedule": "hbm"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_int8_half_sweep", "avg_us": avg_us,
                     "bytes_per_launch": bytes_per_launch},
        "kernels_us": {"half_sweep": avg_us, "measure0": 1e3 * ms_m0 / max(1, n_m0),
                       "levels_accumulate": 1e3 * ms_lv / max(1, n_lv)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
