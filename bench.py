#!/usr/bin/env python3
"""Benchmark of the MCRG hot path on one MI355X.

Headline workload (batched): the 4096-replica manifest of
mcrg_amd.workloads.batch_manifest -- L drawn from {8,16,32,64,128}, |K| from
{0.4320459, 0.4406868, 0.4496804, 0.4688157, 0.489652}, Philox seeds 0..4095,
bit-packed spins held in LDS, blocking down to L = 4 with int64 NN/NNN/plaquette
correlations per level, every sweep measured (the measurement phase that is
50000 of the 55000 sweeps of each replica).  One step = one launch advancing
every replica by `--sweeps-per-step` measured sweeps.  W untimed warmup steps,
then exactly K timed steps between device synchronisations.  value = spin
updates / ns over the timed steps.

Measured beside it (same JSON line):
  plain_mixed  the same replicas run one at a time through the plain int8 chain
               path (one launch per half-sweep), on a sample of the manifest
  matched_yt_error  the plain-path sample again through the batched schedule:
               identical records, hence identical y_t and jackknife error at
               equal sweeps (the comparison is at matched y_t error)
  plain_l2048  the plain path on one L = 2048 lattice, 5 blockings
  roofline     the resident kernel's algorithmic bytes per launch over its
               HIP-event-timed average duration (events on its own stream)
  cpu_baseline the oracle (scalar C) on the host's cores over a sample of the
               manifest (threads over independent replicas)
  bisection    the critical-coupling bisection (mcrg_run --bisect): per round 16
               couplings spanning the bracket (start [0.43, 0.45]), 64 lattices
               of L = 1024 (int8, HBM schedule, Swendsen-Wang, 1000 + 10000
               sweeps) and 1024 of L = 128 (bit-packed, LDS-resident, Metropolis,
               50000 + 10000 sweeps) in one batch; NN per site of L = 1024
               blocked 3 + m times against L = 128 blocked m times (m = 0..3);
               8 rounds; reports the final K_c with its jackknife error (groups
               of independent chains) and the change between the two deepest
               matched levels, against the exact 0.4406868, and the aggregate
               updates/ns
  serving      open-loop Poisson job streams (mcrg_amd.workloads.serving_stream:
               L = 2^k, k in 4..9, |K| ~ U[0.42, 0.47], 1e3 thermalisation plus
               U[1e3, 2e4] measured sweeps, blocking to L = 4, seed = job id)
               submitted to the job server at each --serving-rates rate for
               --serving-duration seconds; reports p50/p99 latency (submit to
               records on the host, y_t computed by the client) and aggregate
               spin updates / ns from the first arrival to the last completion
"""
import argparse
import concurrent.futures as cf
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "spin-updates/ns (batched vs plain) at matched y_t error; p50/p99 job latency"
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of k_resident<threads>, keyed by (replicas, sweeps per
# launch, threads, min_team, max_slot_lds, threads_large, sub_chunks), from two rocprofv3
# --pmc passes over `bench.py --only batched --sub-chunks 1` (FETCH_SIZE x 2.00
# + WRITE_SIZE x 1.00 KiB per dispatch, the scale factors read off the same
# counters on a 1 GiB read + 1 GiB write copy kernel, --copy-calibration).
# Configurations without a counter run report null.  The passes ran with
# every slot in one part (sub_chunks 1); the default splits slots into parts,
# each of which reloads and writes back its lattices, and is not counted.
# Every configuration the autotuner can keep has its passes (threads, min_team,
# max_slot_lds); threads_large only concerns L >= 512 and leaves this manifest's
# slots as they are.  A launch of two classes (max_slot_lds 65536) sums both.
MEASURED_TRAFFIC = {
    (4096, 100, 256, 16, 0, 0, 1): (3282.0 * 2.0 + 5048.0) * 1024.0,
    (4096, 100, 256, 16, 0, 512, 1): (3282.0 * 2.0 + 5048.0) * 1024.0,
    (4096, 100, 256, 32, 0, 0, 1): (3285.7 * 2.0 + 5047.8) * 1024.0,
    (4096, 100, 256, 8, 0, 0, 1): (3284.3 * 2.0 + 5047.8) * 1024.0,
    (4096, 100, 128, 16, 0, 0, 1): (3176.6 * 2.0 + 5066.3) * 1024.0,
    (4096, 100, 512, 16, 0, 0, 1): (3267.4 * 2.0 + 5012.3) * 1024.0,
    (4096, 100, 256, 16, 65536, 0, 1): (2.0 * 1746.6 * 2.0 + 2.0 * 2523.5) * 1024.0,
}
# The same kernel is bound by VALU issue, not HBM: rocprofv3 --pmc SQ_INSTS_VALU
# and GRBM_GUI_ACTIVE over the same one-part launches give the wave-level VALU
# instructions per launch and the busy cycles per XCD (8 XCDs, 1024 SIMDs that
# issue one wave64 VALU instruction per 4 cycles); parts add only their
# reloads.  The 128- and 512-thread shapes and the 64 KiB cap were counted on
# a kernel that read each lattice's two sweep counters with two loads (about
# 0.5 % more VALU instructions than this one).
MEASURED_VALU = {
    (4096, 100, 256, 16, 0, 0, 1): {"valu_wave_instr_per_launch": 4.6509e8, "gui_active_cycles_per_xcd": 2.4379e7 / 8},
    (4096, 100, 256, 16, 0, 512, 1): {"valu_wave_instr_per_launch": 4.6509e8, "gui_active_cycles_per_xcd": 2.4379e7 / 8},
    (4096, 100, 256, 32, 0, 0, 1): {"valu_wave_instr_per_launch": 4.6509e8, "gui_active_cycles_per_xcd": 2.4261e7 / 8},
    (4096, 100, 256, 8, 0, 0, 1): {"valu_wave_instr_per_launch": 4.6509e8, "gui_active_cycles_per_xcd": 2.4248e7 / 8},
    (4096, 100, 128, 16, 0, 0, 1): {"valu_wave_instr_per_launch": 4.9458e8, "gui_active_cycles_per_xcd": 3.4708e7 / 8},
    (4096, 100, 512, 16, 0, 0, 1): {"valu_wave_instr_per_launch": 4.6728e8, "gui_active_cycles_per_xcd": 3.7474e7 / 8},
    (4096, 100, 256, 16, 65536, 0, 1): {"valu_wave_instr_per_launch": 2.0 * 2.3104e8,
                                        "gui_active_cycles_per_xcd": 2.4222e7 / 8},
}


KIND_STRIP = 4  # mcrg_ctx_kernel_time kind of the strip kernel


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bytes_per_update(measured):
    # bit-packed: 3 * (1/8) B per site update, plus (5/3) * (1/8) B per site per
    # measured sweep (level-0 read + the geometric series of blocked levels)
    return 3.0 / 8.0 + (5.0 / 24.0 if measured else 0.0)


def run_batched(mcrg_amd, ctx, jobs, steps, warmup, sps, threads, autotune=False, cfg=None):
    """The autotuner (when on) tries its candidates during the warmup steps;
    otherwise `cfg` (min_team, max_slot_lds, threads_large) fixes the configuration."""
    n_sweeps = (warmup + steps) * sps
    bjobs = [dict(j, n_therm=0, n_meas=n_sweeps, sweeps_per_bin=max(1, n_sweeps)) for j in jobs]
    sites = sum(j["L"] * j["L"] for j in bjobs)
    with mcrg_amd.Batch(ctx, bjobs, threads=threads, sweeps_per_launch=sps, autotune=autotune, **(cfg or {})) as b:
        b.run(warmup * sps)
        b.sync()
        opts = b.options()
        if opts["autotune"]:
            raise RuntimeError("autotuning needs more warmup steps")
        t0 = time.perf_counter()
        b.run(steps * sps)
        b.sync()
        dt = time.perf_counter() - t0
    return sites * sps * steps / (dt * 1e9), dt, sites, opts


def autotune_pass(mcrg_amd, ctx, jobs, sps, threads):
    """Untimed autotuning on the same manifest: launches of `sps` measured
    sweeps until the tuner has timed every candidate (HIP events on the
    engine's stream) and kept the fastest; returns the configuration it kept
    and the launches it took."""
    n_sweeps = 64 * sps
    bjobs = [dict(j, n_therm=0, n_meas=n_sweeps, sweeps_per_bin=max(1, n_sweeps)) for j in jobs]
    with mcrg_amd.Batch(ctx, bjobs, threads=threads, sweeps_per_launch=sps, autotune=True) as b:
        launches = 0
        while b.options()["autotune"] and launches < 64:
            b.run(sps)
            b.sync()
            launches += 1
        opts = b.options()
    if opts["autotune"]:
        raise RuntimeError("the autotuner did not finish")
    return {k: opts[k] for k in ("threads", "min_team", "max_slot_lds", "threads_large", "sub_chunks")}, launches


def run_batched_timed(mcrg_amd, ctx, jobs, sps, tuned, launches):
    """Kernel duration of the resident kernel from HIP events on its stream,
    in the configuration `tuned` (the one the timed leg ran)."""
    n_sweeps = launches * sps
    bjobs = [dict(j, n_therm=0, n_meas=n_sweeps, sweeps_per_bin=max(1, n_sweeps)) for j in jobs]
    with mcrg_amd.Batch(ctx, bjobs, threads=tuned["threads"], sweeps_per_launch=sps, min_team=tuned["min_team"],
                        max_slot_lds=tuned["max_slot_lds"], threads_large=tuned["threads_large"],
                        sub_chunks=tuned["sub_chunks"]) as b:
        b.run(sps)  # first launch builds the slot table
        b.sync()
        n0, ms0 = ctx.kernel_time(3)
        ctx.set_timing(True)
        b.run((launches - 1) * sps)
        b.sync()
        n1, ms1 = ctx.kernel_time(3)
        ctx.set_timing(False)
    return (ms1 - ms0) / max(1, n1 - n0), n1 - n0


def run_plain_mixed(mcrg_amd, ctx, jobs, sweeps):
    """Plain path, one replica at a time: every chain is created, initialised
    and warmed (graphs captured) before the clock starts; the timed region is
    each chain's `sweeps` measured sweeps in turn, each run to completion.
    Records are read after it.  Returns updates/ns, seconds and the records."""
    chains = []
    try:
        for j in jobs:
            ch = mcrg_amd.Chain(ctx, j["L"], j["coupling"], j["seed"], n_levels=j["n_levels"],
                                sweeps_per_bin=2 * sweeps, max_bins=1)
            chains.append(ch)
            ch.sweep(sweeps, measure=True)  # warm-up: captures the replayed graph, loads the kernels
            ch.sync()
        updates = sum(j["L"] * j["L"] * sweeps for j in jobs)
        t0 = time.perf_counter()
        for ch in chains:
            ch.sweep(sweeps, measure=True)
            ch.sync()
        dt = time.perf_counter() - t0
        recs = [ch.records() for ch in chains]
    finally:
        for ch in chains:
            ch.close()
    return updates / (dt * 1e9), dt, recs


def matched_yt(mcrg_amd, ctx, jobs, sweeps, plain_recs):
    """The same replicas and sweeps through the batched schedule: the records
    (every sum and product the y_t estimate and its jackknife error are formed
    from) equal the plain path's bit for bit, so both paths reach the same y_t
    error after the same sweeps and updates/ns compare at matched error."""
    bjobs = [dict(j, n_therm=0, n_meas=sweeps, sweeps_per_bin=sweeps) for j in jobs]
    with mcrg_amd.Batch(ctx, bjobs) as b:
        b.run_all()
        same = all(bool((b.records(k) == r).all()) for k, r in enumerate(plain_recs))
    return {"jobs": len(jobs), "sweeps": sweeps, "records_identical": same}


def run_plain(mcrg_amd, ctx, L, nl, steps, measure=True, warm=36, kernels=None):
    """One plain int8 chain (|K| = 0.4406868, seed 7): `warm` sweeps (which
    also capture the replayed graph), then `steps` timed sweeps.  When
    `kernels` is a dict it receives the strip kernel's launches and duration
    over the timed sweeps (HIP events on the chain's stream; none when the
    chain runs on the per-half-sweep schedule)."""
    total = warm + 2 * steps  # the timed sweeps run twice: wall clock, then HIP events
    spb = max(1, min(total, mcrg_amd.max_sweeps_per_bin(L)))  # int64-safe bins (L = 16384: 31 sweeps)
    with mcrg_amd.Chain(ctx, L, 0.4406868, 7, n_levels=nl, sweeps_per_bin=spb,
                        max_bins=-(-total // spb)) as ch:
        ch.sweep(warm, measure=measure)
        ch.sync()
        n0, ms0 = ctx.kernel_time(KIND_STRIP)
        t0 = time.perf_counter()
        ch.sweep(steps, measure=measure)
        ch.sync()
        dt = time.perf_counter() - t0
        # the same sweeps once more with the kernel timed (events cost a little)
        ctx.set_timing(True)
        ch.sweep(steps, measure=measure)
        ch.sync()
        ctx.set_timing(False)
        n1, ms1 = ctx.kernel_time(KIND_STRIP)
    if kernels is not None and n1 > n0:
        kernels.update({"kernel": "k_strip", "launches": n1 - n0, "sweeps_per_launch": steps,
                        "avg_ms": (ms1 - ms0) / (n1 - n0),
                        "updates_per_ns_kernel": L * L * steps * (n1 - n0) / ((ms1 - ms0) * 1e6),
                        "source": "HIP events on the chain's stream"})
    return L * L * steps / (dt * 1e9)


def run_plain_2048(mcrg_amd, ctx, steps, kernels=None):
    return run_plain(mcrg_amd, ctx, 2048, 5, steps, True, 20, kernels)


def run_serving(mcrg_amd, workloads, ctx, rate, duration, drain_limit, quantum, threads):
    stream = workloads.serving_stream(rate, duration)
    srv = mcrg_amd.Server(ctx, launch_quantum=quantum, threads=threads)
    try:
        return _serve(mcrg_amd, srv, stream, rate, duration, drain_limit, quantum)
    finally:
        srv.close()  # stops the scheduler thread


def _serve(mcrg_amd, srv, stream, rate, duration, drain_limit, quantum):
    srv.start()
    lat, yts, updates = [], [], 0.0
    yt_fail = [0]
    n_done = 0
    next_log = 10.0

    def collect():
        nonlocal n_done, updates
        for jid in srv.completed():
            job = srv.jobs[jid]
            rec = srv.records(jid)
            nl = job["n_levels"]
            try:  # y_t from the 2-operator RG matrix between L/2 and L/4
                yts.append(mcrg_amd.rg_exponent(rec, nl, 1, 2)["yt"])
            except mcrg_amd.MCRGError:
                yt_fail[0] += 1
            p = srv.poll(jid)
            lat.append(p["finish_s"] - p["submit_s"])
            updates += job["L"] ** 2 * (job["n_therm"] + job["n_meas"])
            srv.release(jid)
            del srv.jobs[jid]
            n_done += 1

    t0 = time.perf_counter()
    late = 0.0
    for ta, job in stream:
        now = time.perf_counter() - t0
        if ta > now:
            time.sleep(ta - now)
        else:
            late = max(late, now - ta)
        srv.submit(job)
        if time.perf_counter() - t0 > next_log:
            collect()
            log(f"bench: serving {rate}/s t={time.perf_counter() - t0:.0f}s submitted "
                f"{srv.stats()['submitted']} completed {srv.stats()['completed']} collected {n_done}")
            next_log += 10.0
    t_last_arrival = time.perf_counter() - t0
    while n_done < len(stream) and time.perf_counter() - t0 < duration + drain_limit:
        collect()
        time.sleep(0.005)
        if time.perf_counter() - t0 > next_log:
            log(f"bench: serving {rate}/s draining t={time.perf_counter() - t0:.0f}s collected {n_done}/{len(stream)}")
            next_log += 10.0
    collect()
    t_end = time.perf_counter() - t0
    st = srv.stats()
    srv.stop()
    lat_ms = sorted(1e3 * x for x in lat)

    def pct(q):
        return lat_ms[min(len(lat_ms) - 1, int(q * len(lat_ms)))] if lat_ms else None

    return {"rate_per_s": rate, "duration_s": duration, "jobs": len(stream), "completed": n_done,
            "p50_ms": pct(0.50), "p99_ms": pct(0.99), "max_ms": lat_ms[-1] if lat_ms else None,
            "updates_per_ns": updates / (t_end * 1e9), "wall_s": t_end, "last_arrival_s": t_last_arrival,
            "max_submit_lag_s": late, "launches": st["launches"], "busy_s": st["busy_s"],
            "device_wait_s": st["device_wait_s"],
            "yt_median": sorted(yts)[len(yts) // 2] if yts else None, "yt_undefined": yt_fail[0],
            "launch_quantum_us": quantum or 4000}


def host_cpu():
    """What the host reports about its cores."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = None
    return {"model": model, "os_cpu_count": os.cpu_count(), "affinity_cpus": affinity}


def cpu_baseline(jobs, sweeps, threads):
    """The oracle (the same Philox Metropolis, blocking and int64 records in
    scalar C) on the host's cores:
      value       threads over independent replicas of the batching manifest
      one_thread  the plain configurations on one thread: L = 8 in full
                  (1e4 + 1e5 sweeps, one blocking, seed 1) and the first 100
                  sweeps of L = 2048 (seed 7)"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as orc
    orc.lib()

    def one(j):
        orc.run_chain(j["L"], j["coupling"], j["seed"], 0, sweeps, j["n_levels"], sweeps)
        return j["L"] * j["L"] * sweeps

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:  # ctypes releases the GIL
        updates = sum(ex.map(one, jobs))
    dt = time.perf_counter() - t0
    t1 = time.perf_counter()
    orc.run_chain(8, 0.4406868, 1, 10000, 100000, 1, 10000)
    dt1 = time.perf_counter() - t1
    t2 = time.perf_counter()
    orc.run_chain(2048, 0.4406868, 7, 100, 0, 5, 1)
    dt2 = time.perf_counter() - t2
    host = host_cpu()
    return {"value": updates / (dt * 1e9), "unit": "spin-updates/ns", "cores": threads, "kind": "port",
            "sample": f"oracle (scalar C, -O3 -march=x86-64-v3) on {len(jobs)} replicas of the batching manifest x "
                      f"{sweeps} measured sweeps each, {threads} threads over independent replicas: {dt:.1f} s",
            "host": dict(host, threads_used=threads),
            "one_thread": {
                "L8_full": {"updates_per_ns": 64 * 110000 / (dt1 * 1e9), "seconds": dt1,
                            "sample": "L=8, |K|=0.4406868, seed 1, 1e4 + 1e5 sweeps, every measured sweep "
                                      "blocked once (8->4), int64 records"},
                "L2048_first_100": {"updates_per_ns": 2048 * 2048 * 100 / (dt2 * 1e9), "seconds": dt2,
                                    "sample": "L=2048, |K|=0.4406868, seed 7, sweeps 1..100 (unmeasured)"}}}


def valu_roofline(cfg_key, sites, sps, avg_ms, clock_khz, cus):
    """VALU issue roof of the resident kernel: wave64 VALU instructions per
    launch (rocprofv3 SQ_INSTS_VALU) x 64 lanes over the launch's event-timed
    duration, against cus x 4 SIMDs x 16 lanes x the engine clock."""
    m = MEASURED_VALU.get(cfg_key) or MEASURED_VALU.get(cfg_key[:-1] + (1,))
    if not m:
        return None
    peak = cus * 4 * 16 * clock_khz * 1e3 / 1e12  # T lane-instructions / s
    achieved = 64.0 * m["valu_wave_instr_per_launch"] / (avg_ms * 1e-3) / 1e12
    exact = cfg_key in MEASURED_VALU
    return {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "T lane-instr/s", "frac": achieved / peak,
            "lane_instr_per_update": 64.0 * m["valu_wave_instr_per_launch"] / (sites * sps),
            "source": ("rocprofv3 --pmc SQ_INSTS_VALU on bench.py --only batched "
                       + ("(this configuration)" if exact else "with every slot in one part (the parts of this "
                          "configuration add only their reloads)"))
                      + f"; duration {avg_ms:.4f} ms from HIP events on the kernel's stream in this run"}


def run_bisection(rounds):
    import subprocess
    cmd = [os.path.join(ROOT, "mcrg_amd", "mcrg_run"), "--bisect", str(rounds), "--eq", "1000", "--eq-small",
           "50000", "--samples", "10000", "--bins", "20", "--quiet"]
    log("bench: " + " ".join(cmd[1:]))
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=1200)
    if out.returncode != 0:
        raise RuntimeError(f"mcrg_run --bisect failed: {out.stderr.strip()}")
    res = json.loads(out.stdout.strip().splitlines()[-1])
    res["config"] = ("per round 16 couplings; 64 x L=1024 int8 Swendsen-Wang (1000 + 10000 sweeps) and "
                     "1024 x L=128 bit-packed Metropolis (50000 + 10000 sweeps) in one batch; NN per site of "
                     "L=1024 blocked 3+m times vs L=128 blocked m times, m = 0..3, crossing at m = 3")
    return res


def copy_calibration(n_bytes):
    """A plain elementwise kernel that reads and writes n_bytes each (torch.neg
    on int32), five times: the reference point for the FETCH_SIZE / WRITE_SIZE
    counters of a profiled run."""
    import torch
    x = torch.ones(n_bytes // 4, dtype=torch.int32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        torch.neg(x, out=y)
    torch.cuda.synchronize()
    log(f"bench: copy calibration {n_bytes} bytes read + {n_bytes} bytes written per launch, 5 launches")
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--sweeps-per-step", type=int, default=100)
    ap.add_argument("--threads", type=int, default=256)
    ap.add_argument("--replicas", type=int, default=4096)
    ap.add_argument("--plain-sample", type=int, default=50)
    ap.add_argument("--cpu-sweeps", type=int, default=4000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--serving-rates", default="200,1000,5000",
                    help="comma-separated arrival rates in jobs/s for the serving leg ('' = none)")
    ap.add_argument("--serving-duration", type=float, default=60.0)
    ap.add_argument("--plain-sweeps", type=int, default=64, help="timed measured sweeps per plain_mixed chain")
    ap.add_argument("--no-autotune", action="store_true", help="skip the untimed autotuning pass")
    ap.add_argument("--serving-drain", type=float, default=120.0,
                    help="seconds after the last arrival to wait for completions")
    ap.add_argument("--serving-quantum", type=int, default=0)
    ap.add_argument("--bisect-rounds", type=int, default=8, help="rounds of the bisection leg (0 = skip)")
    ap.add_argument("--min-team", type=int, default=16, help="--only batched: lanes per lattice at least")
    ap.add_argument("--max-slot-lds", type=int, default=0, help="--only batched: LDS cap per slot (0 = default)")
    ap.add_argument("--threads-large", type=int, default=0, help="--only batched: workgroup size for L = 512")
    ap.add_argument("--max-team", type=int, default=0, help="--only batched: lanes per lattice at most (0 = threads)")
    ap.add_argument("--batch-L", type=int, default=0,
                    help="--only batched: every replica at this L (coupling and seed as in the manifest)")
    ap.add_argument("--sub-chunks", type=int, default=0,
                    help="parts per slot and launch (0 = automatic; the autotuned run uses 0)")
    ap.add_argument("--plain-L", type=int, default=16384, help="--only plain: lattice side")
    ap.add_argument("--plain-levels", type=int, default=1, help="--only plain: blockings")
    ap.add_argument("--plain-measure", type=int, default=0, help="--only plain: measure every sweep (1) or not (0)")
    ap.add_argument("--only", choices=["batched", "plain_l2048", "plain"], default=None,
                    help="run one leg alone and print its line to stderr (for profiler runs)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run the counter-calibration copy kernel and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): the replicas are "
              "independent chains and this build runs on a single device. Run with --gpus 1.",
              file=sys.stderr)
        return 2

    # the bisection driver runs as a child process before this one touches the GPU
    bisection = None
    if args.bisect_rounds and not args.only:
        bisection = run_bisection(args.bisect_rounds)
        log(f"bench: bisection Kc = {bisection['Jc']:.7f} +- {bisection['Jc_err']:.7f} (stat) "
            f"+- {bisection['Jc_sys']:.7f} (deepest two levels) "
            f"({bisection['updates_per_ns']:.1f} updates/ns, {bisection['seconds']:.1f} s)")

    import mcrg_amd
    from mcrg_amd import workloads

    jobs = workloads.batch_manifest(args.replicas)
    if args.only == "batched" and args.batch_L:
        jobs = [dict(j, L=args.batch_L, n_levels=workloads.levels_to(args.batch_L)) for j in jobs]
    ctx = mcrg_amd.Context(0)
    if args.only == "batched":
        cfg = dict(min_team=args.min_team, max_slot_lds=args.max_slot_lds, threads_large=args.threads_large,
                   sub_chunks=args.sub_chunks, max_team=args.max_team)
        value, dt, _, _ = run_batched(mcrg_amd, ctx, jobs, args.steps, args.warmup, args.sweeps_per_step,
                                      args.threads, cfg=cfg)
        log(f"bench: batched {value:.1f} updates/ns ({dt:.3f} s timed), {cfg}")
        ctx.close()
        return 0
    if args.only == "plain":
        kern = {}
        rate = run_plain(mcrg_amd, ctx, args.plain_L, args.plain_levels, args.steps, bool(args.plain_measure),
                         kernels=kern)
        if kern:
            log(f"bench: {kern['kernel']} {kern['avg_ms']:.4f} ms per launch of {kern['sweeps_per_launch']} sweeps "
                f"({kern['updates_per_ns_kernel']:.1f} updates/ns inside the kernel, {kern['source']})")
        log(f"bench: plain L={args.plain_L} levels={args.plain_levels} measure={args.plain_measure}: "
            f"{rate:.1f} updates/ns, {args.plain_L ** 2 / rate / 1e3:.1f} us per sweep")
        ctx.close()
        return 0
    if args.only == "plain_l2048":
        log(f"bench: plain L=2048: {run_plain_2048(mcrg_amd, ctx, args.steps):.1f} updates/ns")
        ctx.close()
        return 0
    dev = ctx.device_info()
    log(f"bench: {dev['name']} {dev['compute_units']} CUs; {len(jobs)} replicas, "
        f"{args.warmup}+{args.steps} steps x {args.sweeps_per_step} sweeps")
    autotuned = {"ran": False}
    cfg = dict(min_team=16, max_slot_lds=0, threads_large=0, sub_chunks=0)
    threads = args.threads
    if not args.no_autotune:
        chosen, tune_launches = autotune_pass(mcrg_amd, ctx, jobs, args.sweeps_per_step, args.threads)
        threads = chosen.pop("threads")
        cfg = chosen
        autotuned = {"ran": True, "launches": tune_launches, "kept": dict(cfg, threads=threads)}
        log(f"bench: autotuner kept threads={threads} {cfg} after {tune_launches} untimed launches")
    value, dt, sites, tuned = run_batched(mcrg_amd, ctx, jobs, args.steps, args.warmup, args.sweeps_per_step,
                                          threads, cfg=cfg)
    log(f"bench: batched {value:.1f} updates/ns ({dt:.3f} s timed); configuration "
        f"threads={tuned['threads']} min_team={tuned['min_team']} max_slot_lds={tuned['max_slot_lds']} "
        f"threads_large={tuned['threads_large']} sub_chunks={tuned['sub_chunks']}")

    avg_ms, n_launch = run_batched_timed(mcrg_amd, ctx, jobs, args.sweeps_per_step, tuned, 11)
    cfg_key = (len(jobs), args.sweeps_per_step, tuned["threads"], tuned["min_team"], tuned["max_slot_lds"],
               tuned["threads_large"], tuned["sub_chunks"])
    bpl = bytes_per_update(True) * sites * args.sweeps_per_step
    achieved = bpl / (avg_ms * 1e-3) / 1e9
    log(f"bench: resident kernel {avg_ms:.3f} ms/launch over {n_launch} launches -> {achieved:.1f} GB/s algorithmic")

    plain_mixed = plain_2048 = matched = None
    plain_2048_kernel = {}
    if not args.no_plain:
        sample = jobs[:args.plain_sample]
        plain_mixed, pdt, plain_recs = run_plain_mixed(mcrg_amd, ctx, sample, args.plain_sweeps)
        log(f"bench: plain path on {len(sample)} replicas x {args.plain_sweeps} sweeps: {plain_mixed:.3f} updates/ns "
            f"({pdt:.3f} s timed)")
        matched = matched_yt(mcrg_amd, ctx, sample, 2 * args.plain_sweeps, plain_recs)
        log(f"bench: batched records of the same replicas identical: {matched['records_identical']}")
        plain_2048 = run_plain_2048(mcrg_amd, ctx, 500, plain_2048_kernel)
        log(f"bench: plain L=2048: {plain_2048:.1f} updates/ns")
    valu = valu_roofline(cfg_key, sites, args.sweeps_per_step, avg_ms, dev["clock_khz"], dev["compute_units"])
    serving = []
    for rate in [float(r) for r in args.serving_rates.split(",") if r.strip()]:
        res = run_serving(mcrg_amd, workloads, ctx, rate, args.serving_duration, args.serving_drain,
                          args.serving_quantum, args.threads)
        log(f"bench: serving {rate}/s: {res['completed']}/{res['jobs']} jobs, p50 {res['p50_ms']} ms, "
            f"p99 {res['p99_ms']} ms, {res['updates_per_ns']:.1f} updates/ns")
        serving.append(res)
    ctx.close()

    cpu = None
    if not args.no_cpu:
        # every core this process may run on, at most 16 (one GPU's share of a shared host)
        host = host_cpu()
        threads = max(1, min(16, host["affinity_cpus"] or os.cpu_count() or 1))
        sample = jobs[:1024]
        log(f"bench: cpu baseline on {len(sample)} replicas x {args.cpu_sweeps} sweeps, {threads} threads ({host})")
        cpu = cpu_baseline(sample, args.cpu_sweeps, threads)

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
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": f"batched: {len(jobs)} replicas, L in {{8..128}}, 5 couplings, seeds 0..{len(jobs)-1}, "
                               f"bit-packed LDS-resident, blocking to L=4, every sweep measured",
                   "replicas": len(jobs), "sweeps_per_step": args.sweeps_per_step, "threads": tuned["threads"],
                   "autotuned": autotuned,
                   "run_config": {k: tuned[k] for k in ("threads", "min_team", "max_slot_lds", "threads_large",
                                                        "sub_chunks")},
                   "sites": sites, "layout": "packed", "schedule": "resident", "device": dev["name"]},
        "roofline": dict(valu or {"bound": "valu", "achieved": None, "peak": None, "unit": "T lane-instr/s",
                                  "frac": None, "source": "no counter run for this configuration"},
                         traffic=MEASURED_TRAFFIC.get(cfg_key), kernel=f"k_resident<{tuned['threads']}>",
                         avg_ms=avg_ms,
                         hbm={"achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": achieved / HBM_PEAK_GBS, "bytes_per_launch": bpl,
                              "counted_bytes_one_part": MEASURED_TRAFFIC.get(cfg_key[:-1] + (1,)),
                              "counted_basis": "rocprofv3 FETCH_SIZE x 2 + WRITE_SIZE on the same configuration "
                                               "with every slot in one part (parts add their reloads)",
                              "basis": "algorithmic bytes (3/8 B per site update + 5/24 B per site per measured "
                                       "sweep, bit-packed streaming model)"}),
        "plain_mixed": plain_mixed,
        "plain_l2048": plain_2048,
        "plain_l2048_kernel": plain_2048_kernel or None,
        "batched_over_plain_mixed": (value / plain_mixed) if plain_mixed else None,
        "matched_yt_error": matched,
        "serving": serving,
        "bisection": bisection,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
