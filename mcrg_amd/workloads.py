"""Deterministic synthetic workloads (job manifests) for benchmarks and tests.

Every draw comes from splitmix64 of a fixed index, so a manifest is the same on
every machine and run.
"""
import math

MASK64 = (1 << 64) - 1
KC = 0.5 * math.log(1.0 + math.sqrt(2.0))  # 0.44068679350977...

# the five couplings of the batching workload (T/Tc = 1.02, 1.00, 0.98, 0.94, 0.90)
BATCH_COUPLINGS = (0.4320459, 0.4406868, 0.4496804, 0.4688157, 0.489652)
BATCH_SIZES = (8, 16, 32, 64, 128)


def splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & MASK64
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31)


def unit(x):
    """Uniform double in [0, 1) from splitmix64(x)."""
    return (splitmix64(x) >> 11) * (1.0 / (1 << 53))


def levels_to(L, L_min=4):
    """Blockings from L down to L_min (b = 2)."""
    return int(round(math.log2(L // L_min)))


def batch_manifest(n=4096, n_therm=5000, n_meas=50000, sweeps_per_bin=1000):
    """Replica i: L = BATCH_SIZES[h % 5], |K| = BATCH_COUPLINGS[(h >> 32) % 5] with
    h = splitmix64(i), Philox seed i, blocking down to L = 4."""
    jobs = []
    for i in range(n):
        h = splitmix64(i)
        L = BATCH_SIZES[h % 5]
        jobs.append(dict(L=L, coupling=BATCH_COUPLINGS[(h >> 32) % 5], seed=i, n_therm=n_therm,
                         n_meas=n_meas, n_levels=levels_to(L), sweeps_per_bin=sweeps_per_bin))
    return jobs


def serving_stream(rate, duration=60.0, n_therm=1000, stream_seed=2024, sweeps_per_bin=500):
    """Open-loop Poisson arrivals at `rate` jobs/s for `duration` s.  Job id j:
    L = 2^k, k uniform on {4..9}; |K| ~ U[0.42, 0.47]; measured sweeps uniform
    on [1000, 20000]; n_therm thermalisation sweeps; blocking down to L = 4;
    seed = j.  Returns a list of (arrival time in s, job)."""
    out = []
    t = 0.0
    j = 0
    base = (stream_seed << 32) & MASK64
    while True:
        u = unit(base + 4 * j)
        t += -math.log(1.0 - u) / rate
        if t >= duration:
            break
        k = 4 + splitmix64(base + 4 * j + 1) % 6
        L = 1 << k
        coupling = 0.42 + 0.05 * unit(base + 4 * j + 2)
        n_meas = 1000 + splitmix64(base + 4 * j + 3) % 19001
        out.append((t, dict(L=L, coupling=coupling, seed=j, n_therm=n_therm, n_meas=n_meas,
                            n_levels=levels_to(L), sweeps_per_bin=sweeps_per_bin)))
        j += 1
    return out
