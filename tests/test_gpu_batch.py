"""GPU parity of the batched LDS-resident schedule.

Bit-exact against the oracle for mixed L / coupling / level counts / bin sizes,
for several workgroup sizes and launch lengths; bit-identical to the plain int8
chain path (a different layout and schedule); and every job's result is the
same alone and in a mix.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KC = 0.5 * math.log(1.0 + math.sqrt(2.0))


@pytest.fixture(scope="module")
def amd():
    import mcrg_amd
    return mcrg_amd


@pytest.fixture(scope="module")
def ctx(amd):
    c = amd.Context(0)
    yield c
    c.close()


def _jobs(seed, n, Ls=(8, 16, 32, 64, 128), meas_max=25):
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n):
        L = int(rng.choice(Ls))
        jobs.append(dict(L=L, coupling=float(rng.choice([0.4320459, KC, 0.4496804, 0.4688157, 0.489652, 0.0])),
                         seed=int(rng.integers(0, 2**63)), n_therm=int(rng.integers(0, 8)),
                         n_meas=int(rng.integers(0, meas_max)), n_levels=int(rng.integers(0, int(math.log2(L)))),
                         sweeps_per_bin=int(rng.integers(1, 9))))
    return jobs


def _check_vs_oracle(orc, b, jobs):
    for j, job in enumerate(jobs):
        s_ref, r_ref = orc.run_chain(job["L"], job["coupling"], job["seed"], job["n_therm"], job["n_meas"],
                                     job["n_levels"], job["sweeps_per_bin"])
        assert np.array_equal(b.spins(j), s_ref), f"spins differ for job {j}: {job}"
        got = b.records(j)
        assert np.array_equal(got, r_ref), f"records differ for job {j}: {job} at {np.argwhere(got != r_ref)[:5]}"
        assert b.state(j) == (job["n_therm"] + job["n_meas"], job["n_meas"])


@pytest.mark.parametrize("threads,chunk", [(256, 250), (64, 1), (128, 7), (512, 3)])
def test_mixed_batch_bit_exact(amd, ctx, orc, threads, chunk):
    jobs = _jobs(threads + chunk, 40)
    print("batch", threads, chunk, flush=True)
    with amd.Batch(ctx, jobs, threads=threads, sweeps_per_launch=chunk) as b:
        b.run_all()
        b.sync()
        _check_vs_oracle(orc, b, jobs)


@pytest.mark.parametrize("threads,chunk,parts", [(256, 250, 2), (256, 13, 3), (64, 5, 7), (128, 9, 64), (512, 40, 4)])
def test_parted_slots_bit_exact(amd, ctx, orc, threads, chunk, parts):
    """Each slot's sweeps of a launch split into `parts` sequential parts that
    any workgroup may take (waiting for the slot's previous part): the chains
    are the oracle's, also with parts of no sweeps (64 parts of 9 sweeps)."""
    jobs = _jobs(7 * parts + threads, 48, meas_max=40)
    print("parted batch", threads, chunk, parts, flush=True)
    with amd.Batch(ctx, jobs, threads=threads, sweeps_per_launch=chunk, sub_chunks=parts) as b:
        assert b.options()["sub_chunks"] == parts
        b.run_all()
        b.sync()
        _check_vs_oracle(orc, b, jobs)


def test_large_lattices_bit_exact(amd, ctx, orc):
    jobs = [dict(L=256, coupling=KC, seed=11, n_therm=3, n_meas=9, n_levels=6, sweeps_per_bin=4),
            dict(L=512, coupling=0.4496804, seed=12, n_therm=2, n_meas=6, n_levels=7, sweeps_per_bin=5),
            dict(L=512, coupling=0.3, seed=13, n_therm=4, n_meas=0, n_levels=2, sweeps_per_bin=1)]
    with amd.Batch(ctx, jobs, sweeps_per_launch=5) as b:
        b.run_all()
        _check_vs_oracle(orc, b, jobs)


def test_batch_equals_plain_chain(amd, ctx):
    """Same seed on two layouts and schedules: int8 HBM chain vs packed LDS batch."""
    for L, nl in ((16, 3), (64, 4), (128, 5)):
        job = dict(L=L, coupling=KC, seed=77 + L, n_therm=11, n_meas=40, n_levels=nl, sweeps_per_bin=10)
        with amd.Batch(ctx, [job], sweeps_per_launch=13) as b:
            b.run_all()
            s_b, r_b = b.spins(0), b.records(0)
        with amd.Chain(ctx, L, KC, job["seed"], n_levels=nl, sweeps_per_bin=10, max_bins=4) as ch:
            ch.sweep(11)
            ch.sweep(40, measure=True)
            assert np.array_equal(ch.spins(), s_b)
            assert np.array_equal(ch.records(), r_b)


def test_alone_equals_in_mix(amd, ctx):
    probe = dict(L=64, coupling=KC, seed=123456789, n_therm=20, n_meas=60, n_levels=4, sweeps_per_bin=15)
    with amd.Batch(ctx, [probe], sweeps_per_launch=25) as b:
        b.run_all()
        alone = (b.spins(0), b.records(0))
    mix = _jobs(5, 300, meas_max=80)
    mix.insert(137, probe)
    with amd.Batch(ctx, mix, sweeps_per_launch=25) as b:
        b.run_all()
        assert np.array_equal(b.spins(137), alone[0])
        assert np.array_equal(b.records(137), alone[1])


def test_incremental_runs(amd, ctx):
    jobs = _jobs(9, 12)
    with amd.Batch(ctx, jobs, sweeps_per_launch=4) as a, amd.Batch(ctx, jobs, sweeps_per_launch=4) as b:
        a.run_all()
        while b.run(3) > 0:
            pass
        for j in range(len(jobs)):
            assert np.array_equal(a.spins(j), b.spins(j))
            assert np.array_equal(a.records(j), b.records(j))


@pytest.mark.parametrize("bad,msg", [(dict(L=12), "power of two"), (dict(L=32768), "[2, 16384]"),
                                     (dict(L=4), "n_levels"),
                                     (dict(n_levels=6), "n_levels"), (dict(coupling=-0.1), "coupling"),
                                     (dict(sweeps_per_bin=0), "sweeps_per_bin"),
                                     # one bin of 2^17 samples at L = 2048 could pass 2^63 (int64 records)
                                     (dict(L=2048, sweeps_per_bin=1 << 17), "2^63"),
                                     (dict(L=256, sweeps_per_bin=(1 << 31) - 1, n_meas=1), "2^63")])
def test_invalid_jobs(amd, ctx, bad, msg):
    job = dict(L=32, coupling=KC, seed=1, n_therm=1, n_meas=1, n_levels=2, sweeps_per_bin=1)
    job.update(bad)
    with pytest.raises(amd.MCRGError) as e:
        amd.Batch(ctx, [job])
    assert e.value.status == amd.ERR_INVALID_ARG and msg in str(e.value)


@pytest.mark.parametrize("parts", [-1, 65])
def test_invalid_sub_chunks(amd, ctx, parts):
    job = dict(L=32, coupling=KC, seed=1, n_therm=1, n_meas=1, n_levels=2, sweeps_per_bin=1)
    with pytest.raises(amd.MCRGError) as e:
        amd.Batch(ctx, [job], sub_chunks=parts)
    assert e.value.status == amd.ERR_INVALID_ARG and "sub_chunks" in str(e.value)


@pytest.mark.parametrize("threads_large", [0, 256, 512, 1024])
def test_launch_classes_mixed_lds(amd, ctx, orc, threads_large):
    """Slots of very different LDS size (L = 8 groups next to single L = 512
    lattices) launch as concurrent classes, with L = 512 on 256- or 512-thread
    workgroups; results stay bit-exact."""
    jobs = [dict(L=L, coupling=KC, seed=300 + i, n_therm=2, n_meas=5, n_levels=min(nl, int(math.log2(L)) - 1),
                 sweeps_per_bin=2)
            for i, (L, nl) in enumerate([(512, 7), (8, 1), (256, 6), (8, 0), (16, 2), (512, 3), (32, 3)] * 3)]
    print("classes", len(jobs), threads_large, flush=True)
    with amd.Batch(ctx, jobs, sweeps_per_launch=3, threads_large=threads_large) as b:
        b.run_all()
        _check_vs_oracle(orc, b, jobs)


@pytest.mark.parametrize("chunk", [250, 7])
def test_hbm_groups_bit_exact(amd, ctx, orc, chunk):
    """L = 2, 4 and L >= 1024 run on the HBM schedule, grouped by shape, beside
    resident jobs; every job equals the oracle (graph replays included when a
    chunk covers 32 sweeps or more)."""
    jobs = [dict(L=1024, coupling=KC, seed=501, n_therm=3, n_meas=40, n_levels=4, sweeps_per_bin=16),
            dict(L=1024, coupling=0.45, seed=502, n_therm=3, n_meas=40, n_levels=4, sweeps_per_bin=16),
            dict(L=32, coupling=KC, seed=503, n_therm=3, n_meas=40, n_levels=3, sweeps_per_bin=16),
            dict(L=4, coupling=KC, seed=504, n_therm=10, n_meas=70, n_levels=1, sweeps_per_bin=8),
            dict(L=2, coupling=0.3, seed=505, n_therm=0, n_meas=33, n_levels=0, sweeps_per_bin=33),
            dict(L=1024, coupling=0.44, seed=506, n_therm=35, n_meas=0, n_levels=1, sweeps_per_bin=1),
            dict(L=4, coupling=0.2, seed=507, n_therm=10, n_meas=70, n_levels=1, sweeps_per_bin=8)]
    print("hbm groups", chunk, flush=True)
    with amd.Batch(ctx, jobs, sweeps_per_launch=chunk) as b:
        b.run_all()
        b.sync()
        _check_vs_oracle(orc, b, jobs)


def test_swendsen_wang_jobs_in_batch(amd, ctx, orc):
    """Swendsen-Wang jobs run on the HBM schedule beside Metropolis jobs of the
    same L on the resident schedule; both equal the oracle."""
    jobs = [dict(L=64, coupling=KC, seed=700 + i, n_therm=4, n_meas=36, n_levels=3, sweeps_per_bin=12,
                 sampler=amd.SAMPLER_SWENDSEN_WANG if i % 2 else amd.SAMPLER_METROPOLIS) for i in range(6)]
    jobs.append(dict(L=16, coupling=0.6, seed=799, n_therm=0, n_meas=40, n_levels=2, sweeps_per_bin=40,
                     sampler=amd.SAMPLER_SWENDSEN_WANG))
    with amd.Batch(ctx, jobs, sweeps_per_launch=250) as b:
        b.run_all()
        b.sync()
        for j, job in enumerate(jobs):
            s_ref, r_ref = orc.run_chain(job["L"], job["coupling"], job["seed"], job["n_therm"], job["n_meas"],
                                         job["n_levels"], job["sweeps_per_bin"], sampler=job["sampler"])
            assert np.array_equal(b.spins(j), s_ref), job
            assert np.array_equal(b.records(j), r_ref), job


def test_bisection_batch_shape_matches_oracle(amd, ctx, orc):
    """The K-bisection's batch (configuration C5) at a few sweeps: L = 1024
    Swendsen-Wang int8 jobs blocked six times (1024 -> 16) beside L = 128
    bit-packed Metropolis jobs blocked three times (128 -> 16), at four
    couplings of the starting bracket [0.43, 0.45], in one batch.  Every
    job's spins and records equal the oracle's."""
    jobs = []
    for q in range(4):
        J = 0.43 + 0.02 * (q + 0.5) / 4
        jobs.append(dict(L=1024, coupling=J, seed=5000 + q, n_therm=2, n_meas=4, n_levels=6, sweeps_per_bin=2,
                         sampler=amd.SAMPLER_SWENDSEN_WANG))
        for r in range(4):
            jobs.append(dict(L=128, coupling=J, seed=6000 + 16 * q + r, n_therm=3, n_meas=4, n_levels=3,
                             sweeps_per_bin=2, sampler=amd.SAMPLER_METROPOLIS))
    print("bisection batch shape", len(jobs), flush=True)
    with amd.Batch(ctx, jobs) as b:
        b.run_all()
        b.sync()
        for j, job in enumerate(jobs):
            s_ref, r_ref = orc.run_chain(job["L"], job["coupling"], job["seed"], job["n_therm"], job["n_meas"],
                                         job["n_levels"], job["sweeps_per_bin"], sampler=job["sampler"])
            assert np.array_equal(b.spins(j), s_ref), job
            assert np.array_equal(b.records(j), r_ref), job


def test_autotune_keeps_results_and_picks_a_candidate(amd, ctx, orc):
    """The autotuner runs its candidate configurations on the first launches of
    the real jobs; every job still equals the oracle, and it settles on one of
    the candidates."""
    jobs = _jobs(31, 40, Ls=(8, 16, 32, 64, 128))
    jobs = [dict(j, n_meas=j["n_meas"] + 20) for j in jobs]  # long enough for all candidates
    with amd.Batch(ctx, jobs, sweeps_per_launch=2, autotune=True) as b:
        assert b.options()["autotune"] == 1
        b.run_all()
        b.sync()
        opts = b.options()
        assert opts["autotune"] == 0 and opts["threads"] in (128, 256, 512)
        _check_vs_oracle(orc, b, jobs)
