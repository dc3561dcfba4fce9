"""Configurations at their full size, checked through properties that do not
need the oracle (which would take hours on them).

Batching: 4096 replicas of
mcrg_amd.workloads.batch_manifest (L from {8..128}, five couplings, seeds
0..4095), 5e3 thermalisation and 5e4 measured sweeps each, bins of 1e3.

* every replica's records count exactly 5e4 samples in 50 bins;
* replicas run again alone give records bit-identical to the ones they got
  in the mix;
* the mean nearest-neighbour correlation of the L = 8, 16 and 32 replicas at
  each coupling equals the exact finite-lattice value from Kaufman's
  partition function (tests/exact_ising.py) within 5 standard errors of the
  replica spread; those lattices equilibrate well within 5e3 sweeps.
"""
import math

import numpy as np
import pytest

from exact_ising import nn_mean

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def amd():
    import mcrg_amd
    return mcrg_amd


@pytest.fixture(scope="module")
def ctx(amd):
    c = amd.Context(0)
    yield c
    c.close()


def test_batching_config_full_size(amd, ctx):
    from mcrg_amd import workloads
    jobs = workloads.batch_manifest(4096)
    assert jobs[0]["n_therm"] == 5000 and jobs[0]["n_meas"] == 50000 and jobs[0]["sweeps_per_bin"] == 1000
    print("full batching configuration: 4096 replicas x 55000 sweeps", flush=True)
    with amd.Batch(ctx, jobs) as b:
        b.run_all()
        recs = [b.records(j) for j in range(len(jobs))]
    # sample counts
    for j, (job, r) in enumerate(zip(jobs, recs)):
        nl = job["n_levels"]
        assert r.shape == (50, amd.record_size(nl)), j
        assert np.all(r[:, amd.record_size(nl) - 1] == 1000), j
    # alone against in the mix: one replica of each size plus two more
    picks = [next(j for j, job in enumerate(jobs) if job["L"] == L) for L in (8, 16, 32, 64, 128)] + [1234, 4095]
    print("alone:", [(j, jobs[j]["L"]) for j in picks], flush=True)
    with amd.Batch(ctx, [jobs[j] for j in picks]) as b:
        b.run_all()
        for k, j in enumerate(picks):
            assert np.array_equal(b.records(k), recs[j]), j
    # physics of the small lattices
    for L in (8, 16, 32):
        for K in workloads.BATCH_COUPLINGS:
            sel = [j for j, job in enumerate(jobs) if job["L"] == L and job["coupling"] == K]
            per = np.array([recs[j][:, 0].sum() / 50000.0 / (L * L) for j in sel])
            mean, err = per.mean(), per.std(ddof=1) / math.sqrt(len(per))
            exact = nn_mean(L, K) / (L * L)
            print(f"L={L} K={K}: {len(sel)} replicas, NN/site {mean:.6f} +- {err:.6f}, exact {exact:.6f}", flush=True)
            assert abs(mean - exact) < 5 * err, (L, K, mean, err, exact)


def test_plain_config_l2048_full_size(amd, ctx):
    """The plain single-lattice configuration at full size: L = 2048,
    |K| = 0.4406868, seed 7, 2e3 thermalisation and 2e4 measured sweeps, five
    blockings.  The chain entry point (graph replays of 32 sweeps) and the same
    job inside a batch (an HBM-schedule group advanced in launches of 250
    sweeps) give bit-identical spins and records; every bin counts its samples
    (the first 100 sweeps are checked against the oracle in test_gpu_chain)."""
    L, J, seed, nl, spb = 2048, 0.4406868, 7, 5, 1000
    print("plain configuration L=2048: 2e3 + 2e4 sweeps, chain and batch", flush=True)
    with amd.Chain(ctx, L, J, seed, n_levels=nl, sweeps_per_bin=spb, max_bins=20) as ch:
        ch.sweep(2000)
        ch.sweep(20000, measure=True)
        rec_chain, spins_chain = ch.records(), ch.spins()
    assert rec_chain.shape == (20, amd.record_size(nl))
    assert np.all(rec_chain[:, -1] == spb)
    job = dict(L=L, coupling=J, seed=seed, n_therm=2000, n_meas=20000, n_levels=nl, sweeps_per_bin=spb)
    with amd.Batch(ctx, [job]) as b:
        b.run_all()
        assert np.array_equal(b.records(0), rec_chain)
        assert np.array_equal(b.spins(0), spins_chain)
