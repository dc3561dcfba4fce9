"""The allocation-free sweep (mcrg_chain_sweep_static) against the entry point
it varies (mcrg_chain_sweep): the same spins, counters and every record bin,
bit for bit, on two calls in a row, the second with a smaller lattice; the
device's free memory does not change across its calls; and chains it cannot
run are rejected with a clear error."""
import ctypes
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


def _run(amd, ctx, L, nl, seed, runs, static):
    with amd.Chain(ctx, L, KC, seed=seed, n_levels=nl, sweeps_per_bin=8, max_bins=8) as ch:
        for n, meas in runs:
            (ch.sweep_static if static else ch.sweep)(n, measure=meas)
        return ch.spins(), ch.records(), ch.state()


@pytest.mark.parametrize("L,nl,seed,runs", [
    (2048, 5, 11, [(3, False), (20, True), (1, True)]),   # first call: L = 2048, five blockings
    (256, 4, 12, [(2, False), (9, True)]),                 # then a smaller lattice
])
def test_static_sweep_matches_the_allocating_entry_point(amd, ctx, L, nl, seed, runs):
    print("static sweep", L, nl, runs, flush=True)
    s0, r0, st0 = _run(amd, ctx, L, nl, seed, runs, static=False)
    s1, r1, st1 = _run(amd, ctx, L, nl, seed, runs, static=True)
    assert st0 == st1
    assert np.array_equal(s0, s1)
    assert np.array_equal(r0, r1)


def _free_bytes():
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def test_static_sweep_leaves_device_memory_alone(amd, ctx):
    """A fresh chain's static sweeps: the free device memory reported by the
    runtime is the same before and after each (a workspace allocation of the
    allocating path, about 0.8 MiB at L = 2048 with five levels, is not made)."""
    with amd.Chain(ctx, 512, KC, seed=5, n_levels=3, sweeps_per_bin=8, max_bins=2) as warm:
        warm.sweep_static(4, measure=True)  # loads the kernels
        warm.sync()
    with amd.Chain(ctx, 2048, KC, seed=6, n_levels=5, sweeps_per_bin=8, max_bins=4) as ch:
        ch.sync()
        before = _free_bytes()
        for n in (5, 7, 9):
            ch.sweep_static(n, measure=True)
            ch.sync()
            assert _free_bytes() == before
        assert ch.state() == (21, 21)


@pytest.mark.parametrize("L,nl,measure,sampler", [(32, 2, True, 0), (16384, 1, True, 0), (1024, 2, False, 1)])
def test_static_sweep_rejects_other_chains(amd, ctx, L, nl, measure, sampler):
    with amd.Chain(ctx, L, KC, seed=1, n_levels=nl, sweeps_per_bin=1, max_bins=4, sampler=sampler) as ch:
        with pytest.raises(amd.MCRGError) as e:
            ch.sweep_static(2, measure=measure)
        assert e.value.status == amd.ERR_INVALID_ARG
        assert "strip schedule" in str(e.value)
