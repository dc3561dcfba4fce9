"""GPU parity of the plain (HBM, int8) chain path against the CPU oracle.

Integer/bit work, so the bar is bit-exact: spins after every checked sweep and
every int64 record entry.  Cases cover the smallest lattices (rows shorter than
a 32-site word), the word-aligned fast path, both colours' boundary bytes, the
L = 2048 configuration's first 100 sweeps, chunking and checkpoint/resume.
"""
import math
import sys

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


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("L", [2, 4, 8, 16, 32, 64, 128, 256, 2048])
def test_init_matches_oracle(amd, ctx, orc, L):
    _say("init", L)
    with amd.Chain(ctx, L, KC, seed=12345 + L) as ch:
        assert np.array_equal(ch.spins(), orc.init_spins(L, 12345 + L))


@pytest.mark.parametrize("L,J", [(2, KC), (4, KC), (8, KC), (8, 0.0), (16, 0.3), (32, 0.489652),
                                 (64, KC), (128, 0.4320459), (256, 2.5), (512, KC)])
def test_sweeps_bit_exact(amd, ctx, orc, L, J):
    seed = 0xDEADBEEF00000000 + L
    _say("sweeps", L, J)
    with amd.Chain(ctx, L, J, seed=seed) as ch:
        s = orc.init_spins(L, seed)
        for t in range(12):
            ch.sweep(1)
            orc.sweep(L, s, seed, J, t)
            assert np.array_equal(ch.spins(), s), f"L={L} J={J} diverged at sweep {t}"


def test_config2_first_100_sweeps(amd, ctx, orc):
    """L = 2048, |K| = 0.4406868, seed 7: bit-exact for sweeps 1..100."""
    L, J, seed = 2048, 0.4406868, 7
    _say("L=2048 first 100 sweeps")
    s = orc.init_spins(L, seed)
    with amd.Chain(ctx, L, J, seed=seed) as ch:
        assert np.array_equal(ch.spins(), s)
        done = 0
        for upto in (1, 10, 100):
            for t in range(done, upto):
                orc.sweep(L, s, seed, J, t)
            ch.sweep(upto - done)
            done = upto
            assert np.array_equal(ch.spins(), s), f"diverged by sweep {upto}"


def test_config2_first_100_measured_sweeps_five_levels(amd, ctx, orc):
    """The plain configuration's shape: L = 2048, |K| = 0.4406868, seed 7, five
    blockings (2048 -> 64).  Its first 100 sweeps all measured, in bins of 25,
    enqueued as runs of 1, 9 and 90 sweeps: spins and every record bin equal
    the oracle's bit for bit."""
    L, J, seed, nl, spb = 2048, 0.4406868, 7, 5, 25
    _say("L=2048 first 100 measured sweeps, 5 levels")
    s_ref, rec_ref = orc.run_chain(L, J, seed, 0, 100, nl, spb)
    with amd.Chain(ctx, L, J, seed=seed, n_levels=nl, sweeps_per_bin=spb, max_bins=4) as ch:
        for n in (1, 9, 90):
            ch.sweep(n, measure=True)
        assert ch.state() == (100, 100)
        assert np.array_equal(ch.spins(), s_ref)
        got = ch.records()
        assert np.array_equal(got, rec_ref), np.argwhere(got != rec_ref)[:10]


@pytest.mark.parametrize("L,nl", [(64, 5), (1024, 4), (4096, 6)])
def test_strip_and_hbm_schedules_agree(amd, ctx, L, nl, monkeypatch):
    """Chains of 64 <= L <= 8192 run on the strip schedule (strips held in LDS
    by a cooperative launch); MCRG_STRIP=0 keeps them on the per-half-sweep
    HBM schedule.  Both give the same spins, counters and records after
    unmeasured and measured runs of several lengths."""
    seed, spb = 31 + L, 6
    _say("strip vs hbm", L, nl)
    out = []
    for flag in ("1", "0"):
        monkeypatch.setenv("MCRG_STRIP", flag)
        with amd.Chain(ctx, L, KC, seed=seed, n_levels=nl, sweeps_per_bin=spb, max_bins=4) as ch:
            ch.sweep(3)
            ch.sweep(7, measure=True)
            ch.sweep(2)
            ch.sweep(13, measure=True)
            out.append((ch.spins(), ch.records(), ch.state()))
    assert out[0][2] == out[1][2] == (25, 20)
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("L,nl", [(64, 1), (256, 4), (512, 5), (2048, 5)])
def test_strip_halo_rows_agree(amd, ctx, orc, L, nl, monkeypatch):
    """Strips of at most 512 colour-1 words keep two halo rows a side and
    exchange once per sweep; measured launches of strips of at most 320 of
    them add a measuring workgroup per strip.  MCRG_STRIP_SPLIT=0 keeps the
    measurements in the strip's workgroup, MCRG_STRIP_HALO=0 keeps two
    exchanges per sweep.  Runs of 1, 2, 3 (unmeasured) and 5 sweeps cover
    the shortest launches of the one-sweep-late measurement.  All three equal
    the oracle."""
    seed, spb = 77 + L, 4
    _say("strip halo rows", L, nl)
    s_ref, rec_ref = orc.run_chain(L, KC, seed, 0, 1, nl, spb)
    out = []
    for halo, split in (("1", "1"), ("1", "0"), ("0", "1")):
        monkeypatch.setenv("MCRG_STRIP_HALO", halo)
        monkeypatch.setenv("MCRG_STRIP_SPLIT", split)
        _say("  halo", halo, "split", split)
        with amd.Chain(ctx, L, KC, seed=seed, n_levels=nl, sweeps_per_bin=spb, max_bins=3) as ch:
            ch.sweep(1, measure=True)
            assert np.array_equal(ch.spins(), s_ref)
            assert np.array_equal(ch.records()[:1], rec_ref)
            ch.sweep(2, measure=True)
            ch.sweep(3)
            ch.sweep(5, measure=True)
            out.append((ch.spins(), ch.records(), ch.state()))
    for o in out[1:]:
        assert o[2] == out[0][2] == (11, 8)
        assert np.array_equal(o[0], out[0][0])
        assert np.array_equal(o[1], out[0][1])


@pytest.mark.parametrize("L,nl,J", [(4, 1, KC), (8, 2, KC), (16, 3, 0.42), (32, 4, KC),
                                    (64, 5, 0.4496804), (128, 6, KC), (256, 5, KC)])
def test_measurement_records_bit_exact(amd, ctx, orc, L, nl, J):
    seed = 99 + L
    n_therm, n_meas, spb = 5, 23, 7
    n_bins = -(-n_meas // spb)
    _say("records", L, nl)
    s_ref, rec_ref = orc.run_chain(L, J, seed, n_therm, n_meas, nl, spb)
    with amd.Chain(ctx, L, J, seed=seed, n_levels=nl, sweeps_per_bin=spb, max_bins=n_bins) as ch:
        ch.sweep(n_therm)
        ch.sweep(n_meas, measure=True)
        assert np.array_equal(ch.spins(), s_ref)
        got = ch.records()
        assert np.array_equal(got, rec_ref), np.argwhere(got != rec_ref)[:10]
        assert ch.state() == (n_therm + n_meas, n_meas)


def test_l16384_all_levels_record_bit_exact(amd, ctx, orc):
    """The largest lattice with every blocking (16384 -> 2: five grid-measured
    levels, then eight in one workgroup): one measured sweep's record equals
    the oracle's."""
    L, J, seed, nl = 16384, 0.4406868, 5, 13
    _say("L=16384 13 levels, one measured sweep")
    s_ref, rec_ref = orc.run_chain(L, J, seed, 0, 1, nl, 1)
    with amd.Chain(ctx, L, J, seed=seed, n_levels=nl, sweeps_per_bin=1, max_bins=1) as ch:
        ch.sweep(1, measure=True)
        assert np.array_equal(ch.records(), rec_ref)
        assert np.array_equal(ch.spins(), s_ref)


def test_l16384_strip_sweeps_bit_exact(amd, ctx, orc):
    """L = 16384 (256 MiB of int8 spins) updates strips of 8 rows per thread;
    two sweeps equal the oracle's."""
    L, J, seed = 16384, 0.4406868, 99
    _say("L=16384 two sweeps")
    s = orc.init_spins(L, seed)
    with amd.Chain(ctx, L, J, seed=seed) as ch:
        for t in range(2):
            orc.sweep(L, s, seed, J, t)
        ch.sweep(2)
        assert np.array_equal(ch.spins(), s)


@pytest.mark.parametrize("L,nl,J", [(512, 8, KC), (1024, 5, 0.4688157), (2048, 5, KC), (4096, 6, 0.4320459),
                                    (8192, 7, KC)])
def test_large_lattice_records_bit_exact(amd, ctx, orc, L, nl, J):
    """Levels of side >= 512 are measured by grid launches, the rest by one
    workgroup: L = 1024 has one grid level, 2048 two, 4096 three, 8192 four."""
    seed = 7 + L
    n_therm, n_meas, spb = 2, 5, 2
    _say("records (large)", L, nl)
    s_ref, rec_ref = orc.run_chain(L, J, seed, n_therm, n_meas, nl, spb)
    with amd.Chain(ctx, L, J, seed=seed, n_levels=nl, sweeps_per_bin=spb, max_bins=3) as ch:
        ch.sweep(n_therm)
        ch.sweep(n_meas, measure=True)
        assert np.array_equal(ch.spins(), s_ref)
        got = ch.records()
        assert np.array_equal(got, rec_ref), np.argwhere(got != rec_ref)[:10]


@pytest.mark.parametrize("L", [2, 4, 16, 32, 64, 512])
def test_operators_match_oracle(amd, ctx, orc, L):
    seed = 5 + L
    with amd.Chain(ctx, L, KC, seed=seed) as ch:
        ch.sweep(3)
        s = ch.spins()
        ops = ch.operators()
        assert list(ops[:3]) == list(orc.ops(L, s))
        assert ops[3] == int(s.astype(np.int64).sum())


def test_chunking_is_invisible(amd, ctx):
    L, seed = 128, 31
    with amd.Chain(ctx, L, KC, seed=seed, n_levels=4, sweeps_per_bin=5, max_bins=4) as a, \
            amd.Chain(ctx, L, KC, seed=seed, n_levels=4, sweeps_per_bin=5, max_bins=4) as b:
        a.sweep(6)
        a.sweep(20, measure=True)
        for _ in range(6):
            b.sweep(1)
        for _ in range(20):
            b.sweep(1, measure=True)
        assert np.array_equal(a.spins(), b.spins())
        assert np.array_equal(a.records(), b.records())


def test_checkpoint_resume(amd, ctx):
    L, seed, nl = 256, 4242, 3
    with amd.Chain(ctx, L, KC, seed=seed, n_levels=nl, sweeps_per_bin=10, max_bins=3) as a:
        a.sweep(15)
        a.sweep(30, measure=True)
        want_s, want_r = a.spins(), a.records()
    with amd.Chain(ctx, L, KC, seed=seed, n_levels=nl, sweeps_per_bin=10, max_bins=3) as b:
        b.sweep(15)
        b.sweep(12, measure=True)
        snap = (b.spins(), b.state(), b.records())
    with amd.Chain(ctx, L, KC, seed=seed, n_levels=nl, sweeps_per_bin=10, max_bins=3) as c:
        c.set_spins(snap[0])
        c.set_state(snap[1][0], snap[1][1], snap[2])
        c.sweep(18, measure=True)
        assert np.array_equal(c.spins(), want_s)
        assert np.array_equal(c.records(), want_r)


def test_set_coupling_matches_oracle(amd, ctx, orc):
    L, seed = 64, 8
    s = orc.init_spins(L, seed)
    with amd.Chain(ctx, L, 0.3, seed=seed) as ch:
        ch.sweep(4)
        for t in range(4):
            orc.sweep(L, s, seed, 0.3, t)
        ch.set_coupling(0.5)
        ch.sweep(4)
        for t in range(4, 8):
            orc.sweep(L, s, seed, 0.5, t)
        assert np.array_equal(ch.spins(), s)


@pytest.mark.parametrize("kw,msg", [
    (dict(L=12), "power of two"), (dict(L=1), "power of two"), (dict(coupling=-0.44), "coupling"),
    (dict(n_levels=7, L=64), "n_levels"), (dict(sweeps_per_bin=0), "sweeps_per_bin"),
    (dict(L=4096, sweeps_per_bin=8192), "2^63"),  # max_sweeps_per_bin(4096) = 8191
])
def test_invalid_chain_arguments(amd, ctx, kw, msg):
    args = dict(L=64, coupling=KC, seed=1, n_levels=1)
    args.update(kw)
    with pytest.raises(amd.MCRGError) as e:
        amd.Chain(ctx, **args)
    assert e.value.status == amd.ERR_INVALID_ARG
    assert msg in str(e.value)


def test_record_capacity_error(amd, ctx):
    with amd.Chain(ctx, 16, KC, seed=1, n_levels=1, sweeps_per_bin=4, max_bins=2) as ch:
        ch.sweep(8, measure=True)
        with pytest.raises(amd.MCRGError) as e:
            ch.sweep(1, measure=True)
        assert e.value.status == amd.ERR_CAPACITY


def test_graph_replays_match_oracle(amd, ctx, orc):
    """Runs of 32 sweeps or more replay a captured graph.  Several replays,
    with eager reads, eager short runs and a coupling change between them,
    each checked against the oracle (spins, counters and every record bin)."""
    L, nl, seed, spb = 64, 3, 4242, 20
    _say("graph replays", L)
    s = orc.init_spins(L, seed)
    rec_ref = np.zeros((8, orc.record_size(nl)), dtype=np.int64)
    t = q = 0
    with amd.Chain(ctx, L, KC, seed, n_levels=nl, sweeps_per_bin=spb, max_bins=8) as ch:
        for n, meas, J in [(64, False, KC), (40, True, KC), (5, True, KC), (70, True, 0.42), (33, False, 0.42)]:
            ch.set_coupling(J)
            ch.sweep(n, measure=meas)
            for _ in range(n):
                if meas:
                    s, r = orc.run_chain(L, J, seed, 0, 1, nl, 1, s=s, t0=t)
                    rec_ref[q // spb] += r[0]
                    q += 1
                else:
                    orc.sweep(L, s, seed, J, t)
                t += 1
            assert np.array_equal(ch.spins(), s), (n, meas, J)
            assert ch.state() == (t, q)
            assert np.array_equal(ch.records(), rec_ref), (n, meas, J)


@pytest.mark.parametrize("L,J,n_therm,n_meas,nl", [(2, KC, 3, 5, 0), (4, 0.3, 2, 6, 1), (8, KC, 1, 40, 2),
                                                   (32, KC, 5, 9, 3), (64, 0.5, 33, 3, 4), (256, KC, 2, 4, 5),
                                                   (1024, KC, 1, 2, 4)])
def test_swendsen_wang_bit_exact(amd, ctx, orc, L, J, n_therm, n_meas, nl):
    """Swendsen-Wang chains: spins and records equal the oracle's union-find
    restatement (cluster roots are the smallest site index, so the partition
    and every flip are schedule-independent)."""
    seed = 0x5EED0000 + L
    _say("swendsen-wang", L, J)
    assert amd.bond_threshold(J) == orc.sw_threshold(J)
    spb = max(1, n_meas // 2)
    nb = amd.n_bins(n_meas, spb)
    with amd.Chain(ctx, L, J, seed, n_levels=nl, sweeps_per_bin=spb, max_bins=nb,
                   sampler=amd.SAMPLER_SWENDSEN_WANG) as ch:
        ch.sweep(n_therm)
        ch.sweep(n_meas, measure=True)
        s_ref, r_ref = orc.run_chain(L, J, seed, n_therm, n_meas, nl, spb, sampler=orc.SAMPLER_SW)
        assert np.array_equal(ch.spins(), s_ref)
        assert np.array_equal(ch.records(), r_ref)
