"""Pins for the CPU oracle (oracle/oracle.c) that need no GPU.

Each check is chosen so that a plausible slip in the oracle (a wrong neighbour,
a dropped operator term, a swapped flatten order, a wrong acceptance class)
fails it: published known-answer vectors, a library routine (rocRAND, numpy),
closed forms, and exact enumeration of small lattices.
"""
import itertools
import json
import math
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
KC = 0.5 * math.log(1.0 + math.sqrt(2.0))  # 0.44068679350977...


# ------------------------------------------------------------------ Philox
# Random123 known-answer vectors for philox4x32-10 (kat_vectors, R = 10):
# (ctr0..3, key0..1) -> out0..3.
RANDOM123_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,want", RANDOM123_KAT)
def test_philox_random123_kat(orc, ctr, key, want):
    seed = key[0] | (key[1] << 32)
    assert tuple(int(v) for v in orc.philox(seed, ctr)) == want


def test_philox_matches_rocrand(orc):
    vecs = json.load(open(os.path.join(GOLDEN, "philox_rocrand.json")))
    assert len(vecs) >= 8
    for v in vecs:
        assert [int(x) for x in orc.philox(v["seed"], v["ctr"])] == v["out"], v


# ------------------------------------------------------------------ thresholds
def test_thresholds_closed_form_at_kc(orc):
    # exp(-2 Kc) = sqrt(2) - 1 exactly, so exp(-4 Kc) = 3 - 2 sqrt 2 and
    # exp(-8 Kc) = 17 - 12 sqrt 2.
    thr = orc.thresholds(KC)
    p4 = 3.0 - 2.0 * math.sqrt(2.0)
    p8 = 17.0 - 12.0 * math.sqrt(2.0)
    assert abs(int(thr[0]) - p4 * 2**32) <= 2.0
    assert abs(int(thr[1]) - p8 * 2**32) <= 2.0


def test_thresholds_edges(orc):
    assert list(orc.thresholds(0.0)) == [2**32, 2**32]  # infinite temperature: always flip
    assert list(orc.thresholds(50.0)) == [0, 0]


# ------------------------------------------------------------------ init
def test_init_balance_and_determinism(orc):
    a = orc.init_spins(256, 7)
    b = orc.init_spins(256, 7)
    c = orc.init_spins(256, 8)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)
    assert set(np.unique(a)) <= {-1, 1}
    n = a.size
    assert abs(int(a.sum())) < 5 * math.sqrt(n)  # binomial balance


# ------------------------------------------------------------------ operators
def _ops_numpy(L, s):
    g = s.reshape(L, L).astype(np.int64)  # g[y, x]
    right = np.roll(g, -1, axis=1)
    down = np.roll(g, -1, axis=0)
    diag = np.roll(right, -1, axis=0)
    return np.array([(g * right + g * down).sum(), (g * diag + right * down).sum(),
                     (g * right * down * diag).sum()])


@pytest.mark.parametrize("L", [2, 4, 8, 16])
def test_ops_closed_forms(orc, L):
    n = L * L
    up = np.ones(n, dtype=np.int8)
    y, x = np.divmod(np.arange(n), L)
    neel = np.where((x + y) % 2 == 0, 1, -1).astype(np.int8)
    stripes = np.where(x % 2 == 0, 1, -1).astype(np.int8)
    assert list(orc.ops(L, up)) == [2 * n, 2 * n, n]
    assert list(orc.ops(L, -up)) == [2 * n, 2 * n, n]
    assert list(orc.ops(L, neel)) == [-2 * n, 2 * n, n]
    assert list(orc.ops(L, stripes)) == [0, -2 * n, n]


@pytest.mark.parametrize("L", [2, 4, 8, 32])
def test_ops_bruteforce_and_refcount(orc, L):
    rng = np.random.default_rng(L)
    for _ in range(5):
        s = rng.choice(np.array([-1, 1], dtype=np.int8), size=L * L)
        got = orc.ops(L, s)
        assert list(got) == list(_ops_numpy(L, s))
        # The reference's calc_interactions counts every bond twice.
        ref = orc.ops_refcount(L, s)
        assert ref[0] == 2 * got[0] and ref[1] == 2 * got[1]


# ------------------------------------------------------------------ blocking
def test_block_all_16_patterns(orc):
    # A 2x2 lattice is a single b=2 block.  sum > 0 -> +1, < 0 -> -1, 0 -> coin.
    for bits in itertools.product([1, -1], repeat=4):
        s = np.array(bits, dtype=np.int8)
        tot = int(s.sum())
        outs = {int(orc.block(2, s, seed=3, t=t, level=0)[0]) for t in range(64)}
        if tot > 0:
            assert outs == {1}
        elif tot < 0:
            assert outs == {-1}
        else:
            assert outs == {1, -1}


def test_block_tie_coin_fair(orc):
    s = np.array([1, -1, -1, 1], dtype=np.int8)
    v = np.array([orc.block(2, s, seed=11, t=t, level=0)[0] for t in range(20000)])
    assert abs(v.mean()) < 5 / math.sqrt(v.size)


def test_block_matches_reshape(orc):
    L = 16
    rng = np.random.default_rng(5)
    s = rng.choice(np.array([-1, 1], dtype=np.int8), size=L * L)
    g = s.reshape(L, L).astype(int)
    tot = g.reshape(L // 2, 2, L // 2, 2).sum(axis=(1, 3)).reshape(-1)
    sb = orc.block(L, s, seed=9, t=4, level=2)
    assert np.all(sb[tot > 0] == 1) and np.all(sb[tot < 0] == -1)
    assert set(np.unique(sb[tot == 0])) <= {-1, 1}


# ------------------------------------------------------------------ exact 4x4 ensemble
def _enumerate(L):
    n = L * L
    idx = np.arange(2**n, dtype=np.int64)
    bits = (idx[:, None] >> np.arange(n)[None, :]) & 1
    return (1 - 2 * bits).astype(np.int64).reshape(-1, L, L)  # [state, y, x]


def _ops_batch(g):
    right = np.roll(g, -1, axis=2)
    down = np.roll(g, -1, axis=1)
    diag = np.roll(right, -1, axis=1)
    nn = (g * right + g * down).sum(axis=(1, 2))
    nnn = (g * diag + right * down).sum(axis=(1, 2))
    p = (g * right * down * diag).sum(axis=(1, 2))
    return np.stack([nn, nnn, p], axis=1).astype(np.float64)


def _exact_L4(J):
    g = _enumerate(4)
    S0 = _ops_batch(g)
    w = np.exp(J * S0[:, 0] - np.max(J * S0[:, 0]))
    w /= w.sum()
    tot = g.reshape(-1, 2, 2, 2, 2).sum(axis=(2, 4))  # [state, by, bx]
    # average over the 16 equally likely coin patterns (only ties use them)
    acc = {"S1": 0.0, "S1S1": 0.0, "S1S0": 0.0}
    for coins in itertools.product([1, -1], repeat=4):
        c = np.array(coins).reshape(2, 2)
        b = np.where(tot > 0, 1, np.where(tot < 0, -1, c[None]))
        S1 = _ops_batch(b)
        acc["S1"] = acc["S1"] + (w[:, None] * S1).sum(0) / 16
        acc["S1S1"] = acc["S1S1"] + np.einsum("s,sa,sb->ab", w, S1, S1) / 16
        acc["S1S0"] = acc["S1S0"] + np.einsum("s,sa,sb->ab", w, S1, S0) / 16
    S0m = (w[:, None] * S0).sum(0)
    M = g.sum(axis=(1, 2)).astype(np.float64)
    return {"S0": S0m, "S1": acc["S1"], "S1S1": acc["S1S1"], "S1S0": acc["S1S0"],
            "M2": (w * M * M).sum()}


def _exact_T(ex, n_ops):
    A = ex["S1S1"] - np.outer(ex["S1"], ex["S1"])
    B = ex["S1S0"] - np.outer(ex["S1"], ex["S0"])
    return np.linalg.solve(A[:n_ops, :n_ops], B[:n_ops, :n_ops])


@pytest.mark.parametrize("J,sampler", [(KC, 0), (0.3, 0), (KC, 1), (0.6, 1)])
def test_sampler_samples_exact_4x4_ensemble(orc, J, sampler):
    """The sweep's stationary distribution is Boltzmann (Metropolis = 0,
    Swendsen-Wang = 1): MC means of every operator and of M^2 agree with exact
    enumeration of all 2^16 states."""
    ex = _exact_L4(J)
    nl = 1
    bin_len = 500
    _, rec = orc.run_chain(4, J, seed=2024, n_therm=200, n_meas=100000, nl=nl,
                           sweeps_per_bin=bin_len, sampler=sampler)
    n = rec[:, orc.off_N(nl)].astype(float)
    for a in range(3):
        means = rec[:, orc.off_S(nl, 0, a)] / n
        err = means.std(ddof=1) / math.sqrt(len(means))
        assert abs(means.mean() - ex["S0"][a]) < 5 * err + 1e-12, (a, means.mean(), ex["S0"][a], err)
    m2 = rec[:, orc.off_M(nl) + 2] / n
    err = m2.std(ddof=1) / math.sqrt(len(m2))
    assert abs(m2.mean() - ex["M2"]) < 5 * err


@pytest.mark.parametrize("J", [0.1, KC])
def test_uphill_flip_rate_is_the_threshold(orc, J):
    """From the all-up lattice every colour-0 site has e = 4, so one colour-0
    half-sweep flips each with probability exactly thr[1] / 2^32, through the
    4 top bits and, for the 1/16 of sites they leave open, the refinement
    bits.  A comparison that mishandles the open sites (accepting or rejecting
    them all, or reading the wrong threshold bits) moves the rate by a
    sizeable fraction of 1/16, many standard errors at this sample size."""
    L, n_t = 64, 600
    p = float(orc.thresholds(J)[1]) / 2.0**32
    flips = 0
    for t in range(n_t):
        s = np.ones(L * L, dtype=np.int8)
        orc.half_sweep(L, s, 7, J, t, 0)
        flips += int((s < 0).sum())
    n = n_t * L * L // 2
    err = math.sqrt(p * (1.0 - p) / n)
    assert abs(flips / n - p) < 5 * err, (flips / n, p, err)


def test_rg_matrix_converges_to_exact_4x4(orc):
    """The whole measurement chain (operators, blocking with coins, record layout,
    connected correlators, T = A^-1 B, leading eigenvalue) reproduces the exact
    L = 4 -> 2 RG matrix obtained by enumeration."""
    ex = _exact_L4(KC)
    nl = 1
    _, rec = orc.run_chain(4, KC, seed=77, n_therm=200, n_meas=200000, nl=nl, sweeps_per_bin=5000)
    for n_ops in (2, 3):
        lam_exact = max(np.linalg.eigvals(_exact_T(ex, n_ops)).real)
        mean, err = orc.jackknife_yt(rec, nl, 0, n_ops)
        yt_exact = math.log(lam_exact) / math.log(2.0)
        assert abs(mean - yt_exact) < 5 * err, (n_ops, mean, err, yt_exact)


# ------------------------------------------------------------------ records
def test_record_layout_against_per_sample_sums(orc):
    """Recompute every record entry with Python integers from per-sample operator
    values (S(n) S(n-1)^T flattened column-major, as src/definitions.cpp:9-19;
    NN(n) NN(0) as src/mcrg.cpp:262-263)."""
    L, nl, J, seed = 8, 2, KC, 5
    s0 = orc.init_spins(L, seed)
    _, rec = orc.run_chain(L, J, seed, n_therm=3, n_meas=7, nl=nl, sweeps_per_bin=100, s=s0)
    rec = rec[0]
    s = s0.copy()
    want = [0] * orc.record_size(nl)
    for t in range(10):
        orc.sweep(L, s, seed, J, t)
        if t < 3:
            continue
        levels = [s.copy()]
        for l in range(nl):
            levels.append(orc.block(L >> l, levels[-1], seed, t, l))
        S = [[int(v) for v in orc.ops(L >> l, levels[l])] for l in range(nl + 1)]
        for l in range(nl + 1):
            for a in range(3):
                want[orc.off_S(nl, l, a)] += S[l][a]
                for b in range(3):
                    want[orc.off_self(nl, l, a, b)] += S[l][a] * S[l][b]
                    if l >= 1:
                        want[orc.off_cross(nl, l, a, b)] += S[l][a] * S[l - 1][b]
            if l >= 1:
                want[orc.off_nn0(nl, l)] += S[l][0] * S[0][0]
        M = int(s.astype(np.int64).sum())
        want[orc.off_M(nl)] += M
        want[orc.off_M(nl) + 1] += abs(M)
        want[orc.off_M(nl) + 2] += M * M
        want[orc.off_N(nl)] += 1
    assert [int(v) for v in rec] == want


def test_chain_equals_single_sweeps(orc):
    L, seed, J = 16, 99, 0.42
    s = orc.init_spins(L, seed)
    s_chain, _ = orc.run_chain(L, J, seed, n_therm=12, n_meas=0, nl=1, sweeps_per_bin=1, s=s)
    for t in range(12):
        orc.sweep(L, s, seed, J, t)
    assert np.array_equal(s, s_chain)


# ------------------------------------------------------------------ eigen-solve
def test_leading_eigenvalue_against_numpy(orc):
    rng = np.random.default_rng(0)
    for n in (2, 3):
        for _ in range(300):
            T = rng.normal(size=(n, n))
            want = max(np.linalg.eigvals(T).real)
            assert abs(orc.leading_eigenvalue(T) - want) < 1e-9 * max(1.0, abs(want))


def test_leading_eigenvalue_special_cases(orc):
    assert orc.leading_eigenvalue(np.diag([2.0, 3.0])) == pytest.approx(3.0)
    assert orc.leading_eigenvalue(np.array([[0.0, -1.0], [1.0, 0.0]])) == pytest.approx(0.0)
    rot = np.array([[1.0, -2.0, 0.0], [2.0, 1.0, 0.0], [0.0, 0.0, 0.5]])  # 1 +- 2i, 0.5
    assert orc.leading_eigenvalue(rot) == pytest.approx(1.0)
    assert orc.leading_eigenvalue(np.diag([1.5, 1.5, 1.5])) == pytest.approx(1.5)


def test_split_samples(orc):
    assert orc.split_samples(0, 8, 10) == -4  # the reference's quirk is kept
    assert orc.split_samples(3, 8, 10) == 2
    assert sum(orc.split_samples(r, 4, 1000003) for r in range(4)) == 1000003


# ------------------------------------------------------------------ jackknife, critical-point step
def _numpy_yt(rec, nl, level, n_ops):
    """y_t from one record with numpy: connected correlators, T = A^-1 B,
    log2 of the largest real eigenvalue (src/mcrg.cpp:110-131)."""
    n = float(rec[orc_off_N(nl)])
    S = lambda l: np.array([rec[l * 3 + a] for a in range(n_ops)], dtype=float) / n
    self_ = np.array([[rec[3 * (nl + 1) + (level + 1) * 9 + b * 3 + a] for b in range(n_ops)]
                      for a in range(n_ops)], dtype=float) / n
    cross = np.array([[rec[12 * (nl + 1) + level * 9 + b * 3 + a] for b in range(n_ops)]
                      for a in range(n_ops)], dtype=float) / n
    A = self_ - np.outer(S(level + 1), S(level + 1))
    B = cross - np.outer(S(level + 1), S(level))
    lam = max(np.linalg.eigvals(np.linalg.solve(A, B)).real)
    return math.log2(lam)


def orc_off_N(nl):
    return 12 * (nl + 1) + 10 * nl + 3


def test_jackknife_against_numpy(orc):
    nl, n_ops = 2, 2
    _, rec = orc.run_chain(8, KC, seed=5, n_therm=200, n_meas=4000, nl=nl, sweeps_per_bin=500)
    assert rec.shape[0] == 8
    tot = rec.sum(axis=0)
    for level in range(nl):
        assert orc.yt(tot, nl, level, n_ops) == pytest.approx(_numpy_yt(tot, nl, level, n_ops), rel=1e-9)
        loo = [_numpy_yt(tot - rec[b], nl, level, n_ops) for b in range(8)]
        want_err = math.sqrt(7.0 / 8.0 * sum((y - np.mean(loo)) ** 2 for y in loo))
        m, e = orc.jackknife_yt(rec, nl, level, n_ops)
        assert m == pytest.approx(_numpy_yt(tot, nl, level, n_ops), rel=1e-9)
        assert e == pytest.approx(want_err, rel=1e-6)
    # identical bins: no spread
    same = np.repeat(rec[:1], 5, axis=0)
    m, e = orc.jackknife_yt(same, nl, 0, n_ops)
    assert e == pytest.approx(0.0, abs=1e-12) and m == pytest.approx(orc.yt(rec[0], nl, 0, n_ops), rel=1e-12)


def kc_known_answer_records(orc):
    """Two samples per lattice, single-count NN values chosen by hand.
    Large (levels 0..2): NN_0 = 10, 14 and NN_1 = 6, 4.  Small (levels 0..1):
    NN_0 = 5, 7.  With the reference's double-counted S (src/mcrg.cpp:240-267):
    <SL> = 24, <SLb_0> = 10, <SLb_0 SL> = (12*20 + 8*28)/2 = 232 -> dSL = -8;
    <SS> = <SSb_0> = 12, <SSb_0 SS> = (100 + 196)/2 = 148 -> dSS = 4;
    dK = (10 - 12) / (-8 - 4) = 1/6 (src/mcrg.cpp:292-295)."""
    rl = np.zeros(orc.record_size(2), dtype=np.int64)
    rl[orc.off_N(2)] = 2
    rl[orc.off_S(2, 0, 0)] = 10 + 14
    rl[orc.off_S(2, 1, 0)] = 6 + 4
    rl[orc.off_nn0(2, 1)] = 6 * 10 + 4 * 14
    rs = np.zeros(orc.record_size(1), dtype=np.int64)
    rs[orc.off_N(1)] = 2
    rs[orc.off_S(1, 0, 0)] = 5 + 7
    rs[orc.off_self(1, 0, 0, 0)] = 25 + 49
    return rl, rs


def test_kc_step_known_answer(orc):
    rl, rs = kc_known_answer_records(orc)
    assert orc.kc_step(rl, 2, rs, 1, 0) == pytest.approx(1.0 / 6.0, rel=1e-12)


# ------------------------------------------------------------------ Swendsen-Wang
def _coin(orc, seed, r, t, tag=5):
    out = orc.philox(seed, [r >> 7, t, 0, tag << 28])
    return (int(out[(r >> 5) & 3]) >> (r & 31)) & 1


def test_sw_threshold_closed_form(orc):
    # exp(-2 K_c) = sqrt(2) - 1, so the bond probability is 2 - sqrt(2)
    assert orc.sw_threshold(KC) == math.floor((2.0 - math.sqrt(2.0)) * 2**32)
    assert orc.sw_threshold(0.0) == 0
    assert orc.sw_threshold(50.0) == 2**32 - 1


def test_sw_singletons_at_zero_coupling(orc):
    """J = 0: no bonds, every site is its own cluster and flips with its own coin."""
    L, seed, t = 8, 77, 5
    s0 = orc.init_spins(L, seed)
    s = s0.copy()
    orc.sw_sweep(L, s, seed, 0.0, t)
    want = np.array([-v if _coin(orc, seed, m, t) else v for m, v in enumerate(s0)], dtype=np.int8)
    assert np.array_equal(s, want)


def test_sw_clusters_and_roots(orc):
    """J large: every aligned pair bonds, so clusters are the aligned domains.
    All-up is one cluster with root 0; vertical stripes of width 2 are L/2
    clusters whose roots are the sites (2k, 0)."""
    L, seed, t = 8, 91, 3
    up = np.ones(L * L, dtype=np.int8)
    s = up.copy()
    orc.sw_sweep(L, s, seed, 50.0, t)
    assert np.array_equal(s, up * (-1 if _coin(orc, seed, 0, t) else 1))
    y, x = np.divmod(np.arange(L * L), L)
    stripes = np.where((x // 2) % 2 == 0, 1, -1).astype(np.int8)
    s = stripes.copy()
    orc.sw_sweep(L, s, seed, 50.0, t)
    root = 2 * (x // 2)
    want = np.array([-v if _coin(orc, seed, int(r), t) else v for v, r in zip(stripes, root)], dtype=np.int8)
    assert np.array_equal(s, want)
