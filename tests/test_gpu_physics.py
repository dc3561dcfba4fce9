"""Physics parity with the reference's published MCRG results, and with the
exact finite-lattice energy (Kaufman's partition function, tests/exact_ising.py).

The reference publishes nu = ln 2 / ln lambda per blocking level for N = 16 at
the exact critical coupling, NN + NNN operators, b = 2 majority blocking
(slides/MCRG.pdf p.22, from 1e7 Wolff samples):
    n = 0: 1.0371936855,  n = 1: 1.0058615220,  n = 2: 0.9832937022.
These are equilibrium expectations of the estimator, so any correct sampler
reproduces them within statistical error.  The reference states no error for
them; its 1e7 samples put them near 3e-4 (our own 2e7-sample runs at L = 16
have jackknife errors of 3e-4), so the bounds below are about 6 combined
errors for Swendsen-Wang and 5 of our jackknife errors for Metropolis, whose
samples are correlated over ~L^2.2 sweeps.
"""
import math

import numpy as np
import pytest

from exact_ising import nn_mean

pytestmark = pytest.mark.gpu

KC = 0.5 * math.log(1.0 + math.sqrt(2.0))
NU_PUBLISHED_N16 = (1.0371936855, 1.0058615220, 0.9832937022)


@pytest.fixture(scope="module")
def amd():
    import mcrg_amd
    return mcrg_amd


@pytest.fixture(scope="module")
def ctx(amd):
    c = amd.Context(0)
    yield c
    c.close()


def _replicas(amd, ctx, sampler, n_rep, n_therm, n_meas, L=16, J=KC, nl=3):
    """One record row (one bin) per independent replica."""
    jobs = [dict(L=L, coupling=J, seed=90000 + i, n_therm=n_therm, n_meas=n_meas, n_levels=nl,
                 sweeps_per_bin=n_meas, sampler=sampler) for i in range(n_rep)]
    with amd.Batch(ctx, jobs) as b:
        b.run_all()
        return np.stack([b.records(j)[0] for j in range(n_rep)])


_RUNS = {}


def _sw16(amd, ctx):
    if "sw16" not in _RUNS:
        print("physics: swendsen-wang L=16, 2048 replicas", flush=True)
        _RUNS["sw16"] = _replicas(amd, ctx, amd.SAMPLER_SWENDSEN_WANG, 2048, 500, 10000)
    return _RUNS["sw16"]


def _metro16(amd, ctx):
    if "metro16" not in _RUNS:
        print("physics: metropolis L=16 (LDS-resident), 1024 replicas", flush=True)
        _RUNS["metro16"] = _replicas(amd, ctx, amd.SAMPLER_METROPOLIS, 1024, 10000, 50000)
    return _RUNS["metro16"]


def _print_tau(amd, recs, nl, what):
    """Binning tau_int of NN per level over the replicas' bins (one bin each)."""
    for lvl in range(nl + 1):
        t = amd.tau_int(recs, nl, lvl, amd.OP_NN)
        print(f"{what}: tau_int(NN, level {lvl}) = {t['tau_int']:.2f} +- {t['tau_err']:.2f} sweeps "
              f"(bins of {t['bin_samples']} samples)", flush=True)


def _nu(amd, recs, groups):
    # independent replicas summed into jackknife groups
    grp = recs.reshape(groups, recs.shape[0] // groups, -1).sum(axis=1)
    return [amd.rg_exponent(grp, 3, lvl, 2) for lvl in range(3)]


def _check_nn(amd, recs, L, J, nl, n_sigma=5.0):
    """Replica means of NN per site against Kaufman's exact finite-lattice value."""
    off_s, off_n = 0, 12 * (nl + 1) + 10 * nl + 3  # S[0][NN] and the sample count
    assert amd.record_size(nl) == off_n + 1
    per = recs[:, off_s] / recs[:, off_n] / (L * L)
    mean, err = per.mean(), per.std(ddof=1) / math.sqrt(len(per))
    exact = nn_mean(L, J) / (L * L)
    print(f"NN/site L={L} J={J:.5f}: {mean:.6f} +- {err:.6f} (exact {exact:.6f})", flush=True)
    assert abs(mean - exact) < n_sigma * err, (mean, err, exact)


def test_nn_swendsen_wang_matches_kaufman(amd, ctx):
    _check_nn(amd, _sw16(amd, ctx), 16, KC, 3)
    _print_tau(amd, _sw16(amd, ctx), 3, "swendsen-wang L=16")
    print("physics: swendsen-wang L=32 J=0.42", flush=True)
    _check_nn(amd, _replicas(amd, ctx, amd.SAMPLER_SWENDSEN_WANG, 512, 500, 5000, L=32, J=0.42, nl=1), 32, 0.42, 1)


def test_nn_metropolis_batched_matches_kaufman(amd, ctx):
    _check_nn(amd, _metro16(amd, ctx), 16, KC, 3)
    _print_tau(amd, _metro16(amd, ctx), 3, "metropolis L=16")
    print("physics: metropolis L=32 J=0.46", flush=True)
    _check_nn(amd, _replicas(amd, ctx, amd.SAMPLER_METROPOLIS, 512, 5000, 20000, L=32, J=0.46, nl=1), 32, 0.46, 1)


def test_nu_swendsen_wang_matches_published_n16(amd, ctx):
    res = _nu(amd, _sw16(amd, ctx), 64)
    for lvl, (r, pub) in enumerate(zip(res, NU_PUBLISHED_N16)):
        print(f"level {lvl}: nu = {r['nu']:.5f} +- {r['nu_err']:.5f} (published {pub:.5f})", flush=True)
        assert 0 < r["nu_err"] < 0.001
        assert abs(r["nu"] - pub) < 0.0025, (lvl, r["nu"], pub)


def test_nu_metropolis_batched_matches_published_n16(amd, ctx):
    res = _nu(amd, _metro16(amd, ctx), 64)
    for lvl, (r, pub) in enumerate(zip(res, NU_PUBLISHED_N16)):
        print(f"level {lvl}: nu = {r['nu']:.5f} +- {r['nu_err']:.5f} (published {pub:.5f})", flush=True)
        assert 0 < r["nu_err"] < 0.004
        assert abs(r["nu"] - pub) < 5 * r["nu_err"] + 0.0005, (lvl, r["nu"], pub, r["nu_err"])


def test_config_l8_plain_yt_matches_oracle(amd, ctx, orc):
    """The plain single-lattice case: L = 8, |K| = 0.4406868, seed 1, 1e4
    thermalisation and 1e5 measured sweeps, one 8 -> 4 blocking, on the
    device's HBM schedule.  Records equal the oracle's bit for bit, so y_t and
    its jackknife error match exactly; y_t is also within its error of the
    value of an independent seed."""
    L, J, nl, spb = 8, 0.4406868, 1, 2000
    print("config L=8 plain", flush=True)
    with amd.Chain(ctx, L, J, 1, n_levels=nl, sweeps_per_bin=spb, max_bins=50) as ch:
        ch.sweep(10000)
        ch.sweep(100000, measure=True)
        rec = ch.records()
    _, rec_ref = orc.run_chain(L, J, 1, 10000, 100000, nl, spb)
    assert np.array_equal(rec, rec_ref)
    for n_ops in (2, 3):
        r = amd.rg_exponent(rec, nl, 0, n_ops)
        m, e = orc.jackknife_yt(rec_ref, nl, 0, n_ops)
        assert r["yt"] == pytest.approx(m, rel=1e-12) and r["yt_err"] == pytest.approx(e, rel=1e-9)
        print(f"n_ops={n_ops}: y_t = {r['yt']:.5f} +- {r['yt_err']:.5f}", flush=True)
    _, rec2 = orc.run_chain(L, J, 2, 10000, 100000, nl, spb)
    r2 = amd.rg_exponent(rec2, nl, 0, 2)
    r1 = amd.rg_exponent(rec, nl, 0, 2)
    assert abs(r1["yt"] - r2["yt"]) < 4 * math.hypot(r1["yt_err"], r2["yt_err"])


def test_magnetisation_and_yt_agree_between_samplers(amd, ctx):
    """Swendsen-Wang (a cluster algorithm, like the reference's Wolff sampler)
    and checkerboard Metropolis sample the same ensemble: at L = 16, K_c the
    mean |M| and M^2 per site and y_t at every level agree within 5 combined
    standard errors (replica spread for the moments, jackknife for y_t)."""
    nl, L = 3, 16
    off_m = 12 * (nl + 1) + 10 * nl  # M, |M|, M^2, count
    out = {}
    for name, recs in (("sw", _sw16(amd, ctx)), ("metropolis", _metro16(amd, ctx))):
        n = recs[:, off_m + 3].astype(float)
        absm = recs[:, off_m + 1] / n / (L * L)
        m2 = recs[:, off_m + 2] / n / float(L * L) ** 2
        yts = _nu(amd, recs, 64)
        out[name] = (absm.mean(), absm.std(ddof=1) / math.sqrt(len(absm)),
                     m2.mean(), m2.std(ddof=1) / math.sqrt(len(m2)), yts)
        print(f"{name}: <|M|>/N = {absm.mean():.5f} +- {out[name][1]:.5f}, <M^2>/N^2 = {m2.mean():.5f} "
              f"+- {out[name][3]:.5f}, y_t = " + ", ".join(f"{r['yt']:.4f}({r['yt_err']:.4f})" for r in yts),
              flush=True)
    a, b = out["sw"], out["metropolis"]
    assert abs(a[0] - b[0]) < 5 * math.hypot(a[1], b[1])
    assert abs(a[2] - b[2]) < 5 * math.hypot(a[3], b[3])
    for ra, rb in zip(a[4], b[4]):
        assert abs(ra["yt"] - rb["yt"]) < 5 * math.hypot(ra["yt_err"], rb["yt_err"])
