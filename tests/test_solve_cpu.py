"""The product's FP64 solve (host code in libmcrg_amd.so) against the oracle's
independent implementation, numpy, and closed forms.  No device work."""
import math

import numpy as np
import pytest

KC = 0.5 * math.log(1.0 + math.sqrt(2.0))


@pytest.fixture(scope="module")
def amd():
    import mcrg_amd
    return mcrg_amd


@pytest.fixture(scope="module")
def chains(orc):
    """Oracle records for a few (L, J, levels)."""
    out = {}
    for L, J, nl in ((8, KC, 2), (16, 0.42, 3), (16, KC, 3), (8, 0.3, 1)):
        _, rec = orc.run_chain(L, J, seed=L * 1000 + nl, n_therm=200, n_meas=4000, nl=nl, sweeps_per_bin=500)
        out[(L, J, nl)] = rec
    return out


def test_rg_matrix_matches_oracle(amd, orc, chains):
    for (L, J, nl), rec in chains.items():
        tot = rec.sum(axis=0)
        for level in range(nl):
            for n_ops in (1, 2, 3):
                a = amd.rg_matrix(tot, nl, level, n_ops)
                b = orc.rg_matrix(tot, nl, level, n_ops)
                assert np.allclose(a, b, rtol=1e-9, atol=1e-12), (L, J, nl, level, n_ops)


def test_leading_eigenvalue_against_numpy(amd):
    rng = np.random.default_rng(3)
    for n in (1, 2, 3):
        for _ in range(500):
            T = rng.normal(size=(n, n)) * rng.choice([0.01, 1.0, 30.0])
            want = max(np.linalg.eigvals(T).real)
            assert abs(amd.leading_eigenvalue(T) - want) < 1e-9 * max(1.0, abs(want))


def test_leading_eigenvalue_closed_forms(amd):
    assert amd.leading_eigenvalue(np.diag([2.0, 3.0, -7.0])) == pytest.approx(3.0)
    rot = np.array([[1.0, -2.0, 0.0], [2.0, 1.0, 0.0], [0.0, 0.0, 0.5]])  # 1 +- 2i, 0.5
    assert amd.leading_eigenvalue(rot) == pytest.approx(1.0)
    rot2 = np.array([[0.2, -2.0, 0.0], [2.0, 0.2, 0.0], [0.0, 0.0, 0.9]])  # 0.2 +- 2i, 0.9
    assert amd.leading_eigenvalue(rot2) == pytest.approx(0.9)
    jordan = np.array([[2.0, 1.0, 0.0], [0.0, 2.0, 1.0], [0.0, 0.0, 2.0]])
    assert amd.leading_eigenvalue(jordan) == pytest.approx(2.0, rel=1e-5)


def test_exponent_and_jackknife_match_oracle(amd, orc, chains):
    for (L, J, nl), rec in chains.items():
        for level in range(nl):
            for n_ops in (2, 3):
                got = amd.rg_exponent(rec, nl, level, n_ops)
                m, e = orc.jackknife_yt(rec, nl, level, n_ops)
                assert got["yt"] == pytest.approx(m, rel=1e-9)
                assert got["yt_err"] == pytest.approx(e, rel=1e-6, abs=1e-12)
                assert got["nu"] == pytest.approx(1.0 / m, rel=1e-9)
                assert got["lambda"] == pytest.approx(2.0 ** m, rel=1e-9)
                assert got["n_jackknife"] == rec.shape[0]


def test_critical_point_step_matches_oracle(amd, orc):
    L, J = 16, 0.44
    _, rl = orc.run_chain(L, J, seed=1, n_therm=300, n_meas=3000, nl=3, sweeps_per_bin=3000)
    _, rs = orc.run_chain(L // 2, J, seed=2, n_therm=300, n_meas=3000, nl=2, sweeps_per_bin=3000)
    for level in range(3):
        got = amd.critical_point_step(rl[0], 3, rs[0], 2, level)
        want = orc.kc_step(rl[0], 3, rs[0], 2, level)
        assert got == pytest.approx(want, rel=1e-9)


def test_critical_point_step_is_zero_for_identical_statistics(amd, orc):
    # Matching a lattice against itself: blocked L statistics equal the
    # unblocked "small" ones exactly when both records are built to agree.
    nl = 2
    rec = np.zeros(orc.record_size(nl), dtype=np.int64)
    n = 100
    rec[orc.off_N(nl)] = n
    rec[orc.off_S(nl, 0, 0)] = 50 * n
    rec[orc.off_S(nl, 1, 0)] = 30 * n
    rec[orc.off_nn0(nl, 1)] = 1600 * n
    rec[orc.off_self(nl, 0, 0, 0)] = 2600 * n
    small = np.zeros(orc.record_size(nl - 1), dtype=np.int64)
    small[orc.off_N(nl - 1)] = n
    small[orc.off_S(nl - 1, 0, 0)] = 30 * n
    small[orc.off_self(nl - 1, 0, 0, 0)] = 1100 * n
    # <S_L^(1)> = <S_S^(0)>: the numerator vanishes
    assert amd.critical_point_step(rec, nl, small, nl - 1, 0) == 0.0


def test_solve_rejects_bad_arguments(amd):
    rec = np.zeros(amd.record_size(2), dtype=np.int64)
    with pytest.raises(amd.MCRGError):
        amd.rg_matrix(rec, 2, 0, 2)  # no samples
    with pytest.raises(amd.MCRGError):
        amd.rg_matrix(rec, 2, 2, 2)  # level out of range


def test_critical_point_step_known_answer(amd, orc):
    from test_oracle import kc_known_answer_records
    rl, rs = kc_known_answer_records(orc)
    assert amd.critical_point_step(rl, 2, rs, 1, 0) == pytest.approx(1.0 / 6.0, rel=1e-12)


def _ar1_records(rho, n, bin_len, seed, shift=0.0):
    """Per-bin records (no blocking level) of an AR(1) series x_t = rho x_{t-1}
    + e_t, scaled by 1000 and rounded, in the NN slot and in the |M| slots."""
    from scipy.signal import lfilter
    rng = np.random.default_rng(seed)
    e = rng.standard_normal(n + 2000)
    x = lfilter([1.0], [1.0, -rho], e)[2000:]
    xi = np.rint(1000.0 * (x + shift)).astype(np.int64).reshape(-1, bin_len)
    rs = 16  # record_size(0)
    rec = np.zeros((xi.shape[0], rs), dtype=np.int64)
    rec[:, 0] = xi.sum(axis=1)          # S[0][NN]
    rec[:, 3] = (xi * xi).sum(axis=1)   # SELF[0][NN][NN]
    rec[:, 13] = np.abs(xi).sum(axis=1)  # |M|
    rec[:, 14] = (xi * xi).sum(axis=1)   # M^2
    rec[:, 15] = bin_len
    return rec


@pytest.mark.parametrize("rho", [0.0, 0.5, 0.9])
def test_tau_int_ar1_closed_form(amd, rho):
    """tau_int = (1 + rho) / (2 (1 - rho)) for an AR(1) series (1/2 for white
    noise); the estimate from given bins of 1024 samples is within 5 %, the
    coarse one within 4 of its own errors."""
    exact = (1.0 + rho) / (2.0 * (1.0 - rho))
    rec = _ar1_records(rho, 1 << 22, 1024, seed=7)
    for op, shift in ((amd.OP_NN, 0.0), (amd.OP_ABS_M, 50.0)):
        r = amd.tau_int(_ar1_records(rho, 1 << 22, 1024, seed=7, shift=shift) if shift else rec, 0, 0, op)
        assert r["tau_fine"] == pytest.approx(exact, rel=0.05), (op, r)
        assert abs(r["tau_int"] - exact) < 4 * r["tau_err"], (op, r)
        assert r["n_bins"] == 32 and r["bin_samples"] == (1 << 22) // 32
    # a planted factor-of-two mistake in the series' correlation would be seen
    assert abs(amd.tau_int(rec, 0, 0, amd.OP_NN)["tau_fine"] - 2 * exact) > 0.2 * exact


def test_tau_int_partial_last_bin_and_errors(amd):
    rec = _ar1_records(0.5, 1 << 16, 512, seed=3)
    full = amd.tau_int(rec, 0, 0, amd.OP_NN)
    short = rec.copy()
    short[-1, 15] = 100  # a shorter last bin is left out
    r = amd.tau_int(short, 0, 0, amd.OP_NN)
    full_minus = amd.tau_int(rec[:-1], 0, 0, amd.OP_NN)
    assert r == full_minus and r != full
    with pytest.raises(amd.MCRGError):
        amd.tau_int(rec[:1], 0, 0, amd.OP_NN)
    with pytest.raises(amd.MCRGError):
        amd.tau_int(rec, 0, 0, 4)


def _scaled_bins(orc, rec, nl, c, n_bins):
    """The total of `rec` with every operator multiplied by the integer c (sums
    by c, products by c^2), spread over n_bins int64 bins whose exact sum it
    is.  Scaling the operators is a similarity transform of T (same
    eigenvalues) and divides the locator step by c."""
    tot = [int(v) for v in rec.sum(axis=0)]
    n_s = 3 * (nl + 1)
    i_m = orc.off_N(nl) - 3  # M, |M|, M^2, count
    scaled = []
    for k, v in enumerate(tot):
        scaled.append(v * c if k < n_s else (v if k >= i_m else v * c * c))
    out = np.zeros((n_bins, len(tot)), dtype=np.int64)
    for k, v in enumerate(scaled):
        q, r = divmod(v, n_bins)
        assert abs(q) < 2**62
        out[:, k] = q
        out[0, k] += r
    return scaled, out


def test_solve_is_exact_beyond_int64_totals(amd, orc, chains):
    """Bins whose totals pass 2^63 (as the NN self sums of L = 2048 do over 1e6
    samples) give the same lambda as the unscaled records: the sums over bins
    are formed in 128-bit integers.  Summed in int64 they would wrap."""
    L, J, nl = 16, KC, 3
    rec = chains[(L, J, nl)]
    c = 1 << 20
    scaled, bins = _scaled_bins(orc, rec, nl, c, 4096)
    assert max(abs(v) for v in scaled) > 2**63  # the totals do not fit int64
    for level in range(nl):
        for n_ops in (2, 3):
            want = amd.rg_exponent(rec, nl, level, n_ops)["lambda"]
            got = amd.rg_exponent(bins, nl, level, n_ops)["lambda"]
            assert got == pytest.approx(want, rel=1e-12), (level, n_ops)
    # the locator step: one record of each size, scaled alike
    _, rl = orc.run_chain(16, 0.44, seed=1, n_therm=300, n_meas=3000, nl=3, sweeps_per_bin=3000)
    _, rs = orc.run_chain(8, 0.44, seed=2, n_therm=300, n_meas=3000, nl=2, sweeps_per_bin=3000)
    _, bl = _scaled_bins(orc, rl, 3, c, 512)
    _, bs = _scaled_bins(orc, rs, 2, c, 256)
    for level in range(3):
        want = amd.critical_point_step(rl[0], 3, rs[0], 2, level)
        assert amd.critical_point_step(bl, 3, bs, 2, level) == pytest.approx(want / c, rel=1e-10)


def test_max_sweeps_per_bin_closed_form(amd):
    """floor((2^63 - 1) / (2 L^2)^2): 2^17 - 1 at L = 2048, 31 at L = 16384."""
    assert amd.max_sweeps_per_bin(2048) == 2**17 - 1
    assert amd.max_sweeps_per_bin(4096) == 2**13 - 1
    assert amd.max_sweeps_per_bin(16384) == 31
    assert amd.max_sweeps_per_bin(64) == 2**31 - 1  # capped at int32
    assert amd.max_sweeps_per_bin(1) == 0 and amd.max_sweeps_per_bin(32768) == 0
    for L in (256, 1024, 8192):
        m = amd.max_sweeps_per_bin(L)
        assert m * (2 * L * L) ** 2 < 2**63 <= (m + 1) * (2 * L * L) ** 2
