"""Pins of the exact finite-lattice Ising results (Kaufman) and their use on
the oracle's samplers at L = 8 and 16, beyond the 4 x 4 enumeration."""
import math

import numpy as np
import pytest

from exact_ising import kaufman_log_z, nn_mean, onsager_nn_per_site

KC = 0.5 * math.log(1.0 + math.sqrt(2.0))


def _brute_L4(K):
    n = 16
    idx = np.arange(2**n, dtype=np.int64)
    s = (1 - 2 * ((idx[:, None] >> np.arange(n)[None, :]) & 1)).reshape(-1, 4, 4)
    nn = (s * np.roll(s, -1, axis=2) + s * np.roll(s, -1, axis=1)).sum(axis=(1, 2)).astype(np.float64)
    top = K * nn.max()
    w = np.exp(K * nn - top)
    return top + math.log(w.sum()), float((w * nn).sum() / w.sum())


@pytest.mark.parametrize("K", [0.1, 0.3, KC, 0.5, 0.9])
def test_kaufman_matches_enumeration_L4(K):
    lz, nn = _brute_L4(K)
    assert kaufman_log_z(4, 4, K) == pytest.approx(lz, rel=1e-12, abs=1e-12)
    assert nn_mean(4, K) == pytest.approx(nn, rel=1e-8)


@pytest.mark.parametrize("K", [0.3, 0.6])
def test_kaufman_large_L_approaches_onsager(K):
    """Away from K_c the finite-size correction is exponentially small."""
    assert nn_mean(512, K) / 512**2 == pytest.approx(onsager_nn_per_site(K), abs=1e-9)


def test_onsager_at_kc_is_sqrt2_and_finite_size_approach():
    assert onsager_nn_per_site(KC) == pytest.approx(math.sqrt(2.0), abs=1e-12)
    e = [nn_mean(L, KC) / L**2 for L in (16, 64, 256)]
    d = [abs(x - math.sqrt(2.0)) for x in e]
    assert d[0] > d[1] > d[2] and d[2] < 0.01


@pytest.mark.parametrize("L,J,sampler", [(8, KC, 0), (8, 0.3, 0), (8, KC, 1), (16, KC, 1), (16, 0.5, 1)])
def test_oracle_samplers_match_kaufman(orc, L, J, sampler):
    """Mean NN per site of the oracle's Metropolis (0) and Swendsen-Wang (1)
    chains against the exact finite-lattice value, bins as error estimate."""
    nl, spb = 1, 1000
    n_meas = 100000 if L == 8 else 40000
    _, rec = orc.run_chain(L, J, seed=31 + L, n_therm=1000, n_meas=n_meas, nl=nl, sweeps_per_bin=spb,
                           sampler=sampler)
    n = rec[:, orc.off_N(nl)].astype(float)
    means = rec[:, orc.off_S(nl, 0, 0)] / n / (L * L)
    err = means.std(ddof=1) / math.sqrt(len(means))
    exact = nn_mean(L, J) / (L * L)
    assert abs(means.mean() - exact) < 5 * err, (means.mean(), exact, err)
    # the check can see a 2 % error in the coupling
    assert abs(means.mean() - nn_mean(L, 1.02 * J) / (L * L)) > 5 * err
