"""Exact results for the 2D Ising model used as physics pins by the tests.

* kaufman_log_z: the partition function of an m x n periodic lattice in
  closed form (B. Kaufman, Phys. Rev. 76, 1232 (1949); written out as in
  P. D. Beale, Phys. Rev. Lett. 76, 78 (1996)):
      Z = 1/2 (2 sinh 2K)^(mn/2) (Z1 + Z2 + Z3 + Z4),
      Z1 = prod_{r<n} 2 cosh(m g_{2r+1} / 2),  Z2 = prod_{r<n} 2 sinh(m g_{2r+1} / 2),
      Z3 = prod_{r<n} 2 cosh(m g_{2r} / 2),    Z4 = prod_{r<n} 2 sinh(m g_{2r} / 2),
      cosh g_l = cosh 2K coth 2K - cos(pi l / n)  (l >= 1, g_l > 0),
      g_0 = 2K + ln tanh K  (negative above the critical temperature).
  Z = sum over states of exp(K sum_<ij> s_i s_j), so <NN> = d ln Z / dK with
  NN the single-count nearest-neighbour sum (2mn bonds).
* onsager_nn_per_site: the infinite lattice's <NN>/N (L. Onsager, Phys. Rev.
  65, 117 (1944)): coth 2K [1 + (2/pi)(2 tanh^2 2K - 1) K(k)], k = 2 sinh 2K /
  cosh^2 2K, K(k) the complete elliptic integral of the first kind.  At the
  critical coupling it is sqrt(2).

The tests check kaufman_log_z against brute-force enumeration at L = 4 and
against Onsager's limit at large L before using it.
"""
import math

import numpy as np


def _log_2cosh(x):
    a = np.abs(x)
    return a + np.log1p(np.exp(-2.0 * a))


def _gammas(n, K):
    l = np.arange(2 * n, dtype=np.float64)
    c = math.cosh(2 * K) / math.tanh(2 * K) - np.cos(math.pi * l / n)
    g = np.arccosh(c)
    g[0] = 2 * K + math.log(math.tanh(K))
    return g


def kaufman_log_z(m, n, K):
    """ln Z of an m x n periodic lattice (m rows, n columns) at coupling K > 0."""
    g = _gammas(n, K)
    odd, even = 0.5 * m * g[1::2], 0.5 * m * g[0::2]
    terms = []  # (sign, log|Z_i|)
    terms.append((1.0, float(np.sum(np.log(2.0) + _log_2cosh(odd) - np.log(2.0)))))
    terms.append((1.0, float(np.sum(odd + np.log(-np.expm1(-2.0 * odd))))))
    terms.append((1.0, float(np.sum(_log_2cosh(even)))))
    if even[0] != 0.0:
        sgn = math.copysign(1.0, even[0])
        a = np.abs(even)
        terms.append((sgn, float(np.sum(a + np.log(-np.expm1(-2.0 * a))))))
    top = max(t[1] for t in terms)
    s = sum(sg * math.exp(lg - top) for sg, lg in terms)
    return -math.log(2.0) + 0.5 * m * n * math.log(2.0 * math.sinh(2 * K)) + top + math.log(s)


def nn_mean(L, K, h=2e-3):
    """<NN> (single count) on the L x L torus: d ln Z / dK by Richardson-
    extrapolated central differences (error far below 1e-9 per site)."""
    def d(hh):
        return (kaufman_log_z(L, L, K + hh) - kaufman_log_z(L, L, K - hh)) / (2 * hh)
    return (4.0 * d(h / 2) - d(h)) / 3.0


def onsager_nn_per_site(K):
    from scipy.special import ellipk
    t2 = math.tanh(2 * K) ** 2
    k = 2 * math.sinh(2 * K) / math.cosh(2 * K) ** 2
    if k * k >= 1.0 - 1e-15:  # K = K_c: (2 tanh^2 2K - 1) K(k) -> 0 (a logarithm times a zero)
        return 1.0 / math.tanh(2 * K)
    return (1.0 / math.tanh(2 * K)) * (1.0 + (2.0 / math.pi) * (2 * t2 - 1.0) * float(ellipk(k * k)))
