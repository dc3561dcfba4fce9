"""The K-bisection's crossing fit (mcrg_cpp::zero_crossing, the host library's
C++ code through a small probe program) on curves whose zero is known.  No GPU.

D(J), the difference of the blocked nearest-neighbour correlations of the two
lattice sizes, saturates on both sides of the critical coupling, so a straight
line through a wide bracket is biased; each curve here is built so that the
straight-line answer misses by far more than the tolerance the fit is held to.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROBE = os.path.join(ROOT, "mcrg_amd", "_build", "crossing_probe")
JC = 0.4406868


def crossing(J, D, w):
    if not os.path.exists(PROBE):
        pytest.skip("mcrg_amd/_build/crossing_probe not built (make -C mcrg_amd)")
    text = "".join(f"{float(a)!r} {float(b)!r} {float(c)!r}\n" for a, b, c in zip(J, D, w))
    out = subprocess.run([PROBE], input=text, capture_output=True, text=True, check=True, timeout=30)
    return float(out.stdout)


def line_crossing(J, D, w):
    J, D, w = map(np.asarray, (J, D, w))
    b, a = np.polyfit(J, D, 1, w=np.sqrt(w))
    return -a / b


def grid(lo=0.43, hi=0.45, n=16):
    return lo + (hi - lo) * (np.arange(n) + 0.5) / n


def test_straight_line_is_exact():
    J = grid()
    D = 3.0 * (J - JC)
    assert crossing(J, D, np.ones_like(J)) == pytest.approx(JC, abs=1e-12)


def test_straight_line_without_sign_change_extrapolates():
    J = grid(0.445, 0.45, 8)
    D = 2.0 * (J - JC)
    assert crossing(J, D, np.ones_like(J)) == pytest.approx(JC, abs=1e-12)


def test_asymmetric_saturation():
    # saturates at -1 below and +2 above, variance largest at the crossing
    J = grid()
    t = np.tanh((J - JC) / 0.002)
    D = t * (1.5 + 0.5 * t)
    w = 1e6 / (1.0 - t * t + 0.01)
    assert abs(line_crossing(J, D, w) - JC) > 1e-3
    assert crossing(J, D, w) == pytest.approx(JC, abs=2e-5)


def test_bump_with_sign_noise_far_away():
    # D returns to zero far from the crossing, where tiny values of either
    # sign with tiny errors must not be taken for the crossing
    J = grid(0.40, 0.48, 64)
    x = (J - JC) / 0.004
    D = x * np.exp(-x * x) + 1e-9 * np.where(np.arange(J.size) % 2, 1.0, -1.0)
    w = 1.0 / (1e-10 + 1e-4 * np.exp(-x * x))
    got = crossing(J, D, w)
    assert got == pytest.approx(JC, abs=2e-5)


@pytest.mark.parametrize("seed", range(5))
def test_noisy_sigmoid(seed):
    rng = np.random.default_rng(seed)
    J = grid()
    t = np.tanh((J - JC) / 0.003)
    sd = 1e-3 * (1.0 + (1.0 - t * t))
    D = t * (1.3 + 0.3 * t) + rng.normal(0.0, sd)
    got = crossing(J, D, 1.0 / sd ** 2)
    # the slope at the crossing is 1.3 / 0.003: a 1e-3 error in D moves the
    # zero by about 2.3e-6
    assert got == pytest.approx(JC, abs=2e-5)


def test_narrow_bracket_takes_the_line():
    """A bracket in which D(J) is straight within its errors gets the straight
    line's zero, even outside the bracket (a fit confined to the bracket would
    pin it there).  sd = 1e-3 and slope 1.3 / 0.003 make the zero's
    statistical error about 2.3e-6; the line's curvature bias here is far
    below it."""
    J = grid(JC - 1e-4, JC + 3e-4)
    t = np.tanh((J - JC) / 0.003)
    D = t * (1.3 + 0.3 * t)
    w = np.full_like(J, 1e6)
    assert crossing(J, D, w) == pytest.approx(line_crossing(J, D, w), abs=1e-12)
    assert crossing(J, D, w) == pytest.approx(JC, abs=5e-7)
    # a bracket entirely above the zero: the line extrapolates below it
    J2 = grid(JC + 1e-4, JC + 5e-4)
    t2 = np.tanh((J2 - JC) / 0.003)
    got = crossing(J2, t2 * (1.3 + 0.3 * t2), w)
    assert got < J2.min()
    assert got == pytest.approx(JC, abs=5e-6)


@pytest.mark.parametrize("bracket", [(0.43, 0.45), (0.438, 0.443), (0.4402, 0.4412)])
@pytest.mark.parametrize("noise", [1e-4, 1e-3, 1e-2])
def test_unbiased_over_noisy_replicas(bracket, noise):
    """Over 100 noisy copies of an asymmetric sigmoid the crossing's bias stays
    below its spread (one standard deviation), for wide and narrow brackets (a
    straight line through every point is off by 1.7e-3 on the widest, 30 to
    4000 times its spread)."""
    rng = np.random.default_rng(7)
    J = grid(*bracket)
    t = np.tanh((J - JC) / 0.003)
    sd = noise * (2.0 - t * t)
    D0 = t * (1.3 + 0.3 * t)
    got = np.array([crossing(J, D0 + rng.normal(0.0, sd), 1.0 / sd ** 2) for _ in range(100)])
    assert abs(got.mean() - JC) < got.std() + 1e-7


def test_nan_without_slope():
    J = grid()
    assert np.isnan(crossing(J, np.ones_like(J), np.ones_like(J)))
