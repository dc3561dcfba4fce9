"""Parity with the numbers the reference publishes for this path, at every
lattice size it publishes them for.

The reference (Wolff sampler, b = 2 majority rule with a random tie coin, NN
and NNN operators) publishes:

* nu = ln 2 / ln lambda per blocking level n = 0 .. floor(log2 N) - 2, at its
  own K_c(N) and at the exact K_c, for N = 16, 32, 64, 128 (slides/MCRG.pdf
  pp.22-23; estimator src/mcrg.cpp:110-137);
* the first iteration of its critical-point locator from K = -0.44, per level,
  for N = 16 .. 128 (slides p.20; src/mcrg.cpp:191-310);
* K_c(N), the average starting K of the last 50 locator iterations, for
  N = 16 .. 128 (src/main.cpp:24-30, averaged by plot/critical_point.py:22-23).

All three are equilibrium expectations of estimators that depend only on the
Boltzmann ensemble and the blocking rule, so any correct sampler reproduces
them within statistical error.  The runs below use Swendsen-Wang (the
all-clusters form of Wolff's update) on the device, n_ops = 2 and the
reference's level count floor(log2 N) - 1.

The reference states no errors.  They are taken as follows (stated, not
measured by the reference):

* nu: the published table holds each N >= 32 twice, at K_c(N) and at the exact
  K_c, which differ by less than 7e-5 in K; the rms of those 15 differences
  over sqrt(2) is the error of one published value, SIGMA_REF_NU (6.0e-4).
* one locator iteration: 1e6 Wolff samples.  A Wolff step at K_c flips about
  half of a lattice of these sizes (0.33-0.55 of the sites), so it is counted
  as half a Swendsen-Wang sweep: the reference's error is our jackknife error
  scaled to 5e5 sweeps.
* K_c(N): an average of 50 iterates of K_{i+1} = K_i + dK_i.  The
  reference's step is half the Newton step (double-counted NN,
  src/mcrg.cpp:295-298), so K_i - K* is an AR(1) series with rho = 1/2 and
  the mean of n iterates has error 2 sigma_step / sqrt(n); sigma_step is that
  of one iteration (as above, 1e6 Wolff samples).

Bounds are 3 combined standard errors (ours and the reference's).
"""
import math
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "mcrg_amd", "mcrg_run")

KC = 0.5 * math.log(1.0 + math.sqrt(2.0))
# |K_c(N)| (src/main.cpp:24-30)
KC_N = {16: 0.440414806102, 32: 0.440619259540, 64: 0.440675235450, 128: 0.440682772730}
# nu per level: (at K_c(N), at exact K_c) (slides pp.22-23)
NU_PUB = {
    16: [(1.0366474702, 1.0371936855), (1.0055084088, 1.0058615220), (0.9850458988, 0.9832937022)],
    32: [(1.0359594271, 1.0364461518), (1.0030411491, 1.0024276306), (1.0042751683, 1.0040962187),
         (0.9833658883, 0.9817702370)],
    64: [(1.0340840722, 1.0356603405), (1.0013550363, 1.0025377191), (1.0021687106, 1.0029573471),
         (1.0029619937, 1.0039006140), (0.9820802731, 0.9819506190)],
    128: [(1.0343762634, 1.0349921920), (1.0018397044, 1.0006737634), (1.0019303651, 1.0022232598),
          (1.0025651221, 1.0023411428), (1.0035789853, 1.0040042374), (0.9814148379, 0.9819137809)],
}
# first locator iteration from K = -0.44, per level n (slides p.20)
KC_FIRST = {
    16: [-0.4420811188, -0.4404791358, -0.4401272033],
    32: [-0.4417429398, -0.4404577112, -0.4402638848, -0.4402975527],
    64: [-0.4415445759, -0.4404935597, -0.4403468642, -0.4403307834, -0.4403116430],
    128: [-0.4413292658, -0.4404266727, -0.4403131062, -0.4403016019, -0.4402985574, -0.4402807798],
}
REF_LOCATOR_SAMPLES = 1.0e6
WOLFF_SWEEPS_PER_STEP = 0.5


def _sigma_ref_nu():
    d = [a - b for N in (32, 64, 128) for a, b in NU_PUB[N]]
    return math.sqrt(sum(x * x for x in d) / len(d)) / math.sqrt(2.0)


SIGMA_REF_NU = _sigma_ref_nu()


def n_t(N):
    """The reference's blockings per sample, floor(log N / log b) - 1 (src/mcrg.cpp:43)."""
    return int(math.floor(math.log(N) / math.log(2))) - 1


@pytest.fixture(scope="module")
def amd():
    import mcrg_amd
    return mcrg_amd


@pytest.fixture(scope="module")
def ctx(amd):
    c = amd.Context(0)
    yield c
    c.close()


def _sw_records(amd, ctx, L, J, nl, n_rep, n_meas, seed0, n_therm=500):
    """One record (one bin) per independent Swendsen-Wang replica."""
    jobs = [dict(L=L, coupling=J, seed=seed0 + i, n_therm=n_therm, n_meas=n_meas, n_levels=nl,
                 sweeps_per_bin=n_meas, sampler=amd.SAMPLER_SWENDSEN_WANG) for i in range(n_rep)]
    with amd.Batch(ctx, jobs) as b:
        b.run_all()
        return np.stack([b.records(j)[0] for j in range(n_rep)])


def _groups(recs, G):
    return recs.reshape(G, recs.shape[0] // G, -1).sum(axis=1)


def test_sigma_ref_nu_from_the_published_pairs():
    assert 4e-4 < SIGMA_REF_NU < 8e-4


@pytest.mark.parametrize("N", [16, 32, 64, 128])
@pytest.mark.parametrize("at", ["Kc_N", "exact"])
def test_nu_per_level_matches_published(amd, ctx, N, at):
    J = KC_N[N] if at == "Kc_N" else KC
    nl = n_t(N)
    n_rep, n_meas = 1024, 20000
    print(f"nu: N={N} J={J:.9f} ({at}), {n_rep} replicas x {n_meas} sweeps, {nl} levels", flush=True)
    grp = _groups(_sw_records(amd, ctx, N, J, nl, n_rep, n_meas, seed0=700000 + 1000 * N + (at == "exact")
                              * 500000), 64)
    fails = []
    for lvl in range(nl):
        r = amd.rg_exponent(grp, nl, lvl, 2)
        pub = NU_PUB[N][lvl][0 if at == "Kc_N" else 1]
        comb = math.hypot(r["nu_err"], SIGMA_REF_NU)
        print(f"  N={N} {at} n={lvl}: nu = {r['nu']:.6f} +- {r['nu_err']:.6f}  published {pub:.6f}  "
              f"diff {(r['nu'] - pub) / comb:+.2f} combined sigma", flush=True)
        assert 0 < r["nu_err"] < 1e-3
        if abs(r["nu"] - pub) > 3 * comb:
            fails.append((lvl, r["nu"], r["nu_err"], pub))
    assert not fails, fails


_FIRST = {}


def _first_iteration(amd, ctx, orc, N):
    """Records of the locator's first iteration at K = -0.44: N lattices with
    n_t(N) blockings and N/2 lattices with n_t(N) - 1, in 32 jackknife groups
    of independent replicas."""
    if N not in _FIRST:
        n_rep, n_meas, G = 512, 8000, 32
        J = 0.44
        print(f"locator: N={N} and N/2 at J={J}, {n_rep} replicas each x {n_meas} sweeps", flush=True)
        rl = _sw_records(amd, ctx, N, J, n_t(N), n_rep, n_meas, seed0=910000 + 4096 * N)
        rs = _sw_records(amd, ctx, N // 2, J, n_t(N) - 1, n_rep, n_meas, seed0=920000 + 4096 * N)
        _FIRST[N] = (_groups(rl, G), _groups(rs, G), n_rep * n_meas)
    return _FIRST[N]


@pytest.mark.parametrize("N", [16, 32, 64])
def test_locator_first_iteration_matches_published(amd, ctx, orc, N):
    gl, gs, n_samples = _first_iteration(amd, ctx, orc, N)
    G, nl = gl.shape[0], n_t(N)
    K0 = -0.44
    scale = math.sqrt(n_samples / (REF_LOCATOR_SAMPLES * WOLFF_SWEEPS_PER_STEP))
    fails = []
    for n in range(nl):
        dk = amd.critical_point_step(gl, nl, gs, nl - 1, n)
        loo = np.array([amd.critical_point_step(np.delete(gl, g, 0), nl, np.delete(gs, g, 0), nl - 1, n)
                        for g in range(G)])
        err = math.sqrt((G - 1) / G * ((loo - loo.mean()) ** 2).sum())
        kc = K0 + dk
        pub = KC_FIRST[N][n]
        comb = math.hypot(err, err * scale)
        print(f"  N={N} n={n}: Kc = {kc:.7f} +- {err:.7f}  published {pub:.7f}  "
              f"diff {(kc - pub) / comb:+.2f} combined sigma", flush=True)
        if abs(kc - pub) > 3 * comb:
            fails.append((n, kc, err, pub))
    assert not fails, fails


@pytest.mark.parametrize("N", [16, 32])
def test_locator_step_on_device_records_equals_oracle(amd, ctx, orc, N):
    """The locator's own device records, summed, through the library's step
    and the oracle's restatement of src/mcrg.cpp:284-299: the same dK."""
    gl, gs, _ = _first_iteration(amd, ctx, orc, N)
    tl, ts = gl.sum(axis=0), gs.sum(axis=0)
    nl = n_t(N)
    for n in range(nl):
        got = amd.critical_point_step(tl, nl, ts, nl - 1, n)
        want = orc.kc_step(tl, nl, ts, nl - 1, n)
        assert got == pytest.approx(want, rel=1e-9, abs=1e-15), (n, got, want)
        # the bins form sums the same groups in 128-bit integers
        assert amd.critical_point_step(gl, nl, gs, nl - 1, n) == pytest.approx(got, rel=1e-12)


def _read_rows(path):
    rows = [l for l in open(path).read().splitlines() if not l.startswith("#")]
    return [[float(x) for x in r.split(",")] for r in rows]


@pytest.mark.parametrize("N", [16, 32, 64])
def test_locator_fixed_point_matches_published(N, tmp_path):
    """locate_critical_point through the reference-named class and the CLI:
    14 iterations from K = -0.44 with 2e6 Swendsen-Wang samples per lattice
    each; the starting K of iterations 5..14 (after the factor-2^-4 approach)
    averaged as plot/critical_point.py averages the reference's, against
    K_c(N) of src/main.cpp:24-30."""
    iters, samples, replicas = 14, 2000000, 256
    args = [RUN, "--locate", str(iters), "--L", str(N), "--K0", "-0.44", "--eq", "500", "--samples",
            str(samples), "--replicas", str(replicas), "--sampler", "sw", "--seed", str(31 + N), "--quiet",
            "--outdir", str(tmp_path)]
    print("run", " ".join(args[1:]), flush=True)
    out = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=240)
    assert out.returncode == 0, out.stderr
    rows = _read_rows(tmp_path / f"critical_point_L_{N}_K_-0.44.txt")
    nl = n_t(N)
    assert len(rows) == iters * nl
    start = np.array([rows[i * nl][2] for i in range(iters)])  # starting K of each iteration
    deep = np.array([rows[i * nl + nl - 1][3] for i in range(iters)])  # its deepest-level result
    assert np.allclose(start[1:], deep[:-1], rtol=0, atol=1e-9)  # each iteration starts where the last ended
    tail = np.concatenate([start[4:], deep[-1:]])  # K_5 .. K_15
    m = tail.mean()
    # AR(1) innovations of K_{i+1} - m = (K_i - m) / 2 + e_i
    e = (tail[1:] - m) - 0.5 * (tail[:-1] - m)
    sigma_step = e.std(ddof=1)
    err = 2.0 * sigma_step / math.sqrt(len(tail))
    sigma_step_ref = sigma_step * math.sqrt(samples / (REF_LOCATOR_SAMPLES * WOLFF_SWEEPS_PER_STEP))
    err_ref = 2.0 * sigma_step_ref / math.sqrt(50)
    pub = -KC_N[N]
    comb = math.hypot(err, err_ref)
    print(f"  N={N}: K_c(N) = {m:.7f} +- {err:.7f} (step sigma {sigma_step:.2e})  published {pub:.7f}  "
          f"diff {(m - pub) / comb:+.2f} combined sigma", flush=True)
    assert abs(m - pub) <= 3 * comb, (m, err, pub, err_ref)
