"""The reference-named C++ classes (mcrg_amd/host) and the `mcrg_run` driver on
the GPU: bit-exact against the oracle through the classes, output files in the
reference's formats (src/mcrg.cpp:31-39,135 and :157-166,303), and the same
exponents as the batch entry points with the same seeds."""
import json
import math
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "mcrg_amd", "mcrg_run")
PROBE = os.path.join(ROOT, "mcrg_amd", "_build", "class_probe")
KC_REF = -math.log(1.0 + math.sqrt(2.0)) / 2.0


def _run(args, cwd):
    print("run", args, flush=True)
    out = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=240)
    assert out.returncode == 0, out.stderr
    return out.stdout


def test_classes_bit_exact_vs_oracle(orc, tmp_path):
    from mcrg_amd.workloads import splitmix64
    N, K = 16, -0.4406868
    got = json.loads(_run([PROBE, str(N), repr(K)], tmp_path))
    seed = splitmix64(12345)
    assert got["seed"] == seed
    s = orc.init_spins(N, seed)
    for t in range(37 + 5):
        orc.sweep(N, s, seed, -K, t)
    assert got["spins"] == [int(v) for v in s]
    ops = orc.ops(N, s)
    assert got["NN_ref"] == 2 * ops[0] and got["NNN_ref"] == 2 * ops[1] and got["NN_only"] == 2 * ops[0]
    # calc_interactions is the reference's double-counted sum, src/lattice.cpp:102-120
    ref = orc.ops_refcount(N, s)
    assert got["NN_ref"] == ref[0] and got["NNN_ref"] == ref[1]
    assert got["E"] == pytest.approx(K * 2 * ops[0] / N**2, rel=1e-12)
    assert got["M"] == pytest.approx(s.astype(float).sum() / N**2, rel=1e-12)
    assert got["nn"] == [[1, N - 1], [N - 1, N - 1], [0, 0], [0, N - 2]]
    assert got["nnn"] == [[1, 0], [N - 1, 0], [1, N - 2], [N - 1, N - 2]]


def _read_rows(path):
    rows = [l for l in open(path).read().splitlines() if not l.startswith("#")]
    return [[float(x) for x in r.split(",")] for r in rows]


def test_critical_exponent_file_and_batch_parity(tmp_path):
    import mcrg_amd
    from mcrg_amd.workloads import splitmix64
    N, n_eq, n, seed = 16, 200, 4000, 5
    out = _run([RUN, "--N", str(N), "--eq", str(n_eq), "--samples", str(n), "--K", "Kc", "--seed", str(seed),
                "--outdir", str(tmp_path)], tmp_path)
    assert "MONTE CARLO" in out and "RENORMALIZATION GROUP" in out
    assert "* Critical exponent: nu =" in out
    name = tmp_path / "critical_exponent_N_16_K_-0.4406868.txt"
    text = open(name).read().splitlines()
    assert text[0] == "# Number of parallel processes = 1"
    assert text[1] == f"# Number of equilibration samples = {n_eq}"
    assert text[2] == f"# Number of samples = {n}"
    assert text[3] == "# %23s  %25s  %25s" % ("Blocking Level n", "Largest Eigenvalue", "Critical Exponent nu")
    rows = _read_rows(name)
    assert [int(r[0]) for r in rows] == [0, 1, 2]  # floor(log2 16) - 1 levels
    for line in text[4:]:
        assert len(line) == 25 + 2 + 25 + 2 + 25
    # the same chain through the batch entry points: same seed, levels and bins
    job = dict(L=N, coupling=-KC_REF, seed=splitmix64(seed), n_therm=n_eq, n_meas=n, n_levels=3,
               sweeps_per_bin=(n + 49) // 50)
    with mcrg_amd.Context(0) as ctx, mcrg_amd.Batch(ctx, [job]) as b:
        b.run_all()
        rec = b.records(0)
    for lvl, row in enumerate(rows):
        r = mcrg_amd.rg_exponent(rec, 3, lvl, 2)
        assert row[1] == pytest.approx(r["lambda"], abs=1e-9)
        assert row[2] == pytest.approx(r["nu"], abs=1e-9)
    err = _read_rows(tmp_path / "critical_exponent_N_16_K_-0.4406868.err.txt")
    assert len(err) == 3 and all(e[2] > 0 for e in err)
    # the last column is the binning tau_int of NN per level, as mcrg_tau_int gives it
    for lvl, e in enumerate(err):
        t = mcrg_amd.tau_int(rec, 3, lvl, mcrg_amd.OP_NN)
        assert e[4] == pytest.approx(t["tau_int"], abs=1e-4) and t["tau_int"] > 0


@pytest.mark.parametrize("layout", ["auto", "int8"])
def test_replica_records_sum_to_the_oracle_chains(orc, tmp_path, layout):
    """calc_critical_exponent with 3 replicas (the reference's MPI ranks, an
    uneven split_samples share) and 5 bins: each level's lambda in the
    reference-format file equals the one formed from the per-bin sum of the
    three oracle chains with the same seeds, on the batched (auto) and on the
    int8 chain layout."""
    from mcrg_amd.workloads import splitmix64
    N, n_eq, n, R, bins, seed = 16, 50, 301, 3, 5, 9
    K = -0.4406868
    _run([RUN, "--N", str(N), "--eq", str(n_eq), "--samples", str(n), "--K", repr(K), "--replicas", str(R),
          "--bins", str(bins), "--seed", str(seed), "--layout", layout, "--outdir", str(tmp_path), "--quiet"],
         tmp_path)
    rows = _read_rows(tmp_path / "critical_exponent_N_16_K_-0.4406868.txt")
    nl = 3
    loc = [orc.split_samples(r, R, n) for r in range(R)]
    assert loc == [99, 101, 101]
    spb = -(-max(loc) // bins)
    n_bins = -(-max(loc) // spb)
    total = np.zeros((n_bins, orc.record_size(nl)), dtype=np.int64)
    for r in range(R):
        _, rec = orc.run_chain(N, -K, splitmix64(seed + r), n_eq, loc[r], nl, spb)
        total[:rec.shape[0]] += rec
    tot = total.sum(axis=0)
    assert len(rows) == nl
    for lvl, row in enumerate(rows):
        lam = orc.leading_eigenvalue(orc.rg_matrix(tot, nl, lvl, 2))
        assert row[1] == pytest.approx(lam, abs=2e-10), (lvl, row, lam)
        assert row[2] == pytest.approx(math.log(2) / math.log(lam), abs=2e-10)


def test_replicas_and_hbm_schedule(tmp_path):
    # N = 4 runs on the HBM schedule (below the LDS-resident minimum), 3 replicas
    _run([RUN, "--N", "4", "--eq", "50", "--samples", "300", "--K", "-0.3", "--replicas", "3",
          "--outdir", str(tmp_path), "--quiet"], tmp_path)
    text = open(tmp_path / "critical_exponent_N_4_K_-0.3.txt").read()
    assert "# Number of parallel processes = 3" in text
    assert len(_read_rows(tmp_path / "critical_exponent_N_4_K_-0.3.txt")) == 1


def test_locate_critical_point_file(tmp_path):
    out = _run([RUN, "--locate", "2", "--L", "16", "--K0", "-0.44", "--eq", "200", "--samples", "3000",
                "--outdir", str(tmp_path)], tmp_path)
    assert "* Critical point: Kc =" in out
    path = tmp_path / "critical_point_L_16_K_-0.44.txt"
    text = open(path).read().splitlines()
    assert text[3] == "# %23s  %25s  %25s  %25s" % ("Iteration", "Blocking Level n", "Starting K",
                                                    "Approximate Kc")
    rows = _read_rows(path)
    assert [(int(r[0]), int(r[1])) for r in rows] == [(1, 0), (1, 1), (1, 2), (2, 0), (2, 1), (2, 2)]
    assert rows[0][2] == pytest.approx(-0.44)
    assert rows[3][2] == pytest.approx(rows[2][3])  # iteration 2 starts from iteration 1's deepest level
    for r in rows:
        assert -0.5 < r[3] < -0.38


def test_cli_rejects_bad_coupling(tmp_path):
    out = subprocess.run([RUN, "--N", "16", "--K", "0.3", "--eq", "1", "--samples", "10", "--outdir",
                          str(tmp_path)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode != 0 and "ferromagnetic" in out.stderr


def test_bisection_driver(tmp_path):
    """K-bisection with two lattice sizes in one batch (L = 64 on the LDS-resident
    schedule, L = 16 beside it): a crossing inside the first bracket, each
    round's bracket centred on the previous crossing, every round's crossing
    consistent with the last."""
    out = _run([RUN, "--bisect", "2", "--n-k", "8", "--n-large", "16", "--L-large", "64", "--n-small", "64",
                "--L-small", "16", "--eq", "500", "--samples", "4000", "--bins", "20", "--quiet"], tmp_path)
    r = json.loads(out.strip().splitlines()[-1])
    assert r["rounds"] == 2 and len(r["round_Jc"]) == 2
    (lo1, hi1, jc1, e1), (lo2, hi2, jc2, e2) = r["round_Jc"]
    assert (lo1, hi1) == (0.43, 0.45)
    assert lo1 < jc1 < hi1 and e1 > 0
    assert (lo2 + hi2) / 2 == pytest.approx(jc1, abs=1e-9)
    assert r["Jc"] == jc2 and r["Jc_err"] == e2
    assert abs(r["Jc"] - r["exact"]) < 0.01
    # no round's crossing is biased by the width of its bracket: each lies
    # within 5 standard errors of the final one
    for lo, hi, jc, e in r["round_Jc"]:
        assert abs(jc - r["Jc"]) < 5.0 * max(e, r["Jc_err"])
    assert r["site_updates"] == 2 * (16 * 64 * 64 + 64 * 16 * 16) * 4500


def test_equilibrate_log_matches_oracle(orc, tmp_path):
    """IsingModel::equilibrate(.., write = true): the reference's file name,
    header and write_iter cadence (src/ising.cpp:14-84,
    src/definitions.cpp:44-68); E/spin and |M|/spin of the chain after n
    sweeps equal the oracle's."""
    N, K = 16, -0.4406868
    got = json.loads(_run([PROBE, str(N), repr(K), "eq"], tmp_path))
    lines = open(tmp_path / "equilibrate_N_16_K_-0.4406868.txt").read().splitlines()
    assert lines[0] == "# Nearest neighbor coupling K = -0.440687"
    assert lines[2] == f"# Number of lattice sites = {N * N}"
    assert lines[5].startswith("# Iteration, Avg E/spin")
    rows = [l for l in lines[6:] if l.count(",") == 6]
    iters = [int(r.split(",")[0]) for r in rows]
    assert iters == list(range(1, 10)) + list(range(10, 101, 10))
    seed = got["seed2"]
    s = orc.init_spins(N, seed)
    t = 0
    for r in rows:
        n, E, _, _, M = [float(x) for x in r.split(",")[:5]]
        while t < n:
            orc.sweep(N, s, seed, -K, t)
            t += 1
        nn = orc.ops(N, s)[0]
        assert E == pytest.approx(K * 2 * nn / N**2, rel=1e-6, abs=1e-12)
        assert M == pytest.approx(abs(int(s.astype(int).sum())) / N**2, rel=1e-6, abs=1e-12)
