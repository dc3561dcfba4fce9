"""GPU tests of the continuous-batching job server (mcrg_server).

Every job's records and final state are the same whatever else is resident:
jobs admitted mid-stream, under a budget that only holds a few lattices at a
time, or driven by the scheduler thread match the oracle bit for bit.  Error
paths: malformed jobs, jobs larger than the whole budget, unknown ids, results
read before the job is done, and cancellation.
"""
import math
import time

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


def _jobs(seed, n, Ls=(8, 16, 32, 64)):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        L = int(rng.choice(Ls))
        out.append(dict(L=L, coupling=float(rng.choice([0.4320459, KC, 0.4688157, 0.0])),
                        seed=int(rng.integers(0, 2**63)), n_therm=int(rng.integers(0, 30)),
                        n_meas=int(rng.integers(0, 60)), n_levels=int(rng.integers(0, int(math.log2(L)))),
                        sweeps_per_bin=int(rng.integers(1, 9))))
    return out


def _expect(orc, job):
    return orc.run_chain(job["L"], job["coupling"], job["seed"], job["n_therm"], job["n_meas"], job["n_levels"],
                         job["sweeps_per_bin"])[1]


@pytest.mark.parametrize("chunk,quantum", [(7, 0), (0, 300), (0, 1)])
def test_stream_bit_exact_vs_oracle(amd, ctx, orc, chunk, quantum):
    """Jobs arrive in three waves between launches; each equals the oracle,
    with uniform chunks and with time slices that differ per slot."""
    jobs = _jobs(1, 36)
    print("stream", chunk, quantum, flush=True)
    with amd.Server(ctx, sweeps_per_launch=chunk, launch_quantum=quantum) as srv:
        ids = []
        for wave in range(3):
            for job in jobs[wave * 12:(wave + 1) * 12]:
                ids.append(srv.submit(job))
            srv.step()
            srv.step()
        for _ in range(1000):
            if srv.step() == 0 and srv.stats()["queued"] == 0:
                break
        st = srv.stats()
        assert st["submitted"] == 36 and st["admitted"] == 36 and st["completed"] == 36
        assert st["running"] == 0 and st["queued"] == 0 and st["hbm_in_use"] == 0 and st["lds_in_use"] == 0
        done = srv.completed(16) + srv.completed()
        assert sorted(done) == sorted(ids) and srv.completed() == []
        for jid, job in zip(ids, jobs):
            p = srv.poll(jid)
            assert p["state"] == amd.JOB_DONE
            assert p["submit_s"] <= p["start_s"] <= p["finish_s"]
            got = srv.records(jid)
            assert np.array_equal(got, _expect(orc, job)), f"records differ for {job}"
            srv.release(jid)
        with pytest.raises(amd.MCRGError) as e:
            srv.poll(ids[0])
        assert e.value.status == amd.ERR_NOT_FOUND


def test_more_slots_than_cus_bit_exact(amd, ctx, orc):
    """320 L = 128 jobs are 320 single-lattice slots, more than the 256 CUs:
    each launch splits every slot's sweeps into parts that other workgroups
    may take; the jobs' 60..260 sweeps span several parts and launches, and
    the results equal the oracle's."""
    rng = np.random.default_rng(7)
    jobs = [dict(L=128, coupling=float(rng.choice([0.4320459, KC, 0.4688157])), seed=5000 + i,
                 n_therm=int(rng.integers(0, 60)), n_meas=int(rng.integers(60, 200)), n_levels=5,
                 sweeps_per_bin=int(rng.integers(10, 70)))
            for i in range(320)]
    print("server parts", len(jobs), flush=True)
    with amd.Server(ctx) as srv:
        ids = [srv.submit(j) for j in jobs]
        for _ in range(1000):
            if srv.step() == 0 and srv.stats()["queued"] == 0:
                break
        assert srv.stats()["completed"] == len(jobs)
        for jid, job in list(zip(ids, jobs))[::8]:
            assert np.array_equal(srv.records(jid), _expect(orc, job)), f"records differ for {job}"


def test_small_budget_queues_and_reuses_memory(amd, ctx, orc):
    """A budget of three L=64 lattices: admissions wait for retirements, and the
    reused device memory starts clean for every job."""
    jobs = [dict(L=64, coupling=KC, seed=100 + i, n_therm=5 + 3 * i, n_meas=30, n_levels=3, sweeps_per_bin=6)
            for i in range(8)]
    # 64^2 bits of spins, then 5 bins of records, each rounded to 256 bytes
    per_job = 512 + ((5 * amd.record_size(3) * 8 + 255) // 256) * 256
    with amd.Server(ctx, hbm_budget_bytes=3 * per_job, sweeps_per_launch=9) as srv:
        ids = [srv.submit(j) for j in jobs]
        peak = 0
        for _ in range(500):
            n = srv.step()
            st = srv.stats()
            peak = max(peak, st["running"])
            assert st["hbm_in_use"] <= 3 * per_job
            if n == 0 and st["queued"] == 0:
                break
        assert 1 <= peak <= 3
        for jid, job in zip(ids, jobs):
            assert np.array_equal(srv.records(jid), _expect(orc, job))


def test_scheduler_thread(amd, ctx, orc):
    jobs = _jobs(2, 20, Ls=(16, 32, 64, 128))
    with amd.Server(ctx, sweeps_per_launch=16) as srv:
        srv.start()
        ids = [srv.submit(j) for j in jobs]
        deadline = time.time() + 60
        while time.time() < deadline:
            if all(srv.poll(i)["state"] == amd.JOB_DONE for i in ids):
                break
            time.sleep(0.01)
        srv.stop()
        for jid, job in zip(ids, jobs):
            assert srv.poll(jid)["state"] == amd.JOB_DONE
            assert np.array_equal(srv.records(jid), _expect(orc, job))
        st = srv.stats()
        assert st["completed"] == 20 and st["launches"] >= 1 and st["busy_s"] > 0
        assert st["site_updates"] == sum(j["L"] ** 2 * (j["n_therm"] + j["n_meas"]) for j in jobs)


def test_lds_budget_limits_residency(amd, ctx, orc):
    """An LDS budget of about two L=128 lattices holds the rest in the queue."""
    jobs = [dict(L=128, coupling=KC, seed=200 + i, n_therm=3, n_meas=12, n_levels=4, sweeps_per_bin=4)
            for i in range(5)]
    with amd.Server(ctx, sweeps_per_launch=5) as probe:
        probe.submit(jobs[0])
        probe.step()
        share = probe.stats()["lds_in_use"]
    assert share > 0
    with amd.Server(ctx, lds_budget_bytes=2 * share + share // 2, sweeps_per_launch=5) as srv:
        ids = [srv.submit(j) for j in jobs]
        srv.step()
        st = srv.stats()
        assert st["running"] == 2 and st["queued"] == 3 and st["lds_in_use"] == 2 * share
        for _ in range(100):
            if srv.step() == 0 and srv.stats()["queued"] == 0:
                break
        for jid, job in zip(ids, jobs):
            assert np.array_equal(srv.records(jid), _expect(orc, job))
    with amd.Server(ctx, lds_budget_bytes=share - 1) as small:
        with pytest.raises(amd.MCRGError) as e:
            small.submit(jobs[0])
        assert e.value.status == amd.ERR_OUT_OF_BUDGET


def test_large_lattices_follow_the_queue(amd, ctx, orc):
    """L = 256 and 512 jobs under an LDS budget of about one and a half L = 512
    lattices: the large class runs on 512-lane workgroups while jobs wait and
    on 1024-lane ones once the queue is empty, switching while jobs are
    resident; every job equals the oracle."""
    jobs = [dict(L=L, coupling=KC, seed=400 + i, n_therm=2, n_meas=9, n_levels=nl, sweeps_per_bin=3)
            for i, (L, nl) in enumerate([(512, 3), (256, 4), (512, 7), (256, 2), (512, 5), (256, 6)])]
    with amd.Server(ctx, sweeps_per_launch=4) as probe:
        probe.submit(jobs[0])
        probe.step()
        share = probe.stats()["lds_in_use"]
    assert share > 0
    with amd.Server(ctx, lds_budget_bytes=share + share // 2, sweeps_per_launch=4) as srv:
        ids = [srv.submit(j) for j in jobs]
        queued_seen = False
        for _ in range(400):
            active = srv.step()
            queued_seen = queued_seen or srv.stats()["queued"] > 0
            if active == 0 and srv.stats()["queued"] == 0:
                break
        assert queued_seen
        for jid, job in zip(ids, jobs):
            assert np.array_equal(srv.records(jid), _expect(orc, job)), job


def test_errors_and_cancel(amd, ctx):
    with amd.Server(ctx, hbm_budget_bytes=1 << 16, sweeps_per_launch=5) as srv:
        for bad in (dict(L=12, coupling=0.4, seed=1, n_meas=4, sweeps_per_bin=1),
                    dict(L=1024, coupling=0.4, seed=1, n_meas=4, sweeps_per_bin=1),
                    dict(L=16, coupling=-0.4, seed=1, n_meas=4, sweeps_per_bin=1),
                    dict(L=16, coupling=0.4, seed=1, n_meas=4, n_levels=4, sweeps_per_bin=1),
                    dict(L=16, coupling=0.4, seed=1, n_meas=4, sweeps_per_bin=0),
                    dict(L=16, coupling=0.4, seed=1, n_meas=4, sweeps_per_bin=1, sampler=1),
                    dict(L=512, coupling=0.4, seed=1, n_meas=4, sweeps_per_bin=40000000)):
            with pytest.raises(amd.MCRGError) as e:
                srv.submit(bad)
            assert e.value.status == amd.ERR_INVALID_ARG
        # 1000 bins of records (480 KB) exceed the whole 64 KiB budget
        with pytest.raises(amd.MCRGError) as e:
            srv.submit(dict(L=512, coupling=0.4, seed=1, n_meas=1000, n_levels=2, sweeps_per_bin=1))
        assert e.value.status == amd.ERR_OUT_OF_BUDGET
        with pytest.raises(amd.MCRGError) as e:
            srv.poll(12345)
        assert e.value.status == amd.ERR_NOT_FOUND
        long_id = srv.submit(dict(L=32, coupling=KC, seed=3, n_therm=10, n_meas=10000, n_levels=2,
                                  sweeps_per_bin=100))
        queued_id = srv.submit(dict(L=32, coupling=KC, seed=4, n_meas=10, n_levels=2, sweeps_per_bin=5))
        srv.release(queued_id)  # cancelled while queued
        srv.step()
        assert srv.poll(long_id)["state"] == amd.JOB_RUNNING
        assert srv.poll(queued_id)["state"] == amd.JOB_CANCELLED
        with pytest.raises(amd.MCRGError) as e:
            srv.records(long_id)
        assert e.value.status == amd.ERR_NOT_READY
        srv.release(long_id)  # cancelled while resident
        assert srv.step() == 0
        assert srv.poll(long_id)["state"] == amd.JOB_CANCELLED
        st = srv.stats()
        assert st["cancelled"] == 2 and st["hbm_in_use"] == 0 and st["running"] == 0


def test_cancel_after_last_launch_before_retirement(amd, ctx, orc):
    """A job's last launch is queued before its records are read back (the
    next iteration retires it): it still reads as running after that step and
    completes in the next one.  Cancelled in between, it never completes, and a
    job admitted into the memory it frees, beside one that keeps running, both
    match the oracle."""
    short = dict(L=32, coupling=KC, seed=11, n_therm=0, n_meas=5, n_levels=2, sweeps_per_bin=5)
    longer = dict(L=32, coupling=0.4320459, seed=12, n_therm=3, n_meas=40, n_levels=2, sweeps_per_bin=4)
    later = dict(L=32, coupling=0.4688157, seed=13, n_therm=2, n_meas=12, n_levels=3, sweeps_per_bin=3)
    with amd.Server(ctx, sweeps_per_launch=5) as srv:
        a = srv.submit(short)
        b = srv.submit(longer)
        srv.step()                      # a's only launch; b's first
        assert srv.poll(a)["state"] == amd.JOB_RUNNING
        srv.release(a)                  # cancelled after its last launch
        c = srv.submit(later)
        while srv.step() > 0:
            pass
        assert srv.poll(a)["state"] == amd.JOB_CANCELLED
        for jid, job in ((b, longer), (c, later)):
            assert srv.poll(jid)["state"] == amd.JOB_DONE
            assert np.array_equal(srv.records(jid).reshape(-1), _expect(orc, job).reshape(-1))
        st = srv.stats()
        assert st["cancelled"] == 1 and st["completed"] == 2 and st["hbm_in_use"] == 0

    with amd.Server(ctx, sweeps_per_launch=5) as srv:
        a = srv.submit(short)
        srv.step()                      # queued and run; records not read yet
        assert srv.poll(a)["state"] == amd.JOB_RUNNING
        assert srv.step() == 0          # retired
        assert srv.poll(a)["state"] == amd.JOB_DONE
        assert np.array_equal(srv.records(a).reshape(-1), _expect(orc, short).reshape(-1))
