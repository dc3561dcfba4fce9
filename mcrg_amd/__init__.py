"""mcrg_amd -- MI355X-native MCRG hot path (checkerboard Metropolis, b=2 majority
blocking, exact int64 even-operator correlations) behind a C ABI.

Python here is a thin ctypes view of libmcrg_amd.so for tests and benchmarks;
the hot path is the HIP code in csrc/.  Importing this package raises if the
library has not been built (``make -C mcrg_amd``): there is no fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmcrg_amd.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is missing: build it with `make -C {_HERE}` (hipcc, gfx950)")

_lib = ctypes.CDLL(LIB_PATH)

OK = 0
ERR_INVALID_ARG = -1
ERR_OUT_OF_BUDGET = -2
ERR_DEVICE = -3
ERR_NOT_READY = -4
ERR_NOT_FOUND = -5
ERR_CAPACITY = -6
LAYOUT_INT8 = 0
LAYOUT_PACKED = 1
SAMPLER_METROPOLIS = 0
SAMPLER_SWENDSEN_WANG = 1
JOB_QUEUED = 0
JOB_RUNNING = 1
JOB_DONE = 2
JOB_CANCELLED = 3


class MCRGError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"mcrg status {status}: {msg}")
        self.status = status


class ChainDesc(ctypes.Structure):
    _fields_ = [("L", ctypes.c_int32), ("coupling", ctypes.c_double), ("seed", ctypes.c_uint64),
                ("n_levels", ctypes.c_int32), ("layout", ctypes.c_int32),
                ("sweeps_per_bin", ctypes.c_int32), ("max_bins", ctypes.c_int32), ("sampler", ctypes.c_int32)]


class JobDesc(ctypes.Structure):
    _fields_ = [("L", ctypes.c_int32), ("coupling", ctypes.c_double), ("seed", ctypes.c_uint64),
                ("n_therm", ctypes.c_int32), ("n_meas", ctypes.c_int32), ("n_levels", ctypes.c_int32),
                ("sweeps_per_bin", ctypes.c_int32), ("sampler", ctypes.c_int32)]


class BatchOpts(ctypes.Structure):
    _fields_ = [("threads", ctypes.c_int32), ("sweeps_per_launch", ctypes.c_int32),
                ("max_group", ctypes.c_int32), ("max_slot_lds", ctypes.c_int32), ("min_team", ctypes.c_int32),
                ("max_team", ctypes.c_int32), ("threads_large", ctypes.c_int32), ("autotune", ctypes.c_int32),
                ("sub_chunks", ctypes.c_int32)]


class RGResult(ctypes.Structure):
    _fields_ = [("lambda_", ctypes.c_double), ("yt", ctypes.c_double), ("nu", ctypes.c_double),
                ("yt_err", ctypes.c_double), ("nu_err", ctypes.c_double), ("n_jackknife", ctypes.c_int32)]


class TauResult(ctypes.Structure):
    _fields_ = [("mean", ctypes.c_double), ("var", ctypes.c_double), ("tau_int", ctypes.c_double),
                ("tau_err", ctypes.c_double), ("tau_fine", ctypes.c_double), ("bin_samples", ctypes.c_int32),
                ("n_bins", ctypes.c_int32)]


class ServerOpts(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_int64), ("lds_budget_bytes", ctypes.c_int64),
                ("launch_quantum", ctypes.c_int64), ("max_jobs", ctypes.c_int32), ("threads", ctypes.c_int32),
                ("sweeps_per_launch", ctypes.c_int32), ("threads_large", ctypes.c_int32)]


class JobStatus(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int32), ("n_bins", ctypes.c_int32), ("submit_s", ctypes.c_double),
                ("start_s", ctypes.c_double), ("finish_s", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int64), ("admitted", ctypes.c_int64), ("completed", ctypes.c_int64),
                ("cancelled", ctypes.c_int64), ("queued", ctypes.c_int64), ("running", ctypes.c_int64),
                ("launches", ctypes.c_int64), ("site_updates", ctypes.c_double), ("busy_s", ctypes.c_double),
                ("device_wait_s", ctypes.c_double),
                ("hbm_in_use", ctypes.c_int64), ("lds_in_use", ctypes.c_int64)]


class DeviceInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 64), ("compute_units", ctypes.c_int32), ("clock_khz", ctypes.c_int32),
                ("lds_per_block", ctypes.c_int32), ("warp_size", ctypes.c_int32),
                ("global_mem_bytes", ctypes.c_int64)]


_c = ctypes
_p = ctypes.c_void_p
_SIGS = {
    "mcrg_abi_version": (_c.c_int32, []),
    "mcrg_last_error": (_c.c_char_p, []),
    "mcrg_thresholds": (_c.c_int32, [_c.c_double, _c.POINTER(_c.c_uint64)]),
    "mcrg_record_size": (_c.c_int32, [_c.c_int32]),
    "mcrg_bond_threshold": (_c.c_int32, [_c.c_double, _c.POINTER(_c.c_uint32)]),
    "mcrg_ctx_create": (_c.c_int32, [_c.c_int32, _c.POINTER(_p)]),
    "mcrg_ctx_destroy": (_c.c_int32, [_p]),
    "mcrg_ctx_set_timing": (_c.c_int32, [_p, _c.c_int32]),
    "mcrg_ctx_device_info": (_c.c_int32, [_p, _c.POINTER(DeviceInfo)]),
    "mcrg_ctx_kernel_time": (_c.c_int32, [_p, _c.c_int32, _c.POINTER(_c.c_int64), _c.POINTER(_c.c_double)]),
    "mcrg_chain_create": (_c.c_int32, [_p, _c.POINTER(ChainDesc), _c.POINTER(_p)]),
    "mcrg_chain_destroy": (_c.c_int32, [_p]),
    "mcrg_chain_set_coupling": (_c.c_int32, [_p, _c.c_double]),
    "mcrg_chain_sweep": (_c.c_int32, [_p, _c.c_int32, _c.c_int32]),
    "mcrg_chain_sweep_static": (_c.c_int32, [_p, _c.c_int32, _c.c_int32]),
    "mcrg_chain_sync": (_c.c_int32, [_p]),
    "mcrg_chain_get_spins": (_c.c_int32, [_p, _p, _c.c_int64]),
    "mcrg_chain_set_spins": (_c.c_int32, [_p, _p, _c.c_int64]),
    "mcrg_chain_get_records": (_c.c_int32, [_p, _p, _c.c_int64]),
    "mcrg_chain_get_state": (_c.c_int32, [_p, _c.POINTER(_c.c_uint32), _c.POINTER(_c.c_uint32)]),
    "mcrg_chain_set_state": (_c.c_int32, [_p, _c.c_uint32, _c.c_uint32, _p, _c.c_int64]),
    "mcrg_chain_operators": (_c.c_int32, [_p, _c.POINTER(_c.c_int64)]),
    "mcrg_batch_create": (_c.c_int32, [_p, _c.POINTER(JobDesc), _c.c_int32, _c.POINTER(BatchOpts),
                                       _c.POINTER(_p)]),
    "mcrg_batch_destroy": (_c.c_int32, [_p]),
    "mcrg_batch_run": (_c.c_int32, [_p, _c.c_int32, _c.POINTER(_c.c_int32)]),
    "mcrg_batch_sync": (_c.c_int32, [_p]),
    "mcrg_batch_get_records": (_c.c_int32, [_p, _c.c_int32, _p, _c.c_int64]),
    "mcrg_batch_get_spins": (_c.c_int32, [_p, _c.c_int32, _p, _c.c_int64]),
    "mcrg_batch_get_state": (_c.c_int32, [_p, _c.c_int32, _c.POINTER(_c.c_uint32), _c.POINTER(_c.c_uint32)]),
    "mcrg_batch_get_opts": (_c.c_int32, [_p, _c.POINTER(BatchOpts)]),
    "mcrg_server_create": (_c.c_int32, [_p, _c.POINTER(ServerOpts), _c.POINTER(_p)]),
    "mcrg_server_destroy": (_c.c_int32, [_p]),
    "mcrg_server_submit": (_c.c_int32, [_p, _c.POINTER(JobDesc), _c.POINTER(_c.c_uint64)]),
    "mcrg_server_poll": (_c.c_int32, [_p, _c.c_uint64, _c.POINTER(JobStatus)]),
    "mcrg_server_records": (_c.c_int32, [_p, _c.c_uint64, _p, _c.c_int64]),
    "mcrg_server_release": (_c.c_int32, [_p, _c.c_uint64]),
    "mcrg_server_completed": (_c.c_int32, [_p, _c.POINTER(_c.c_uint64), _c.c_int32, _c.POINTER(_c.c_int32)]),
    "mcrg_server_start": (_c.c_int32, [_p]),
    "mcrg_server_stop": (_c.c_int32, [_p]),
    "mcrg_server_step": (_c.c_int32, [_p, _c.POINTER(_c.c_int32)]),
    "mcrg_server_stats_get": (_c.c_int32, [_p, _c.POINTER(ServerStats)]),
    "mcrg_rg_matrix": (_c.c_int32, [_p, _c.c_int32, _c.c_int32, _c.c_int32, _c.POINTER(_c.c_double)]),
    "mcrg_leading_eigenvalue": (_c.c_double, [_c.c_int32, _c.POINTER(_c.c_double)]),
    "mcrg_rg_exponent": (_c.c_int32, [_p, _c.c_int32, _c.c_int32, _c.c_int32, _c.c_int32, _c.POINTER(RGResult)]),
    "mcrg_critical_point_step": (_c.c_int32, [_p, _c.c_int32, _p, _c.c_int32, _c.c_int32,
                                              _c.POINTER(_c.c_double)]),
    "mcrg_critical_point_step_bins": (_c.c_int32, [_p, _c.c_int32, _c.c_int32, _p, _c.c_int32, _c.c_int32,
                                                   _c.c_int32, _c.POINTER(_c.c_double)]),
    "mcrg_max_sweeps_per_bin": (_c.c_int64, [_c.c_int32]),
    "mcrg_tau_int": (_c.c_int32, [_p, _c.c_int32, _c.c_int32, _c.c_int32, _c.c_int32, _c.POINTER(TauResult)]),
}
for _name, (_res, _args) in _SIGS.items():
    _f = getattr(_lib, _name)
    _f.restype = _res
    _f.argtypes = _args


def lib():
    return _lib


def last_error():
    return _lib.mcrg_last_error().decode()


def check(status):
    if status != OK:
        raise MCRGError(status, last_error())
    return status


def abi_version():
    return _lib.mcrg_abi_version()


def thresholds(coupling):
    thr = (ctypes.c_uint64 * 2)()
    check(_lib.mcrg_thresholds(float(coupling), thr))
    return np.array(list(thr), dtype=np.uint64)


def bond_threshold(coupling):
    thr = ctypes.c_uint32()
    check(_lib.mcrg_bond_threshold(float(coupling), ctypes.byref(thr)))
    return thr.value


def record_size(n_levels):
    return _lib.mcrg_record_size(n_levels)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def rg_matrix(record, n_levels, level, n_ops):
    r = np.ascontiguousarray(record, dtype=np.int64).reshape(-1)
    T = (ctypes.c_double * (n_ops * n_ops))()
    check(_lib.mcrg_rg_matrix(_ptr(r), n_levels, level, n_ops, T))
    return np.array(list(T)).reshape(n_ops, n_ops)


def leading_eigenvalue(T):
    T = np.ascontiguousarray(T, dtype=np.float64)
    return _lib.mcrg_leading_eigenvalue(T.shape[0], T.reshape(-1).ctypes.data_as(ctypes.POINTER(ctypes.c_double)))


def rg_exponent(records, n_levels, level, n_ops=2):
    """dict(lambda, yt, nu, yt_err, nu_err, n_jackknife) from per-bin records."""
    r = np.ascontiguousarray(records, dtype=np.int64)
    r2 = r.reshape(-1, record_size(n_levels))
    out = RGResult()
    check(_lib.mcrg_rg_exponent(_ptr(r2), r2.shape[0], n_levels, level, n_ops, ctypes.byref(out)))
    return {"lambda": out.lambda_, "yt": out.yt, "nu": out.nu, "yt_err": out.yt_err, "nu_err": out.nu_err,
            "n_jackknife": out.n_jackknife}


OP_NN, OP_NNN, OP_P, OP_ABS_M = 0, 1, 2, 3


def tau_int(records, n_levels, level=0, op=OP_NN):
    """Integrated autocorrelation time (in samples) of an operator from per-bin
    records, by binning: dict(mean, var, tau_int, tau_err, tau_fine,
    bin_samples, n_bins)."""
    r = np.ascontiguousarray(records, dtype=np.int64)
    r2 = r.reshape(-1, record_size(n_levels))
    out = TauResult()
    check(_lib.mcrg_tau_int(_ptr(r2), r2.shape[0], n_levels, level, op, ctypes.byref(out)))
    return {k: getattr(out, k) for k, _ in TauResult._fields_}


def critical_point_step(rec_large, nl_large, rec_small, nl_small, level):
    """The reference's locator step dK from one record of each chain; with
    several bins per array (shape (bins, record_size)) the bins are summed in
    128-bit integers."""
    a = np.ascontiguousarray(rec_large, dtype=np.int64).reshape(-1, record_size(nl_large))
    b = np.ascontiguousarray(rec_small, dtype=np.int64).reshape(-1, record_size(nl_small))
    d = ctypes.c_double()
    check(_lib.mcrg_critical_point_step_bins(_ptr(a), a.shape[0], nl_large, _ptr(b), b.shape[0], nl_small, level,
                                             ctypes.byref(d)))
    return d.value


def max_sweeps_per_bin(L):
    """Most sweeps per record bin whose int64 sums cannot overflow at side L."""
    return _lib.mcrg_max_sweeps_per_bin(L)


class Context:
    """One device (mcrg_ctx)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        check(_lib.mcrg_ctx_create(device, ctypes.byref(h)))
        self.handle = h

    def device_info(self):
        d = DeviceInfo()
        check(_lib.mcrg_ctx_device_info(self.handle, ctypes.byref(d)))
        return {"name": d.name.decode(), "compute_units": d.compute_units, "clock_khz": d.clock_khz,
                "lds_per_block": d.lds_per_block, "warp_size": d.warp_size,
                "global_mem_bytes": d.global_mem_bytes}

    def set_timing(self, enable=True):
        check(_lib.mcrg_ctx_set_timing(self.handle, 1 if enable else 0))

    def kernel_time(self, kind):
        n = ctypes.c_int64()
        ms = ctypes.c_double()
        check(_lib.mcrg_ctx_kernel_time(self.handle, kind, ctypes.byref(n), ctypes.byref(ms)))
        return n.value, ms.value

    def close(self):
        if self.handle:
            _lib.mcrg_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Chain:
    """One Markov chain on the device (mcrg_chain)."""

    def __init__(self, ctx, L, coupling, seed, n_levels=0, layout=LAYOUT_INT8,
                 sweeps_per_bin=None, max_bins=1, sampler=SAMPLER_METROPOLIS):
        if sweeps_per_bin is None:  # 1000, or fewer where one bin's int64 sums could overflow
            sweeps_per_bin = max(1, min(1000, max_sweeps_per_bin(L)))
        d = ChainDesc(L, float(coupling), seed, n_levels, layout, sweeps_per_bin, max_bins, sampler)
        h = ctypes.c_void_p()
        check(_lib.mcrg_chain_create(ctx.handle, ctypes.byref(d), ctypes.byref(h)))
        self.handle = h
        self.L, self.n_levels, self.max_bins = L, n_levels, max_bins

    def sweep(self, n, measure=False):
        check(_lib.mcrg_chain_sweep(self.handle, n, 1 if measure else 0))

    def sweep_static(self, n, measure=False):
        """mcrg_chain_sweep_static: the same sweeps without allocating."""
        check(_lib.mcrg_chain_sweep_static(self.handle, n, 1 if measure else 0))

    def sync(self):
        check(_lib.mcrg_chain_sync(self.handle))

    def set_coupling(self, coupling):
        check(_lib.mcrg_chain_set_coupling(self.handle, float(coupling)))

    def spins(self):
        a = np.zeros(self.L * self.L, dtype=np.int8)
        check(_lib.mcrg_chain_get_spins(self.handle, _ptr(a), a.size))
        return a

    def set_spins(self, s):
        a = np.ascontiguousarray(s, dtype=np.int8).reshape(-1)
        check(_lib.mcrg_chain_set_spins(self.handle, _ptr(a), a.size))

    def records(self):
        rs = record_size(self.n_levels)
        a = np.zeros(self.max_bins * rs, dtype=np.int64)
        check(_lib.mcrg_chain_get_records(self.handle, _ptr(a), a.size))
        return a.reshape(self.max_bins, rs)

    def state(self):
        t, q = ctypes.c_uint32(), ctypes.c_uint32()
        check(_lib.mcrg_chain_get_state(self.handle, ctypes.byref(t), ctypes.byref(q)))
        return t.value, q.value

    def set_state(self, sweep, n_measured, records=None):
        if records is None:
            check(_lib.mcrg_chain_set_state(self.handle, sweep, n_measured, None, 0))
        else:
            r = np.ascontiguousarray(records, dtype=np.int64).reshape(-1)
            check(_lib.mcrg_chain_set_state(self.handle, sweep, n_measured, _ptr(r), r.size))

    def operators(self):
        out = (ctypes.c_int64 * 4)()
        check(_lib.mcrg_chain_operators(self.handle, out))
        return np.array(list(out), dtype=np.int64)

    def close(self):
        if self.handle:
            _lib.mcrg_chain_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def n_bins(n_meas, sweeps_per_bin):
    return max(1, -(-n_meas // sweeps_per_bin))


class Batch:
    """Independent chains advanced together on the LDS-resident schedule
    (mcrg_batch).  `jobs` is a list of dicts with keys L, coupling, seed,
    n_therm, n_meas, n_levels, sweeps_per_bin."""

    def __init__(self, ctx, jobs, threads=0, sweeps_per_launch=0, max_group=0, max_slot_lds=0, min_team=0,
                 max_team=0, threads_large=0, autotune=False, sub_chunks=0):
        arr = (JobDesc * len(jobs))()
        for i, j in enumerate(jobs):
            arr[i] = JobDesc(j["L"], float(j["coupling"]), j["seed"], j.get("n_therm", 0), j.get("n_meas", 0),
                             j.get("n_levels", 0), j.get("sweeps_per_bin", 1), j.get("sampler", 0))
        opts = BatchOpts(threads, sweeps_per_launch, max_group, max_slot_lds, min_team, max_team, threads_large,
                         1 if autotune else 0, sub_chunks)
        h = ctypes.c_void_p()
        check(_lib.mcrg_batch_create(ctx.handle, arr, len(jobs), ctypes.byref(opts), ctypes.byref(h)))
        self.handle = h
        self.jobs = [dict(j) for j in jobs]

    def run(self, max_sweeps):
        rem = ctypes.c_int32()
        check(_lib.mcrg_batch_run(self.handle, max_sweeps, ctypes.byref(rem)))
        return rem.value

    def run_all(self):
        while self.run(1 << 30) > 0:
            pass

    def sync(self):
        check(_lib.mcrg_batch_sync(self.handle))

    def records(self, j):
        job = self.jobs[j]
        rs = record_size(job.get("n_levels", 0))
        nb = n_bins(job.get("n_meas", 0), job.get("sweeps_per_bin", 1))
        a = np.zeros(nb * rs, dtype=np.int64)
        check(_lib.mcrg_batch_get_records(self.handle, j, _ptr(a), a.size))
        return a.reshape(nb, rs)

    def spins(self, j):
        L = self.jobs[j]["L"]
        a = np.zeros(L * L, dtype=np.int8)
        check(_lib.mcrg_batch_get_spins(self.handle, j, _ptr(a), a.size))
        return a

    def state(self, j):
        t, q = ctypes.c_uint32(), ctypes.c_uint32()
        check(_lib.mcrg_batch_get_state(self.handle, j, ctypes.byref(t), ctypes.byref(q)))
        return t.value, q.value

    def options(self):
        """Configuration in use (the autotuner's choice once it has finished)."""
        o = BatchOpts()
        check(_lib.mcrg_batch_get_opts(self.handle, ctypes.byref(o)))
        return {f: getattr(o, f) for f, _ in BatchOpts._fields_}

    def close(self):
        if self.handle:
            _lib.mcrg_batch_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Server:
    """Continuous-batching job server (mcrg_server): jobs are admitted while the
    device budget has room, advance one chunk per launch of the LDS-resident
    schedule and retire to host memory when done.  Drive it with step() from
    this thread, or start() a scheduler thread of its own."""

    def __init__(self, ctx, hbm_budget_bytes=0, lds_budget_bytes=0, launch_quantum=0, max_jobs=0, threads=0,
                 sweeps_per_launch=0, threads_large=0):
        opts = ServerOpts(hbm_budget_bytes, lds_budget_bytes, launch_quantum, max_jobs, threads, sweeps_per_launch,
                          threads_large)
        h = ctypes.c_void_p()
        check(_lib.mcrg_server_create(ctx.handle, ctypes.byref(opts), ctypes.byref(h)))
        self.handle = h
        self.jobs = {}

    def submit(self, job):
        d = JobDesc(job["L"], float(job["coupling"]), job["seed"], job.get("n_therm", 0), job.get("n_meas", 0),
                    job.get("n_levels", 0), job.get("sweeps_per_bin", 1), job.get("sampler", 0))
        jid = ctypes.c_uint64()
        check(_lib.mcrg_server_submit(self.handle, ctypes.byref(d), ctypes.byref(jid)))
        self.jobs[jid.value] = dict(job)
        return jid.value

    def poll(self, jid):
        st = JobStatus()
        check(_lib.mcrg_server_poll(self.handle, jid, ctypes.byref(st)))
        return dict(state=st.state, n_bins=st.n_bins, submit_s=st.submit_s, start_s=st.start_s,
                    finish_s=st.finish_s)

    def records(self, jid):
        job = self.jobs[jid]
        rs = record_size(job.get("n_levels", 0))
        nb = n_bins(job.get("n_meas", 0), job.get("sweeps_per_bin", 1))
        a = np.zeros(nb * rs, dtype=np.int64)
        check(_lib.mcrg_server_records(self.handle, jid, _ptr(a), a.size))
        return a.reshape(nb, rs)

    def release(self, jid):
        check(_lib.mcrg_server_release(self.handle, jid))

    def completed(self, max_ids=4096):
        """Ids of jobs finished since the last call."""
        buf = (ctypes.c_uint64 * max_ids)()
        n = ctypes.c_int32()
        check(_lib.mcrg_server_completed(self.handle, buf, max_ids, ctypes.byref(n)))
        return list(buf[:n.value])

    def step(self):
        active = ctypes.c_int32()
        check(_lib.mcrg_server_step(self.handle, ctypes.byref(active)))
        return active.value

    def start(self):
        check(_lib.mcrg_server_start(self.handle))

    def stop(self):
        check(_lib.mcrg_server_stop(self.handle))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stats(self):
        st = ServerStats()
        check(_lib.mcrg_server_stats_get(self.handle, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def close(self):
        if self.handle:
            _lib.mcrg_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
