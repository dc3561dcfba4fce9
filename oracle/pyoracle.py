"""ctypes binding of the CPU oracle (oracle.c).  TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py,
where it is the checker (or the timed CPU baseline); the product never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

i8p = np.ctypeslib.ndpointer(dtype=np.int8, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        c_int, c_u32, c_u64, c_d = ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double
        sig = {
            "orc_philox4x32_10": (None, [u32p, u32p, u32p]),
            "orc_thresholds": (None, [c_d, u64p]),
            "orc_init_spins": (None, [c_int, c_u64, i8p]),
            "orc_half_sweep": (None, [c_int, i8p, c_u64, u64p, c_u32, c_int]),
            "orc_sweep": (None, [c_int, i8p, c_u64, u64p, c_u32]),
            "orc_ops": (None, [c_int, i8p, i64p]),
            "orc_ops_refcount": (None, [c_int, i8p, f64p]),
            "orc_magnetization": (ctypes.c_int64, [c_int, i8p]),
            "orc_block": (None, [c_int, i8p, i8p, c_u64, c_u32, c_int]),
            "orc_record_size": (c_int, [c_int]),
            "orc_measure": (None, [c_int, i8p, c_int, c_u64, c_u32, i64p]),
            "orc_run_chain": (None, [c_int, c_d, c_u64, c_u32, c_int, c_int, c_int, c_int, i8p, i64p]),
            "orc_run_chain_sampler": (None, [c_int, c_d, c_u64, c_u32, c_int, c_int, c_int, c_int, c_int, i8p,
                                             i64p]),
            "orc_sw_threshold": (c_u32, [c_d]),
            "orc_sw_sweep": (None, [c_int, i8p, c_u64, c_u32, c_u32]),
            "orc_rg_matrix": (c_int, [i64p, c_int, c_int, c_int, f64p]),
            "orc_leading_eigenvalue": (c_d, [c_int, f64p]),
            "orc_yt": (c_d, [i64p, c_int, c_int, c_int]),
            "orc_jackknife_yt": (c_int, [i64p, c_int, c_int, c_int, c_int,
                                         ctypes.POINTER(c_d), ctypes.POINTER(c_d)]),
            "orc_split_samples": (c_int, [c_int, c_int, c_int]),
            "orc_kc_step": (c_int, [i64p, c_int, i64p, c_int, c_int, ctypes.POINTER(c_d)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


# ---------------------------------------------------------------- helpers
def philox(seed, ctr):
    key = np.array([seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF], dtype=np.uint32)
    c = np.array(ctr, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(c, key, out)
    return out


def thresholds(J):
    thr = np.zeros(2, dtype=np.uint64)
    lib().orc_thresholds(float(J), thr)
    return thr


def init_spins(L, seed):
    s = np.zeros(L * L, dtype=np.int8)
    lib().orc_init_spins(L, seed, s)
    return s


def sweep(L, s, seed, J, t):
    lib().orc_sweep(L, s, seed, thresholds(J), t)


def half_sweep(L, s, seed, J, t, color):
    lib().orc_half_sweep(L, s, seed, thresholds(J), t, color)


def ops(L, s):
    out = np.zeros(3, dtype=np.int64)
    lib().orc_ops(L, np.ascontiguousarray(s, dtype=np.int8), out)
    return out


def ops_refcount(L, s):
    out = np.zeros(2, dtype=np.float64)
    lib().orc_ops_refcount(L, np.ascontiguousarray(s, dtype=np.int8), out)
    return out


def block(L, s, seed, t, level):
    sb = np.zeros((L // 2) * (L // 2), dtype=np.int8)
    lib().orc_block(L, np.ascontiguousarray(s, dtype=np.int8), sb, seed, t, level)
    return sb


def record_size(nl):
    return lib().orc_record_size(nl)


SAMPLER_METROPOLIS = 0
SAMPLER_SW = 1


def sw_threshold(J):
    return lib().orc_sw_threshold(float(J))


def sw_sweep(L, s, seed, J, t):
    lib().orc_sw_sweep(L, s, seed, sw_threshold(J), t)


def run_chain(L, J, seed, n_therm, n_meas, nl, sweeps_per_bin, s=None, t0=0, sampler=SAMPLER_METROPOLIS):
    """Returns (final spins, records[n_bins, record_size])."""
    if s is None:
        s = init_spins(L, seed)
    s = np.ascontiguousarray(s, dtype=np.int8).copy()
    n_bins = max(1, -(-n_meas // sweeps_per_bin)) if n_meas > 0 else 1
    rec = np.zeros(n_bins * record_size(nl), dtype=np.int64)
    lib().orc_run_chain_sampler(L, float(J), seed, t0, n_therm, n_meas, nl, sweeps_per_bin, sampler, s, rec)
    return s, rec.reshape(n_bins, record_size(nl))


def rg_matrix(rec, nl, level, n_ops):
    T = np.zeros(n_ops * n_ops, dtype=np.float64)
    rc = lib().orc_rg_matrix(np.ascontiguousarray(rec, dtype=np.int64), nl, level, n_ops, T)
    if rc != 0:
        raise ValueError("singular RG matrix")
    return T.reshape(n_ops, n_ops)


def leading_eigenvalue(T):
    T = np.ascontiguousarray(T, dtype=np.float64)
    return lib().orc_leading_eigenvalue(T.shape[0], T.reshape(-1))


def yt(rec, nl, level, n_ops):
    return lib().orc_yt(np.ascontiguousarray(rec, dtype=np.int64), nl, level, n_ops)


def jackknife_yt(records, nl, level, n_ops):
    records = np.ascontiguousarray(records, dtype=np.int64)
    m, e = ctypes.c_double(), ctypes.c_double()
    lib().orc_jackknife_yt(records.reshape(-1), records.shape[0], nl, level, n_ops,
                           ctypes.byref(m), ctypes.byref(e))
    return m.value, e.value


def kc_step(rec_large, nl_large, rec_small, nl_small, level):
    d = ctypes.c_double()
    rc = lib().orc_kc_step(np.ascontiguousarray(rec_large, dtype=np.int64), nl_large,
                           np.ascontiguousarray(rec_small, dtype=np.int64), nl_small, level, ctypes.byref(d))
    if rc != 0:
        raise ValueError("undefined step")
    return d.value


def split_samples(rank, n_processes, n_samples):
    return lib().orc_split_samples(rank, n_processes, n_samples)


# record offsets (mirror of the ORC_* macros in oracle.h)
def off_S(nl, l, a):
    return l * 3 + a


def off_self(nl, l, a, b):
    return 3 * (nl + 1) + l * 9 + b * 3 + a


def off_cross(nl, l, a, b):
    return 12 * (nl + 1) + (l - 1) * 9 + b * 3 + a


def off_nn0(nl, l):
    return 12 * (nl + 1) + 9 * nl + l - 1


def off_M(nl):
    return 12 * (nl + 1) + 10 * nl


def off_N(nl):
    return off_M(nl) + 3
