"""ctypes view of the CPU oracle (oracle/liboracle_dto.so).

TEST INFRASTRUCTURE ONLY: the checker for the DTO hot path.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
The restated reference functions and their file:line citations live in
oracle/dto_oracle.c.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_dto.so")
_lib = None


class Record(ctypes.Structure):
    _fields_ = [
        ("rank1", ctypes.c_uint64),
        ("rank2", ctypes.c_uint64),
        ("set1_len", ctypes.c_uint64),
        ("set2_len", ctypes.c_uint64),
        ("population_size", ctypes.c_uint64),
        ("intersection_size", ctypes.c_uint64),
        ("pvalue", ctypes.c_double),
        ("permuted", ctypes.c_int32),
        ("i", ctypes.c_uint32),
        ("j", ctypes.c_uint32),
        ("ties", ctypes.c_uint32),
    ]

    def tie_counts(self):
        """(#cells at the minimum p, #of those at the largest intersection), saturated at 0xFFFF."""
        return (self.ties & 0xFFFF, self.ties >> 16)

    def as_tuple(self):
        return (self.rank1, self.rank2, self.set1_len, self.set2_len,
                self.population_size, self.intersection_size, self.pvalue)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        u64, u32, i32, dbl = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int32, ctypes.c_double
        P = ctypes.POINTER
        L.or_ln_gamma.argtypes = [dbl]; L.or_ln_gamma.restype = dbl
        L.or_ln_factorial.argtypes = [u64]; L.or_ln_factorial.restype = dbl
        L.or_ln_binomial.argtypes = [u64, u64]; L.or_ln_binomial.restype = dbl
        L.or_hypergeom_sf.argtypes = [u64, u64, u64, u64, P(dbl)]; L.or_hypergeom_sf.restype = ctypes.c_int
        L.or_hypergeometric_pvalue.argtypes = [u64, u64, u64, u64, P(dbl)]
        L.or_hypergeometric_pvalue.restype = ctypes.c_int
        L.or_fdr.argtypes = [u64, u64, u64, u64, dbl]; L.or_fdr.restype = dbl
        L.or_generate_thresholds.argtypes = [u32, P(u32), u32]; L.or_generate_thresholds.restype = u32
        L.or_list_new.argtypes = [P(ctypes.c_char_p), P(u32), u64]; L.or_list_new.restype = ctypes.c_void_p
        L.or_list_free.argtypes = [ctypes.c_void_p]
        L.or_list_set_thresholds.argtypes = [ctypes.c_void_p, P(u32), u32]
        L.or_list_set_thresholds.restype = ctypes.c_int
        L.or_list_len.argtypes = [ctypes.c_void_p]; L.or_list_len.restype = u64
        L.or_list_n_thresholds.argtypes = [ctypes.c_void_p]; L.or_list_n_thresholds.restype = u32
        L.or_list_thresholds.argtypes = [ctypes.c_void_p]; L.or_list_thresholds.restype = P(u32)
        L.or_list_ranks.argtypes = [ctypes.c_void_p]; L.or_list_ranks.restype = P(u32)
        L.or_list_id.argtypes = [ctypes.c_void_p, u64]; L.or_list_id.restype = ctypes.c_char_p
        L.or_population_size.argtypes = [ctypes.c_void_p, ctypes.c_void_p, P(ctypes.c_char_p), u64,
                                         ctypes.c_int, P(u64)]
        L.or_population_size.restype = ctypes.c_int
        L.or_process_threshold_pairs.argtypes = [ctypes.c_void_p, ctypes.c_void_p, P(u64), P(u64), u64,
                                                 P(Record)]
        L.or_process_threshold_pairs.restype = ctypes.c_int
        L.or_optimize.argtypes = [ctypes.c_void_p, ctypes.c_void_p, P(u64), P(u64), u64, P(Record)]
        L.or_optimize.restype = ctypes.c_int
        L.or_fast_grid.argtypes = [ctypes.c_void_p, ctypes.c_void_p, P(u64), P(u64), u64, P(i32), P(dbl)]
        L.or_fast_grid.restype = ctypes.c_int
        L.or_fast_optimize.argtypes = [ctypes.c_void_p, ctypes.c_void_p, P(u64), P(u64), u64, P(Record)]
        L.or_fast_optimize.restype = ctypes.c_int
        L.or_optimize_many.argtypes = [ctypes.c_void_p, ctypes.c_void_p, P(u64), P(u64), u64, u64,
                                       ctypes.c_int, ctypes.c_int, P(Record)]
        L.or_optimize_many.restype = ctypes.c_int
        L.or_reference_sample.argtypes = [ctypes.c_void_p, ctypes.c_void_p, P(u64), P(u64), u64, u64,
                                          u32, ctypes.c_int, P(dbl)]
        L.or_reference_sample.restype = u64
        L.or_exp_many.argtypes = [P(dbl), u64, P(dbl)]; L.or_exp_many.restype = None
        L.or_empirical_pvalue.argtypes = [dbl, P(dbl), u64]; L.or_empirical_pvalue.restype = dbl
        _lib = L
    return _lib


# ---- scalar statrs arithmetic ------------------------------------------------

def ln_factorial(x):
    return lib().or_ln_factorial(int(x))


def ln_gamma(x):
    return lib().or_ln_gamma(float(x))


def hypergeometric_pvalue(N, K, n, k):
    out = ctypes.c_double()
    rc = lib().or_hypergeometric_pvalue(int(N), int(K), int(n), int(k), ctypes.byref(out))
    if rc:
        raise ValueError("Failed to create hypergeometric distribution")
    return out.value


def fdr(l1, l2, overlap, N, sensitivity=0.8):
    return lib().or_fdr(int(l1), int(l2), int(overlap), int(N), float(sensitivity))


def generate_thresholds(max_rank):
    n = lib().or_generate_thresholds(int(max_rank), None, 0)
    out = (ctypes.c_uint32 * max(n, 1))()
    lib().or_generate_thresholds(int(max_rank), out, n)
    return [out[i] for i in range(n)]


def exp_many(x):
    a = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(a)
    lib().or_exp_many(a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(a),
                      out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


def empirical_pvalue(p0, perm_p):
    a = np.ascontiguousarray(perm_p, dtype=np.float64)
    return lib().or_empirical_pvalue(float(p0), a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(a))


# ---- ranked lists ------------------------------------------------------------

def _cstrs(ids):
    arr = (ctypes.c_char_p * max(len(ids), 1))()
    for i, s in enumerate(ids):
        arr[i] = s.encode() if isinstance(s, str) else s
    return arr


def _u64p(a):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


class OracleList:
    """RankedFeatureList::from(genes, ranks) restated (stable sort + thresholds)."""

    def __init__(self, ids, ranks):
        ids = list(ids)
        r = np.ascontiguousarray(ranks, dtype=np.uint32)
        self._keep = _cstrs(ids)
        self.h = lib().or_list_new(self._keep, r.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), len(ids))
        self.n = len(ids)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.or_list_free(self.h)
            self.h = None

    def set_thresholds(self, thr):
        t = np.ascontiguousarray(thr, dtype=np.uint32)
        if lib().or_list_set_thresholds(self.h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), len(t)):
            raise ValueError("thresholds must increase")

    @property
    def thresholds(self):
        n = lib().or_list_n_thresholds(self.h)
        p = lib().or_list_thresholds(self.h)
        return [p[i] for i in range(n)]

    @property
    def ranks(self):
        p = lib().or_list_ranks(self.h)
        return [p[i] for i in range(self.n)]

    @property
    def ids(self):
        return [lib().or_list_id(self.h, i).decode() for i in range(self.n)]


def population_size(l1, l2, background=None):
    N = ctypes.c_uint64()
    if background is None:
        rc = lib().or_population_size(l1.h, l2.h, None, 0, 0, ctypes.byref(N))
    else:
        keep = _cstrs(list(background))
        rc = lib().or_population_size(l1.h, l2.h, keep, len(background), 1, ctypes.byref(N))
    if rc == -1:
        raise ValueError("If no background is provided, the feature lists must have identical genes.")
    if rc == -2:
        raise ValueError("The following genes in the first ranked feature list are not in the background")
    if rc == -3:
        raise ValueError("The following genes in the second ranked feature list are not in the background")
    return N.value


def _idx(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.uint64)


def process_threshold_pairs(l1, l2, N, idx1=None, idx2=None):
    T1, T2 = len(l1.thresholds), len(l2.thresholds)
    out = (Record * max(T1 * T2, 1))()
    a, b = _idx(idx1), _idx(idx2)
    rc = lib().or_process_threshold_pairs(l1.h, l2.h, _u64p(a), _u64p(b), int(N), out)
    if rc:
        raise ValueError("Failed to create hypergeometric distribution")
    return [out[c] for c in range(T1 * T2)]


def optimize(l1, l2, N, idx1=None, idx2=None, fast=False):
    best = Record()
    a, b = _idx(idx1), _idx(idx2)
    fn = lib().or_fast_optimize if fast else lib().or_optimize
    rc = fn(l1.h, l2.h, _u64p(a), _u64p(b), int(N), ctypes.byref(best))
    if rc == -2:
        raise ValueError("empty threshold grid")
    if rc:
        raise ValueError("Failed to create hypergeometric distribution")
    return best


def fast_grid(l1, l2, N, idx1=None, idx2=None):
    T1, T2 = len(l1.thresholds), len(l2.thresholds)
    k = np.zeros(T1 * T2, dtype=np.int32)
    p = np.zeros(T1 * T2, dtype=np.float64)
    a, b = _idx(idx1), _idx(idx2)
    rc = lib().or_fast_grid(l1.h, l2.h, _u64p(a), _u64p(b), int(N),
                            k.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                            p.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    if rc:
        raise ValueError("Failed to create hypergeometric distribution")
    return k.reshape(T1, T2), p.reshape(T1, T2)


def optimize_many(l1, l2, N, idx1, idx2, fast=True, threads=None):
    """idx1: (P, N1) and idx2: (P, N2) explicit permutations."""
    a, b = _idx(idx1), _idx(idx2)
    P = a.shape[0]
    out = (Record * max(P, 1))()
    rc = lib().or_optimize_many(l1.h, l2.h, _u64p(a), _u64p(b), P, int(N), 1 if fast else 0,
                                int(threads or min(16, os.cpu_count() or 1)), out)
    if rc:
        raise ValueError("oracle optimize failed (%d)" % rc)
    return [out[p] for p in range(P)]


def reference_sample(l1, l2, N, idx1, idx2, row_stride, threads):
    """Time-able run of the reference cost structure; returns cells evaluated."""
    a, b = _idx(idx1), _idx(idx2)
    P = a.shape[0]
    chk = ctypes.c_double()
    return lib().or_reference_sample(l1.h, l2.h, _u64p(a), _u64p(b), P, int(N), int(row_stride),
                                     int(threads), ctypes.byref(chk))
