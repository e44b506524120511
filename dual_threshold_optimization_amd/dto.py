"""The DTO hot path (src/dto/): optimize, process_threshold_pairs,
compute_population_size and their result types, on the GPU through the C ABI.
"""
import ctypes
import os
import sys
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _lib
from .collections import FeatureList, PermutedRankedFeatureList, RankedFeatureList, philox_permutation


@dataclass
class OptimizationResultRecord:
    """src/dto/results_objects.rs:21-32.  feature_sets is None or the pair of
    feature lists (FeatureSets::Both) in debug mode."""
    rank1: int
    rank2: int
    set1_len: int
    set2_len: int
    population_size: int
    intersection_size: int
    pvalue: float
    permuted: bool
    feature_sets: Optional[tuple] = field(default=None, compare=False, repr=False)

    @classmethod
    def from_c(cls, r):
        return cls(int(r["rank1"]), int(r["rank2"]), int(r["set1_len"]), int(r["set2_len"]),
                   int(r["population_size"]), int(r["intersection_size"]), float(r["pvalue"]),
                   bool(r["permuted"]))


@dataclass
class OptimizationResult:
    """src/dto/results_objects.rs:14-18: Best(record) or Debug(records)."""
    best: Optional[OptimizationResultRecord] = None
    debug: Optional[List[OptimizationResultRecord]] = None

    @property
    def is_best(self):
        return self.best is not None


def _rec_view(raw, n):
    return np.frombuffer(raw, dtype=_lib.RECORD_DTYPE, count=n)


class Job:
    """One uploaded DTO job: two ranked lists (and background) resident on
    the device, with thresholds, bins, join and ln n! tables."""

    def __init__(self, handle):
        self._h = ctypes.c_void_p(handle)
        info = _lib.JobInfo()
        _lib.check(_lib.lib().dto_job_info_get(self._h, ctypes.byref(info)))
        self.population_size = info.population_size
        self.len1, self.len2 = info.len1, info.len2
        self.n_thresholds1, self.n_thresholds2 = info.n_thresholds1, info.n_thresholds2
        self.dropped1, self.dropped2 = info.dropped1, info.dropped2
        self.device_bytes = info.device_bytes

    # ---- construction -------------------------------------------------------
    @classmethod
    def from_lists(cls, ids1, ranks1, ids2, ranks2, background=None, drop_outside_background=False,
                   defer_upload=False):
        """RankedFeatureList::from for both lists + compute_population_size.
        defer_upload: keep the tables on the host until a server admits the
        job (or another run first needs them)."""
        L = _lib.lib()
        a, b = _lib.cstr_array(list(ids1)), _lib.cstr_array(list(ids2))
        r1 = np.ascontiguousarray(ranks1, dtype=np.uint32)
        r2 = np.ascontiguousarray(ranks2, dtype=np.uint32)
        bg = None if background is None else _lib.cstr_array(list(background))
        h = ctypes.c_void_p()
        flags = _lib.DTO_DROP_OUTSIDE_BACKGROUND if drop_outside_background else 0
        if defer_upload:
            flags |= _lib.DTO_DEFER_UPLOAD
        _lib.check(L.dto_job_create(a, _lib.u32p(r1), len(r1), b, _lib.u32p(r2), len(r2), bg,
                                    0 if background is None else len(background), flags, ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def from_c_arrays(cls, ids1, ranks1, ids2, ranks2, background=None, n_background=0,
                      drop_outside_background=False, defer_upload=False):
        """from_lists on id arrays already in C form (_lib.cstr_array) and
        uint32 rank arrays, for building many jobs from the same lists."""
        h = ctypes.c_void_p()
        flags = _lib.DTO_DROP_OUTSIDE_BACKGROUND if drop_outside_background else 0
        if defer_upload:
            flags |= _lib.DTO_DEFER_UPLOAD
        _lib.check(_lib.lib().dto_job_create(ids1, _lib.u32p(ranks1), len(ranks1), ids2, _lib.u32p(ranks2),
                                             len(ranks2), background, n_background if background is not None else 0,
                                             flags, ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def with_population(cls, ids1, ranks1, ids2, ranks2, population_size):
        """Lists as optimize() receives them: the caller's population size, no checks."""
        L = _lib.lib()
        a, b = _lib.cstr_array(list(ids1)), _lib.cstr_array(list(ids2))
        r1 = np.ascontiguousarray(ranks1, dtype=np.uint32)
        r2 = np.ascontiguousarray(ranks2, dtype=np.uint32)
        h = ctypes.c_void_p()
        _lib.check(L.dto_job_create_with_population(a, _lib.u32p(r1), len(r1), b, _lib.u32p(r2), len(r2),
                                                    int(population_size), ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def from_files(cls, path1, path2, background=None, drop_outside_background=False):
        h = ctypes.c_void_p()
        flags = _lib.DTO_DROP_OUTSIDE_BACKGROUND if drop_outside_background else 0
        _lib.check(_lib.lib().dto_job_create_from_files(
            os.fsencode(path1), os.fsencode(path2), None if background is None else os.fsencode(background),
            flags, ctypes.byref(h)))
        return cls(h.value)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().dto_job_destroy(self._h)
        self._h = None

    def release(self):
        """Hands the handle over (to a server that takes ownership)."""
        h, self._h = self._h, None
        return h

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def thresholds(self, which):
        n = self.n_thresholds1 if which == 1 else self.n_thresholds2
        out = np.zeros(max(n, 1), dtype=np.uint32)
        _lib.lib().dto_job_thresholds(self._h, which, _lib.u32p(out), n)
        return out[:n]

    def set_thresholds(self, thr1, thr2):
        t1 = np.ascontiguousarray(thr1, dtype=np.uint32)
        t2 = np.ascontiguousarray(thr2, dtype=np.uint32)
        _lib.check(_lib.lib().dto_job_set_thresholds(self._h, _lib.u32p(t1), len(t1), _lib.u32p(t2), len(t2)))
        self.n_thresholds1, self.n_thresholds2 = len(t1), len(t2)

    # ---- hot path -----------------------------------------------------------
    def optimize(self, idx1=None, idx2=None, mode="screened"):
        a = _lib.idx_array(idx1, self.len1)
        b = _lib.idx_array(idx2, self.len2)
        r = _lib.Record()
        _lib.check(_lib.lib().dto_optimize(self._h, _lib.u32p(a), _lib.u32p(b), _lib.MODES[mode], ctypes.byref(r)))
        return np.frombuffer(bytes(r), dtype=_lib.RECORD_DTYPE)[0]

    def process_threshold_pairs(self, idx1=None, idx2=None):
        a = _lib.idx_array(idx1, self.len1)
        b = _lib.idx_array(idx2, self.len2)
        n = self.n_thresholds1 * self.n_thresholds2
        out = np.zeros(n, dtype=_lib.RECORD_DTYPE)
        _lib.check(_lib.lib().dto_process_threshold_pairs(self._h, _lib.u32p(a), _lib.u32p(b),
                                                          out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def run(self, n_perm, seed=0, idx1=None, idx2=None, mode="screened"):
        """Unpermuted task + n_perm permutations + empirical p-value.
        Returns (dto_result, per-permutation records)."""
        a = _lib.idx_array(idx1, self.len1, n_perm)
        b = _lib.idx_array(idx2, self.len2, n_perm)
        res = _lib.Result()
        perm = np.zeros(max(n_perm, 1), dtype=_lib.RECORD_DTYPE)
        _lib.check(_lib.lib().dto_run(self._h, int(n_perm), int(seed) & (2**64 - 1), _lib.u32p(a), _lib.u32p(b),
                                      _lib.MODES[mode], ctypes.byref(res), perm.ctypes.data_as(ctypes.c_void_p)))
        return res, perm[:n_perm]

    def run_device(self, n_perm, d_out, perm_base=0, seed=0, d_idx1=None, d_idx2=None, mode="screened",
                   stream=None):
        """Asynchronous launch on device pointers (ints), e.g. torch tensors' data_ptr()."""
        _lib.check(_lib.lib().dto_run_device(self._h, int(n_perm), int(perm_base), int(seed) & (2**64 - 1),
                                             d_idx1, d_idx2, _lib.MODES[mode], d_out, stream))


def tie_warnings(rec):
    """The stderr lines optimize prints for a record whose minimum p is shared
    (src/dto/optimize_main.rs:86-108); rec is a dto_record (numpy or ctypes)."""
    ties = int(rec["ties"]) if not hasattr(rec, "ties") else int(rec.ties)
    n_p, n_pk = ties & 0xFFFF, ties >> 16
    out = []
    if n_p > 1:
        p = float(rec["pvalue"]) if not hasattr(rec, "pvalue") else float(rec.pvalue)
        k = int(rec["intersection_size"]) if not hasattr(rec, "intersection_size") else int(rec.intersection_size)
        out.append("Multiple results with the same minimum p-value (%.15f). Choosing the result with the largest "
                   "intersection size." % p)
        if n_pk > 1:
            out.append("Multiple results with the same maximum intersection size (%d). Choosing an arbitrary result "
                       "based on the order of thresholds." % k)
    return out


def _warn_ties(rec):
    for line in tie_warnings(rec):
        print(line, file=sys.stderr)


def _job_for(l1, l2, population_size):
    return Job.with_population(l1.ids(), l1.ranks(), l2.ids(), l2.ranks(), population_size)


def _perm_indices(lst, permute, seed, which):
    if isinstance(lst, PermutedRankedFeatureList):
        return lst.indices
    if not permute:
        return None
    s = int.from_bytes(os.urandom(8), "little") if seed is None else seed
    return philox_permutation(s, 0, 0, which, len(lst))


def _feature_sets(lst, idx, threshold):
    genes = lst.genes()
    ranks = lst.ranks()
    if idx is None:
        return [genes[k] for k in range(len(genes)) if ranks[k] <= threshold]
    return [genes[int(idx[k])] for k in range(len(genes)) if ranks[k] <= threshold]


def process_threshold_pairs(ranked_feature_list1, ranked_feature_list2, use_permutation, population_size,
                            debug=False, *, idx1=None, idx2=None, seed=None):
    """src/dto/process_threshold_pairs.rs:71-131: one record per threshold
    pair (row-major), computed on the GPU."""
    l1 = getattr(ranked_feature_list1, "original", ranked_feature_list1)
    l2 = getattr(ranked_feature_list2, "original", ranked_feature_list2)
    if use_permutation:
        a = idx1 if idx1 is not None else _perm_indices(ranked_feature_list1, True, seed, 1)
        b = idx2 if idx2 is not None else _perm_indices(ranked_feature_list2, True, seed, 2)
    else:
        a = b = None
    job = _job_for(l1, l2, population_size)
    try:
        raw = job.process_threshold_pairs(a, b)
    finally:
        job.close()
    out = []
    for r in raw:
        rec = OptimizationResultRecord.from_c(r)
        rec.permuted = bool(use_permutation)
        if debug:
            rec.feature_sets = (_feature_sets(l1, a, rec.rank1), _feature_sets(l2, b, rec.rank2))
        out.append(rec)
    return out


def optimize(ranked_feature_list1, ranked_feature_list2, permute, population_size, debug=False, *,
             idx1=None, idx2=None, seed=None, mode="screened"):
    """src/dto/optimize_main.rs:53-118: the threshold pair with the smallest
    hypergeometric p-value (ties: larger intersection, then smaller
    (rank1, rank2)); debug=True returns every record."""
    if debug:
        return OptimizationResult(debug=process_threshold_pairs(
            ranked_feature_list1, ranked_feature_list2, permute, population_size, True,
            idx1=idx1, idx2=idx2, seed=seed))
    l1 = getattr(ranked_feature_list1, "original", ranked_feature_list1)
    l2 = getattr(ranked_feature_list2, "original", ranked_feature_list2)
    if permute:
        a = idx1 if idx1 is not None else _perm_indices(ranked_feature_list1, True, seed, 1)
        b = idx2 if idx2 is not None else _perm_indices(ranked_feature_list2, True, seed, 2)
    else:
        a = b = None
    job = _job_for(l1, l2, population_size)
    try:
        r = job.optimize(a, b, mode)
    finally:
        job.close()
    _warn_ties(r)
    rec = OptimizationResultRecord.from_c(r)
    rec.permuted = bool(permute)
    return OptimizationResult(best=rec)


def compute_population_size(ranked_feature_list1, ranked_feature_list2, background=None):
    """src/dto/compute_population_size.rs:66-104 (raises where it panics)."""
    ids1, ids2 = ranked_feature_list1.ids(), ranked_feature_list2.ids()
    if background is None:
        s2 = set(ids2)
        inter = sum(1 for g in ids1 if g in s2)
        if inter != len(ids1) or inter != len(ids2):
            raise ValueError("If no background is provided, the feature lists must have identical genes.")
        return inter
    bg = background.ids() if isinstance(background, FeatureList) else [str(x) for x in background]
    bs = set(bg)
    for ids, name in ((ids1, "first"), (ids2, "second")):
        missing = [g for g in ids if g not in bs]
        if missing:
            raise ValueError("The following genes in the %s ranked feature list are not in the background: %r"
                             % (name, missing))
    return len(bg)
