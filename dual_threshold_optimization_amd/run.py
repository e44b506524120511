"""src/run/: the permutation loop.  run_single_node keeps the reference's
signature (src/run/single_node.rs:83-137) but runs every task on the GPU;
Batch and Server replace run_multi_node's MPI fan-out with on-device batching.
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from .dto import Job, OptimizationResult, OptimizationResultRecord, _warn_ties
from .stat_operations import result_dict


@dataclass
class Task:
    """src/run/task.rs:4-9."""
    id: int
    permute: bool


def run_single_node(tasks, ranked_feature_list1, ranked_feature_list2, population_size, num_threads=1, *,
                    seed=0, mode="screened"):
    """One OptimizationResult per task (order: unpermuted first, then the
    permuted tasks in id order; the reference's order is thread-completion
    order, which empirical_pvalue ignores).  num_threads is accepted for
    signature compatibility."""
    l1, l2 = ranked_feature_list1, ranked_feature_list2
    job = Job.with_population(l1.ids(), l1.ranks(), l2.ids(), l2.ranks(), population_size)
    try:
        n_perm = sum(1 for t in tasks if t.permute)
        res, perm = job.run(n_perm, seed=seed, mode=mode)
    finally:
        job.close()
    # each task's optimize warns about its own ties (optimize_main.rs:86-108)
    n_unperm = sum(1 for t in tasks if not t.permute)
    for _ in range(n_unperm):
        _warn_ties(res.unpermuted)
    for r in perm:
        _warn_ties(r)
    u = res.unpermuted
    out = []
    for _ in range(n_unperm):
        out.append(OptimizationResult(best=OptimizationResultRecord(
            u.rank1, u.rank2, u.set1_len, u.set2_len, u.population_size, u.intersection_size, u.pvalue, False)))
    for r in perm:
        rec = OptimizationResultRecord.from_c(r)
        rec.permuted = True
        out.append(OptimizationResult(best=rec))
    return out


class Batch:
    """Many jobs with their own permutation counts in one persistent,
    hipGraph-replayed launch."""

    def __init__(self, jobs, n_perm, seed=0):
        self.jobs = list(jobs)
        arr = (ctypes.c_void_p * len(self.jobs))(*[j.handle.value for j in self.jobs])
        np_ = np.ascontiguousarray(n_perm, dtype=np.uint64)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().dto_batch_create(arr, _lib.u64p(np_), len(self.jobs), int(seed), ctypes.byref(h)))
        self._h = h

    def autotune(self):
        w = ctypes.c_int()
        _lib.check(_lib.lib().dto_batch_autotune(self._h, ctypes.byref(w)))
        return w.value

    def launch(self, stream=None):
        _lib.check(_lib.lib().dto_batch_launch(self._h, stream))

    def results(self):
        out = (_lib.Result * max(len(self.jobs), 1))()
        _lib.check(_lib.lib().dto_batch_results(self._h, out))
        return [out[i] for i in range(len(self.jobs))]

    def close(self):
        if getattr(self, "_h", None) is not None:
            _lib.lib().dto_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request loop (see dto_server_* in include/dto.h)."""

    def __init__(self, hbm_budget_bytes=8 << 30, max_active_jobs=256, perms_per_slice=1024, seed=0):
        cfg = _lib.ServerConfig(hbm_budget_bytes, max_active_jobs, perms_per_slice, seed)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().dto_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h

    def submit(self, job, n_perm):
        rid = ctypes.c_uint64()
        _lib.check(_lib.lib().dto_server_submit(self._h, job.handle, int(n_perm), ctypes.byref(rid)))
        job.release()
        return rid.value

    def poll(self, timeout_ms=10, cap=1024):
        ids = (ctypes.c_uint64 * cap)()
        res = (_lib.Result * cap)()
        lat = (ctypes.c_double * cap)()
        n = ctypes.c_uint32()
        _lib.check(_lib.lib().dto_server_poll(self._h, int(timeout_ms), ids, res, lat, cap, ctypes.byref(n)))
        return [(ids[i], res[i], lat[i]) for i in range(n.value)]

    def pending(self):
        q, a = ctypes.c_uint64(), ctypes.c_uint64()
        _lib.check(_lib.lib().dto_server_pending(self._h, ctypes.byref(q), ctypes.byref(a)))
        return q.value, a.value

    def close(self):
        if getattr(self, "_h", None) is not None:
            _lib.lib().dto_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["Task", "run_single_node", "Batch", "Server", "result_dict"]
