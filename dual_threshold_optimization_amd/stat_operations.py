"""src/stat_operations/: hypergeometric p-value (on the GPU), intersect_genes,
empirical p-value and FDR."""
import ctypes
import sys

import numpy as np

from . import _lib
from .collections import Feature


def hypergeometric_pvalue(population_size, successes_in_population, sample_size, observed_overlap):
    """src/stat_operations/hypergeometric_pvalue.rs:33-50 (statrs sf replay on the GPU)."""
    return float(hypergeometric_pvalues([population_size], [successes_in_population], [sample_size],
                                        [observed_overlap])[0])


def hypergeometric_pvalues(N, K, n, k):
    """Vectorised hypergeometric_pvalue; raises where Hypergeometric::new fails."""
    a = [np.ascontiguousarray(x, dtype=np.uint64) for x in (N, K, n, k)]
    if len({len(x) for x in a}) != 1:
        raise ValueError("arrays must have equal length")
    out = np.zeros(max(len(a[0]), 1), dtype=np.float64)
    _lib.check(_lib.lib().dto_hypergeometric_pvalue(_lib.u64p(a[0]), _lib.u64p(a[1]), _lib.u64p(a[2]),
                                                    _lib.u64p(a[3]), len(a[0]), _lib.dblp(out)))
    return out[:len(a[0])]


def intersect_genes(genes1, genes2):
    """src/stat_operations/intersect_genes.rs:38-56: entries of genes2 whose id
    is among genes1's ids (duplicates in genes2 count each time)."""
    ids1 = {g.id() if isinstance(g, Feature) else str(g) for g in genes1}
    return sum(1 for g in genes2 if (g.id() if isinstance(g, Feature) else str(g)) in ids1)


def fdr(ranked_list_1_len, ranked_list_2_len, overlap_len, population_size, sensitivity=0.8):
    """src/stat_operations/fdr.rs:29-60."""
    out = ctypes.c_double()
    rc = _lib.lib().dto_fdr_checked(int(ranked_list_1_len), int(ranked_list_2_len), int(overlap_len),
                                    int(population_size), float(sensitivity), ctypes.byref(out))
    if rc == _lib.DTO_ERR_INVALID_ARGUMENT:
        raise ValueError(_lib.lib().dto_last_error().decode())
    _lib.check(rc)
    return out.value


def _to_c_record(r):
    c = _lib.Record()
    c.rank1, c.rank2 = r.rank1, r.rank2
    c.set1_len, c.set2_len = r.set1_len, r.set2_len
    c.population_size, c.intersection_size = r.population_size, r.intersection_size
    c.pvalue, c.permuted = r.pvalue, 1 if r.permuted else 0
    return c


def empirical_pvalue(results):
    """src/stat_operations/empirical_pvalue.rs:109-187 over OptimizationResult
    values; returns the JSON object as a dict."""
    bests = [r.best for r in results if getattr(r, "best", None) is not None]
    unperm = [b for b in bests if not b.permuted]
    perm = [b for b in bests if b.permuted]
    if len(unperm) != 1:
        print("Warning: Expected exactly one unpermuted result, but found %d." % len(unperm), file=sys.stderr)
    if not unperm:
        raise ValueError("No unpermuted result found in the provided results.")
    u = _to_c_record(unperm[0])
    arr = (_lib.Record * max(len(perm), 1))(*[_to_c_record(p) for p in perm])
    res = _lib.Result()
    _lib.check(_lib.lib().dto_empirical_pvalue(ctypes.byref(u), ctypes.cast(arr, ctypes.c_void_p), len(perm),
                                               ctypes.byref(res)))
    return result_dict(res)


def result_dict(res):
    u = res.unpermuted
    return {
        "rank1": u.rank1, "rank2": u.rank2, "set1_len": u.set1_len, "set2_len": u.set2_len,
        "population_size": u.population_size, "unpermuted_intersection_size": u.intersection_size,
        "unpermuted_pvalue": u.pvalue, "empirical_pvalue": res.empirical_pvalue, "fdr": res.fdr,
    }


def result_json(res):
    """serde_json::to_string_pretty of the result object (src/main.rs:165)."""
    buf = ctypes.create_string_buffer(4096)
    n = _lib.lib().dto_result_json(ctypes.byref(res), buf, len(buf))
    if n < 0:
        raise RuntimeError("json formatting failed")
    return buf.value.decode()
