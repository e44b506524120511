"""Writes tests/golden/dense_2k_genome_20k.json: the oracle's records for the
dense-threshold genome-scale job (thresholds 10, 20, ..., 20,000 on both
lists, 4M cells per grid), unpermuted and for 32 explicit permutations.

The oracle needs about a minute per grid at this size, so the GPU test reads
these records instead of recomputing them.  Run from the repository root:

    python tests/golden/make_dense_fixture.py
"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from dual_threshold_optimization_amd import workloads as W  # noqa: E402

PERM_SEED = 21
P = 32


def main():
    ids1, r1, ids2, r2, bg = W.genome_20k()
    keep = set(bg)
    f1 = [(a, b) for a, b in zip(ids1, r1) if a in keep]
    f2 = [(a, b) for a, b in zip(ids2, r2) if a in keep]
    l1 = O.OracleList([a for a, _ in f1], [b for _, b in f1])
    l2 = O.OracleList([a for a, _ in f2], [b for _, b in f2])
    t1, t2 = W.dense_2k_thresholds()
    l1.set_thresholds(t1)
    l2.set_thresholds(t2)
    i1, i2 = W.explicit_permutations(P, len(f1), len(f2), PERM_SEED)
    recs = [O.optimize(l1, l2, len(bg), fast=True)]
    recs += O.optimize_many(l1, l2, len(bg), i1, i2, fast=True, threads=min(16, os.cpu_count() or 1))

    def enc(r):
        t = r.as_tuple()
        return {"rank1": t[0], "rank2": t[1], "set1_len": t[2], "set2_len": t[3], "population_size": t[4],
                "intersection_size": t[5], "pvalue": float(t[6]).hex(), "ties": int(r.ties)}
    out = {"workload": "genome_20k (seed 2), background drop, thresholds 10..20000 step 10 on both lists",
           "permutations": "workloads.explicit_permutations(%d, 20000, 20000, %d)" % (P, PERM_SEED),
           "unpermuted": enc(recs[0]), "permuted": [enc(r) for r in recs[1:]]}
    path = os.path.join(ROOT, "tests", "golden", "dense_2k_genome_20k.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print("wrote", path)


if __name__ == "__main__":
    main()
