"""Shared helpers for the parity tests: build matching product jobs and oracle lists."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def read_csv(path):
    ids, ranks = [], []
    with open(path) as f:
        for line in f:
            a, b = line.rstrip("\n").split(",")
            ids.append(a.strip())
            ranks.append(int(b))
    return ids, np.array(ranks, dtype=np.uint32)


def readme_lists():
    i1, r1 = read_csv(os.path.join(GOLDEN, "ranklist1.csv"))
    i2, r2 = read_csv(os.path.join(GOLDEN, "ranklist2.csv"))
    return i1, r1, i2, r2


def ten_gene_lists():
    """src/dto/optimize_main.rs:126-175."""
    g1 = ["gene%d" % i for i in range(1, 11)]
    g2 = ["gene7", "gene3", "gene9", "gene1", "gene5", "gene10", "gene4", "gene8", "gene6", "gene2"]
    r = np.arange(1, 11, dtype=np.uint32)
    return g1, r, g2, r.copy()


def rec_tuple(r):
    """(rank1, rank2, set1_len, set2_len, N, k, p, ties) of a product record
    (numpy or ctypes) or an oracle record; ties packs the counts behind
    optimize's two tie warnings (0 in grid records)."""
    if hasattr(r, "as_tuple"):
        return r.as_tuple() + (int(r.ties),)
    if hasattr(r, "_fields_"):
        return (int(r.rank1), int(r.rank2), int(r.set1_len), int(r.set2_len), int(r.population_size),
                int(r.intersection_size), float(r.pvalue), int(r.ties))
    return (int(r["rank1"]), int(r["rank2"]), int(r["set1_len"]), int(r["set2_len"]), int(r["population_size"]),
            int(r["intersection_size"]), float(r["pvalue"]), int(r["ties"]))


def random_lists(rng, n, overlap_frac=1.0, tie_frac=0.1, dup=False, extra=0, max_rank=None):
    """Small random lists over n shared ids (+extra list-specific ids)."""
    shared = ["f%d" % i for i in range(n)]
    e1 = ["a%d" % i for i in range(extra)]
    e2 = ["b%d" % i for i in range(extra)]
    ids1 = list(np.array(shared + e1, dtype=object)[rng.permutation(n + extra)])
    ids2 = list(np.array(shared + e2, dtype=object)[rng.permutation(n + extra)])
    if dup and len(ids2) > 2:
        ids2[1] = ids2[0]
    if dup and len(ids1) > 3:
        ids1[3] = ids1[2]
    L = n + extra
    mr = max_rank or L
    r1 = np.sort(rng.integers(1, mr + 1, size=L)).astype(np.uint32) if tie_frac else np.arange(1, L + 1, dtype=np.uint32)
    r2 = np.sort(rng.integers(1, mr + 1, size=L)).astype(np.uint32) if tie_frac else np.arange(1, L + 1, dtype=np.uint32)
    rng.shuffle(r1)
    rng.shuffle(r2)
    return [str(x) for x in ids1], r1, [str(x) for x in ids2], r2
